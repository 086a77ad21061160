"""Benchmark: full plain tree build on one MI355X, plus the sort-dominated and
batched workloads of the same metric.

Headline workload: 2^26 points drawn from a 64-centre Gaussian mixture
(centres uniform in [0,1024)^3, sigma 8, clipped to the box), float32 xyzw,
30-bit Morton codes.  A step is one complete build -- Morton codes with the
sort histogram fused in, 4 OneSweep binning passes, unique, radix tree with
fused edge counts, exclusive scan, octree -- on points already resident in
HBM, queued without host synchronisation.  `value` = points / step time.

Also reported (same JSON line):
  roofline      dominant kernel (the binning pass): algorithmic bytes / average
                launch time, launch times from HIP events around each launch on
                its own stream (outside the timed loop)
  indexed_build the headline build with the point -> leaf map (point index
                sorted along with the codes, run starts from unique)
  pcie_inclusive  the headline build for points and results in host memory:
                points copied in from pinned memory, unique codes and octree
                nodes copied back (the cost of the host boundary; not `value`)
  sort          standalone u32 key sort of the same codes (histogram + 4
                passes) against the 36 B/key algorithmic traffic
  sort_pairs_u64  2^27 63-bit keys with u32 payloads, 8 passes (200 B/pair)
  batched       1024 clouds, sizes log-uniform in [2^10, 2^20], alternating
                uniform and 16-cluster Gaussian (sigma 4): one segmented launch
                per stage replayed as a HIP graph vs one plain build per cloud
  served        Poisson arrivals (1k-50k req/s) of requests sized log-uniform in
                [2^12, 2^22]: the server vs a one-at-a-time plain loop on the
                same traces, p50/p99 latency and points/s
  cpu_baseline  the C port of the pipeline (the oracle's stages) on the same
                input on this host: multi-threaded (the value), single-threaded,
                and std::sort of the codes
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "glsl-tree-construction_amd"))

HBM_PEAK_GBS = 8000.0
TUNING_FILE = os.path.join(ROOT, "glsl-tree-construction_amd", "tc", "tuning_mi355x.json")
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "binning_pass_2p26_gauss_traffic.md")


def measured_traffic():
    """HBM bytes per binning-pass launch at the headline workload, from the
    committed rocprofv3 FETCH_SIZE + WRITE_SIZE summary (None if absent)."""
    try:
        for line in open(TRAFFIC_FILE):
            if line.startswith("traffic_bytes_per_launch:"):
                return int(line.split(":")[1])
    except OSError:
        pass
    return None


def traffic_probe(ctx, pts_h):
    """Three headline builds, then reference kernels of known traffic:
    tc_traffic_reference on a 1 GiB buffer (16-byte plain loads, the same
    non-temporal -- the form the Morton, unique and binning kernels use for data
    read once --, and 16-byte stores), and torch's fill_ / sum on 256 MiB."""
    import ctypes as C

    import torch

    import tc

    n = len(pts_h)
    pts = torch.from_numpy(pts_h).cuda()
    tree = tc.Tree(ctx, n, 30)
    for _ in range(3):
        tree.build(pts, n, 0.0, 1024.0)
    L = tc.lib()
    big = torch.ones(1 << 28, dtype=torch.int32, device="cuda")  # 1 GiB
    sink = torch.zeros(2048, dtype=torch.int32, device="cuda")
    for mode in (0, 1, 2):
        for _ in range(3):
            st = L.tc_traffic_reference(ctx.handle, C.c_void_p(big.data_ptr()), big.numel() * 4, mode,
                                        C.c_void_p(sink.data_ptr()), None)
            assert st == 0, st
    a = torch.empty(1 << 26, dtype=torch.int32, device="cuda")
    for _ in range(3):
        a.fill_(1)
    total = 0
    for _ in range(3):
        total += int(a.sum())
    torch.cuda.synchronize()
    print(json.dumps({"traffic_probe": tree.read_counts(), "check": total}), flush=True)
    tree.free()


def achievable_bandwidth(ctx, reps=20):
    """What a plain streaming kernel reaches on this GPU, the practical ceiling
    under the 8 TB/s peak: tc_traffic_reference's copy (non-temporal 16-byte
    loads, 16-byte stores, one contiguous 32 KiB chunk per workgroup) of 256 MiB
    (the keys of one binning pass at 2^26 u32 keys, in and out) and of 512 MiB,
    and its non-temporal read of 1 GiB (the headline's points); each averaged
    over `reps` back-to-back launches timed by the host."""
    import ctypes as C

    import torch

    import tc

    L = tc.lib()
    buf = torch.ones(1 << 28, dtype=torch.int32, device="cuda")  # 1 GiB
    sink = torch.zeros(2048, dtype=torch.int32, device="cuda")
    out = {}
    for name, mode, nbytes, moved in (("copy_256MiB", 3, 1 << 29, 1 << 29), ("copy_512MiB", 3, 1 << 30, 1 << 30),
                                      ("read_1GiB", 1, 1 << 30, 1 << 30)):
        def launch():
            st = L.tc_traffic_reference(ctx.handle, C.c_void_p(buf.data_ptr()), nbytes, mode,
                                        C.c_void_p(sink.data_ptr()), None)
            assert st == 0, st

        for _ in range(3):
            launch()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            launch()
        torch.cuda.synchronize()
        out[name + "_GBps"] = round(moved / ((time.perf_counter() - t0) / reps) / 1e9, 1)
    del buf
    torch.cuda.empty_cache()
    return out


METRIC = "Mpoints/s full tree build @2^26 (% HBM roofline); batched vs plain p99 latency"


def algorithmic_bytes(n, u, o):
    """Algorithmic HBM bytes per launch of each kernel (reads + writes that the
    algorithm needs, counting each array once)."""
    nb = max(u - 1, 0)
    return {
        "morton": 20 * n,            # 16 B point in, 4 B code out (histogram fused, no extra traffic)
        "histogram": 4 * n,          # one read of the codes (standalone sort)
        "binning": 8 * n,            # one read + one write of the keys, per pass
        "unique": 4 * n + 4 * u,     # sorted codes in, unique codes out
        "radix_tree": 4 * u + 15 * nb,  # codes in; prefix_n, 2 leaf flags, left_child, parent, edge count out
        "scan": 8 * nb,              # edge counts in, offsets out
        "octree": 21 * nb + 64 * o,  # counts, offsets, codes, prefix_n, links in; nodes out
    }


def timed(fn, steps, warmup):
    import torch

    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        fn()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / steps


def profile_stages(ctx, fn, reps):
    import torch

    ctx.profile_reset()
    ctx.profile(True)
    for _ in range(reps):
        fn()
    torch.cuda.synchronize()
    rep = ctx.profile_report()
    ctx.profile(False)
    return {k: (ms / cnt, cnt // reps) for k, (ms, cnt) in rep.items()}


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def host_threads():
    """Threads for the multi-threaded baseline: the cores this process may run
    on, capped at the CPU share of a one-GPU box (16)."""
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    env = int(os.environ.get("OMP_NUM_THREADS") or avail)
    return max(1, min(avail, env, 16))


def run_cpu_baseline(pts, reps=5):
    """The CPU port of the pipeline (the oracle, gcc/g++ -O3 -march=x86-64-v3) on this host, timed
    per stage: one warm-up, then the median of `reps` runs per stage; multi-
    threaded on the host's cores (the value reported), single-threaded, and
    std::sort of the codes (the reference's only CPU stage)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    import numpy as np

    threads = host_threads()
    n = len(pts)
    runs = []
    for r in range(reps + 1):
        st = {}
        oracle.build_tree_mt(pts, 0.0, 1024.0, 30, threads=threads, timings=st)
        if r:
            runs.append(st)  # the stages only, as the GPU build (output arrays preallocated)
    stages = {k: float(np.median([x[k] for x in runs])) for k in runs[0]}
    t_mt = sum(stages.values()) * 1e-3
    st1 = []
    for r in range(reps + 1):
        t0 = time.perf_counter()
        oracle.build_tree(pts, 0.0, 1024.0, 30)
        if r:
            st1.append(time.perf_counter() - t0)
    t_st = float(np.median(st1))
    codes = oracle.compute_morton(pts, 0.0, 1024.0)
    sort_ms = float(np.median([oracle.std_sort_ms(codes) for _ in range(3)]))
    configs = cpu_baseline_configs(oracle, threads)
    return {
        "value": round(n / t_mt / 1e6, 3),
        "unit": "Mpoints/s",
        "cores": threads,
        "kind": "port",
        "cpu": cpu_model(),
        "sample": f"full workload ({n} points): the C port of the pipeline (the oracle's stages, "
                  f"gcc/g++ -O3 -march=x86-64-v3) on "
                  f"{threads} host threads; per stage one warm-up then the median of {reps} runs, summed: "
                  f"{t_mt:.3f} s per build (output arrays allocated beforehand, as for the GPU build)",
        "stage_ms": {k: round(v, 2) for k, v in stages.items()},
        "single_thread_Mpoints_per_s": round(n / t_st / 1e6, 3),
        "single_thread_s": round(t_st, 3),
        "single_thread_method": f"whole oracle build, median of {reps} after one warm-up",
        "std_sort_ms": round(sort_ms, 1),
        "configs": configs,
    }


def cpu_baseline_configs(oracle, threads):
    """The CPU port on the other workloads' inputs (the same generators and
    seeds as the GPU legs): the 2^20 uniform build (per stage median of 5,
    and single-threaded), a 2^24-pair sample of the 2^27 u64 key-value sort
    (the parallel LSD sort, median of 3, and std::stable_sort once), and the
    1,024-cloud batch, whole clouds built single-threaded on the host threads
    in parallel (wall time of the second of two runs)."""
    import numpy as np

    import tc

    out = {}
    pts1 = tc.gen_uniform(1 << 20, 0xC0FFEE01)
    runs = []
    for r in range(6):
        st = {}
        oracle.build_tree_mt(pts1, 0.0, 1024.0, 30, threads=threads, timings=st)
        if r:
            runs.append(sum(st.values()))
    st1 = []
    for r in range(4):
        t0 = time.perf_counter()
        oracle.build_tree(pts1, 0.0, 1024.0, 30)
        if r:
            st1.append(time.perf_counter() - t0)
    out["1"] = {"workload": "2^20 points uniform in [0,1024)^3, 30-bit codes, full build",
                "Mpoints_per_s": round(len(pts1) / (float(np.median(runs)) * 1e-3) / 1e6, 1),
                "single_thread_Mpoints_per_s": round(len(pts1) / float(np.median(st1)) / 1e6, 1)}
    n3 = 1 << 24
    keys = oracle.compute_morton63(tc.gen_uniform(n3, 0xC0FFEE03, 2097152.0), 0.0, 2097152.0)
    vals = np.arange(n3, dtype=np.uint32)
    mt = float(np.median([oracle.sort_pairs_u64_mt(keys, vals, threads)[2] for _ in range(3)]))
    std = oracle.std_stable_sort_pairs_u64(keys, vals)[2]
    out["3"] = {"workload": "sample of the 2^27 u64-key + u32-payload sort: 2^24 63-bit Morton keys (xyz uniform in "
                            "[0,2^21)^3) with their indices",
                "Mpairs_per_s": round(n3 / (mt * 1e-3) / 1e6, 1),
                "std_stable_sort_Mpairs_per_s": round(n3 / (std * 1e-3) / 1e6, 1)}
    rng = np.random.default_rng(0xC0FFEE04)
    sizes = np.round(np.exp(rng.uniform(np.log(2 ** 10), np.log(2 ** 20), 1024))).astype(np.uint32)
    clouds = [tc.gen_uniform(int(s), 0xC0FFEE04 + i) if i % 2 == 0 else
              tc.gen_gaussian_mixture(int(s), 0xC0FFEE04 + i, 16, 4.0) for i, s in enumerate(sizes)]
    from concurrent.futures import ThreadPoolExecutor

    wall = 0.0
    with ThreadPoolExecutor(max_workers=threads) as ex:
        for r in range(2):
            t0 = time.perf_counter()
            list(ex.map(lambda c: oracle.build_tree_mt(c, 0.0, 1024.0, 30, threads=1), clouds))
            wall = time.perf_counter() - t0
    out["4"] = {"workload": "the 1,024-cloud batch (sizes log-uniform in [2^10, 2^20], alternating uniform and "
                            "16-cluster sigma-4 clouds): whole clouds built single-threaded on the host threads "
                            "in parallel, wall time of the second of two runs",
                "Mpoints_per_s": round(float(sizes.sum()) / wall / 1e6, 1)}
    return out


def bench_pcie(ctx, pts_h, steps=5):
    """The headline build for a caller whose points and results live in host
    memory: points copied in from pinned host memory, the build, the counts
    read, then the unique codes and octree nodes copied back into pinned host
    memory (the host-array boundary's cost; never the headline value)."""
    import ctypes as C

    import torch

    import tc
    from tc.api import _hip

    n = len(pts_h)
    host_pts = torch.from_numpy(pts_h).pin_memory()
    dev_pts = torch.empty(host_pts.shape, dtype=host_pts.dtype, device="cuda")
    tree = tc.Tree(ctx, n, 30)
    dev_pts.copy_(host_pts)
    tree.build(dev_pts, n, 0.0, 1024.0)
    cnt = tree.read_counts()
    host_u = torch.empty(max(n, 1), dtype=torch.int32).pin_memory()
    host_nodes = torch.empty(max(cnt["n_nodes"], 1) * 16, dtype=torch.int32).pin_memory()
    hip = _hip()

    def step():
        dev_pts.copy_(host_pts, non_blocking=True)
        tree.build(dev_pts, n, 0.0, 1024.0)
        c = tree.read_counts()
        for dst, src, nbytes in ((host_u, tree.t.unique, 4 * c["n_unique"]), (host_nodes, tree.t.nodes, 64 * c["n_nodes"])):
            if nbytes and hip.hipMemcpy(C.c_void_p(dst.data_ptr()), C.c_void_p(src), C.c_size_t(nbytes), 2) != 0:
                raise RuntimeError("hipMemcpy")
        return c

    step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        c = step()
    torch.cuda.synchronize()
    s = (time.perf_counter() - t0) / steps
    tree.free()
    moved = 16 * n + 4 * c["n_unique"] + 64 * c["n_nodes"]
    return {
        "workload": "the headline build with points in pinned host memory: copy in, build, unique codes and octree "
                    "nodes copied back to pinned host memory",
        "ms": round(s * 1e3, 3),
        "Mpoints_per_s": round(n / s / 1e6, 1),
        "pcie_bytes": moved,
        "steps": steps,
    }


def bench_sort_u32(ctx, codes, steps):
    import torch

    n = codes.shape[0]
    keys = codes.clone()
    alt = torch.empty_like(codes)

    def step():
        keys.copy_(codes)
        ctx.sort(keys, alt)

    for _ in range(3):  # warm-up: first-call costs stay out of the averages, as for the build
        step()
    prof = profile_stages(ctx, step, steps)
    ms = prof["histogram"][0] + prof["binning"][0] * prof["binning"][1]
    return {
        "workload": f"2^{n.bit_length() - 1} u32 Morton codes of the headline cloud, keys only, 4 x 8-bit passes",
        "ms": round(ms, 4),
        "histogram_ms": round(prof["histogram"][0], 4),
        "pass_ms": round(prof["binning"][0], 4),
        "alg_bytes": 36 * n,
        "GBps": round(36 * n / (ms * 1e-3) / 1e9, 1),
        "frac": round(36 * n / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
    }


def bench_config1(ctx, steps=200):
    """The CPU-runnable workload on the GPU: 2^20 points uniform in
    [0,1024)^3, 30-bit codes, the whole build (HBM-resident input)."""
    import torch

    import tc

    n = 1 << 20
    pts = torch.from_numpy(tc.gen_uniform(n, 0xC0FFEE01)).cuda()
    tree = tc.Tree(ctx, n, 30)
    s = timed(lambda: tree.build(pts, n, 0.0, 1024.0), steps, 20)
    c = tree.read_counts()
    tree.free()
    return {"workload": "2^20 points uniform in [0,1024)^3, 30-bit codes, full build", "ms": round(s * 1e3, 4),
            "Mpoints_per_s": round(n / s / 1e6, 1), "unique_codes": c["n_unique"], "octree_nodes": c["n_nodes"]}


def bench_sort_pairs_u64(ctx, log2n, steps):
    import torch

    import tc

    n = 1 << log2n
    pts = torch.from_numpy(tc.gen_uniform(n, 0xC0FFEE03, 2097152.0)).cuda()
    keys0 = torch.empty(n, dtype=torch.int64, device="cuda")
    ctx.morton63(pts, keys0, 0.0, 2097152.0)
    del pts
    vals0 = torch.arange(n, dtype=torch.int32, device="cuda")
    keys, vals = keys0.clone(), vals0.clone()
    ka, va = torch.empty_like(keys), torch.empty_like(vals)

    def step():
        keys.copy_(keys0)
        vals.copy_(vals0)
        ctx.sort(keys, ka, vals, va)

    step()  # warm-up
    prof = profile_stages(ctx, step, steps)
    ms = prof["histogram"][0] + prof["binning"][0] * prof["binning"][1]
    ok = bool(torch.all(keys[1:] >= keys[:-1]).item())
    alg = 200 * n
    return {
        "workload": f"2^{log2n} 63-bit Morton keys (xyz uniform in [0,2^21)^3) + u32 index payload, 8 x 8-bit passes",
        "ms": round(ms, 4),
        "pass_ms": round(prof["binning"][0], 4),
        "alg_bytes": alg,
        "GBps": round(alg / (ms * 1e-3) / 1e9, 1),
        "frac": round(alg / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
        "sorted": ok,
    }


def bench_batched(ctx, n_clouds, steps, seed=0xC0FFEE04, plain=True):
    import numpy as np
    import torch

    import tc

    rng = np.random.default_rng(seed)
    sizes = np.round(np.exp(rng.uniform(np.log(2 ** 10), np.log(2 ** 20), n_clouds))).astype(np.uint32)
    total = int(sizes.sum())
    host = np.empty((total, 4), np.float32)
    off = 0
    for i, s in enumerate(sizes):
        s = int(s)
        if i % 2 == 0:
            host[off:off + s] = tc.gen_uniform(s, seed + i)
        else:
            host[off:off + s] = tc.gen_gaussian_mixture(s, seed + i, 16, 4.0)
        off += s
    pts = torch.from_numpy(host).cuda()
    stream = torch.cuda.Stream()
    batch = tc.Batch(ctx, n_clouds, total, 30)
    with torch.cuda.stream(stream):
        batch.set(sizes, stream=stream)
        batch.capture(pts, stream=stream)
    stream.synchronize()

    # Batched: graph replays of the whole segmented pipeline.
    lat = []
    for i in range(steps + 2):
        t0 = time.perf_counter()
        with torch.cuda.stream(stream):
            batch.replay(stream=stream)
        stream.synchronize()
        if i >= 2:
            lat.append(time.perf_counter() - t0)
    batched_s = float(np.mean(lat))
    if not plain:
        batch.close()
        return {"total_points": total, "batched_ms": round(batched_s * 1e3, 3),
                "batched_Mpoints_per_s": round(total / batched_s / 1e6, 1)}

    # Plain: one tc_build_tree per cloud, one cloud at a time, from the native
    # loop (tc_plain_run_trace with every cloud arriving at once: each build is
    # queued and waited for in C++, no Python between clouds).
    offs = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    views = [pts[offs[i]:offs[i + 1]] for i in range(n_clouds)]
    plain_lat, per_cloud = [], []
    for rep in range(max(2, steps // 4)):
        done = tc.plain_run_trace(ctx, np.zeros(n_clouds), views, sizes)
        plain_lat.append(float(done.max()))
        per_cloud = np.diff(np.concatenate([[0.0], np.sort(done)]))  # each build's own time
    plain_s = float(np.median(plain_lat))
    batch.close()
    return {
        "workload": f"{n_clouds} clouds, sizes log-uniform in [2^10, 2^20] (total {total} points), alternating "
                    "uniform / 16-cluster Gaussian (sigma 4) in [0,1024)^3, 30-bit codes",
        "total_points": total,
        "batched_ms": round(batched_s * 1e3, 3),
        "batched_Mpoints_per_s": round(total / batched_s / 1e6, 1),
        "batched_p99_ms": round(float(np.percentile(lat, 99)) * 1e3, 3),
        "plain_loop": "tc_plain_run_trace: one tc_build_tree per cloud, each waited for, from C++",
        "plain_loop_ms": round(plain_s * 1e3, 3),
        "plain_Mpoints_per_s": round(total / plain_s / 1e6, 1),
        "plain_cloud_p50_ms": round(float(np.percentile(per_cloud, 50)) * 1e3, 4),
        "plain_cloud_p99_ms": round(float(np.percentile(per_cloud, 99)) * 1e3, 4),
        "speedup": round(plain_s / batched_s, 2),
    }


def served_pool(seed, pool):
    """The served workload's request pool: `pool` clouds resident in HBM, sizes
    log-uniform in [2^12, 2^22], uniform xyz in [0,1024)^3."""
    import numpy as np
    import torch

    import tc

    rng = np.random.default_rng(seed)
    sizes = np.round(np.exp(rng.uniform(np.log(2 ** 12), np.log(2 ** 22), pool))).astype(np.uint32)
    pts = [torch.from_numpy(tc.gen_uniform(int(s), seed + i)).cuda() for i, s in enumerate(sizes)]
    return rng, sizes, pts


def bench_served_overload(ctx, rate, n_streams, max_batch_points, requests=2000, pool=256, seed=0xC0FFEE05):
    """The server alone on one Poisson trace of the served workload (no plain
    leg), with a chosen stream count and batch size: the run to profile when
    a configuration's throughput under load is in question."""
    import numpy as np

    import tc

    rng, sizes, pts = served_pool(seed, pool)
    srv = tc.Server(ctx, n_streams=n_streams, max_batch_points=max_batch_points)
    try:
        warm = np.arange(8) * 1e-4
        srv.run_trace(warm, pts[:8], sizes[:8])
        t = np.cumsum(rng.exponential(1.0 / rate, requests))
        t -= t[0]
        idx = rng.integers(0, pool, requests)
        srv.reset_stats()
        lat = srv.run_trace(t, [pts[i] for i in idx], sizes[idx])
        st = srv.stats()
    finally:
        srv.close()
    return {"rate_req_s": rate, "n_streams": n_streams, "max_batch_points": max_batch_points,
            "requests": requests, "batches": st["batches"],
            "Mpoints_per_s": round(float(sizes[idx].sum()) / float(np.max(t + lat)) / 1e6, 1),
            "p50_ms": round(float(np.percentile(lat, 50)) * 1e3, 3),
            "p99_ms": round(float(np.percentile(lat, 99)) * 1e3, 3)}


def bench_served(ctx, rates=(1000, 2000, 5000, 7500, 10000, 12000, 14000, 16000, 18000, 20000, 50000), pool=256,
                 max_req=2000, window_s=0.5, slo_ms=10.0, seed=0xC0FFEE05):
    """Poisson arrivals of build requests (sizes log-uniform in [2^12, 2^22],
    uniform xyz in [0,1024)^3) against the server (64 GiB budget, 2 streams) and
    against a one-request-at-a-time plain loop, on the same traces.  Inputs are
    a pool of distinct clouds resident in HBM, cycled through by the trace."""
    import numpy as np
    import torch

    import tc

    rng, sizes, pts = served_pool(seed, pool)
    srv = tc.Server(ctx)  # defaults: 64 GiB budget, 2 streams, <= 256 clouds / 2^24 points per batch
    # Warm up both paths.
    warm = np.arange(8) * 1e-4
    srv.run_trace(warm, pts[:8], sizes[:8])
    tc.plain_run_trace(ctx, warm, pts[:8], sizes[:8])
    out = {"workload": "Poisson arrivals, request sizes log-uniform in [2^12, 2^22] (uniform xyz in [0,1024)^3, "
                       "30-bit codes), 64 GiB budget, 2 streams; latency = result ready - arrival",
           "slo_p99_ms": slo_ms, "rates": []}
    for lam in rates:
        k = int(min(max_req, lam * window_s))
        t = np.cumsum(rng.exponential(1.0 / lam, k))
        t -= t[0]
        idx = rng.integers(0, pool, k)
        req_pts = [pts[i] for i in idx]
        req_n = sizes[idx]
        total = float(req_n.sum())
        srv.reset_stats()
        lat_s = srv.run_trace(t, req_pts, req_n)
        st = srv.stats()
        lat_p = tc.plain_run_trace(ctx, t, req_pts, req_n)
        span_s = float(np.max(t + lat_s))
        span_p = float(np.max(t + lat_p))
        out["rates"].append({
            "rate_req_s": lam, "requests": k, "points": int(total),
            "served": {"p50_ms": round(float(np.percentile(lat_s, 50)) * 1e3, 3),
                       "p99_ms": round(float(np.percentile(lat_s, 99)) * 1e3, 3),
                       "Mpoints_per_s": round(total / span_s / 1e6, 1),
                       "mean_clouds_per_batch": round(st["mean_clouds_per_batch"], 2)},
            "plain": {"p50_ms": round(float(np.percentile(lat_p, 50)) * 1e3, 3),
                      "p99_ms": round(float(np.percentile(lat_p, 99)) * 1e3, 3),
                      "Mpoints_per_s": round(total / span_p / 1e6, 1)},
        })
    def trial(kind, lam):
        k = int(min(max_req, lam * window_s))
        t = np.cumsum(rng.exponential(1.0 / lam, k))
        t -= t[0]
        idx = rng.integers(0, pool, k)
        req_n = sizes[idx]
        lat = srv.run_trace(t, [pts[i] for i in idx], req_n) if kind == "served" else \
            tc.plain_run_trace(ctx, t, [pts[i] for i in idx], req_n)
        return float(np.percentile(lat, 99)) * 1e3, float(req_n.sum()) / float(np.max(t + lat)) / 1e6

    def bisect(kind, steps=6):
        """Highest rate whose p99 meets the SLO, between the last grid rate that
        meets it and the next one that does not; returns (rate, Mpoints/s, p99)."""
        grid = out["rates"]
        ok = [r for r in grid if r[kind]["p99_ms"] <= slo_ms]
        if not ok:
            return None
        lo = max(ok, key=lambda r: r["rate_req_s"])
        above = [r for r in grid if r["rate_req_s"] > lo["rate_req_s"]]
        best = (lo["rate_req_s"], lo[kind]["Mpoints_per_s"], lo[kind]["p99_ms"])
        if not above:
            return best
        lo_r, hi_r = float(lo["rate_req_s"]), float(min(r["rate_req_s"] for r in above))
        for _ in range(steps):
            mid = 0.5 * (lo_r + hi_r)
            p99, mpts = trial(kind, mid)
            if p99 <= slo_ms:
                lo_r = mid
                if mpts > best[1]:
                    best = (mid, mpts, p99)
            else:
                hi_r = mid
        return best

    cap = {}
    for kind in ("served", "plain"):
        b = bisect(kind)
        cap[kind] = b
        out[f"{kind}_capacity_at_slo"] = None if b is None else {
            "rate_req_s": round(b[0], 1), "Mpoints_per_s": round(b[1], 1), "p99_ms": round(b[2], 3)}
    srv.close()
    out["served_Mpoints_per_s_at_slo"] = round(cap["served"][1], 1) if cap["served"] else 0.0
    out["plain_Mpoints_per_s_at_slo"] = round(cap["plain"][1], 1) if cap["plain"] else 0.0
    out["speedup_at_slo"] = round(out["served_Mpoints_per_s_at_slo"] / out["plain_Mpoints_per_s_at_slo"], 2) \
        if out["plain_Mpoints_per_s_at_slo"] else None
    out["capacity_method"] = ("highest rate with p99 <= SLO: the rate grid, then 6 bisection steps between the "
                              "last grid rate that meets the SLO and the next that does not (new Poisson traces)")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--log2n", type=int, default=26)
    ap.add_argument("--seed", type=int, default=0xC0FFEE02)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-extras", action="store_true", help="headline build only")
    ap.add_argument("--markers", action="store_true",
                    help="roctx ranges around each launch and build (for rocprofv3 --marker-trace)")
    ap.add_argument("--profile-steps", type=int, default=3)
    ap.add_argument("--no-tune", action="store_true", help="fixed sort launch shape instead of the tuned ones")
    ap.add_argument("--retune", action="store_true", help="run the autotuner instead of loading its saved table")
    ap.add_argument("--batched-only", action="store_true",
                    help="only the batched workload's graph replays (for rocprofv3 kernel statistics)")
    ap.add_argument("--served-overload", type=int, nargs=3, metavar=("RATE", "STREAMS", "MAX_BATCH_POINTS"),
                    help="only the server on one Poisson trace of the served workload at RATE req/s, with "
                         "STREAMS slots and batches of at most MAX_BATCH_POINTS (for rocprofv3 kernel traces)")
    ap.add_argument("--traffic-probe", action="store_true",
                    help="only build the headline tree 3 times, then run reference kernels of known traffic "
                         "(for rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE runs)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              "(clouds are independent; the multi-GPU design shards whole clouds)", file=sys.stderr)
        sys.exit(2)

    import torch

    import tc

    n = 1 << args.log2n
    pts_h = tc.gen_gaussian_mixture(n, args.seed, 64, 8.0, 1024.0)
    ctx = tc.Context(0)
    if args.markers:
        ctx.markers(True)
    tuning = None
    if args.retune:
        tuning = tc.tune.autotune(ctx, n)
    elif not args.no_tune and os.path.exists(TUNING_FILE):
        tc.tune.load(ctx, TUNING_FILE)
        tuning = "loaded " + os.path.relpath(TUNING_FILE, ROOT)
    if args.traffic_probe:
        traffic_probe(ctx, pts_h)
        return
    if args.batched_only:
        print(json.dumps({"batched": bench_batched(ctx, 1024, args.steps, plain=False)}), flush=True)
        return
    if args.served_overload:
        rate, streams, mbp = args.served_overload
        print(json.dumps({"served_overload": bench_served_overload(ctx, rate, streams, mbp)}), flush=True)
        return
    pts = torch.from_numpy(pts_h).cuda()
    tree = tc.Tree(ctx, n, 30)

    step_s = timed(lambda: tree.build(pts, n, 0.0, 1024.0), args.steps, args.warmup)
    counts = tree.read_counts()
    u, o = counts["n_unique"], counts["n_nodes"]
    ms_per_step = step_s * 1e3

    prof = profile_stages(ctx, lambda: tree.build(pts, n, 0.0, 1024.0), args.profile_steps)
    alg = algorithmic_bytes(n, u, o)
    stages = {}
    for name, (avg, per_step) in prof.items():
        entry = {"avg_ms": round(avg, 4), "launches_per_step": per_step}
        if name in alg:
            entry["alg_bytes"] = alg[name]
            entry["GBps"] = round(alg[name] / (avg * 1e-3) / 1e9, 1)
            entry["roofline_frac"] = round(entry["GBps"] / HBM_PEAK_GBS, 4)
        stages[name] = entry
    dom = max((k for k in stages if "alg_bytes" in stages[k]),
              key=lambda k: stages[k]["avg_ms"] * stages[k]["launches_per_step"])
    d = stages[dom]
    build_bytes = sum(alg[k] * stages[k]["launches_per_step"] for k in alg if k in stages)

    out = {
        "metric": METRIC,
        "value": round(n / step_s / 1e6, 2),
        "unit": "Mpoints/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic",
        "config": {
            "workload": f"plain full tree build, 2^{args.log2n} points, 64-centre Gaussian mixture (sigma 8) in "
                        "[0,1024)^3, float32 xyzw, 30-bit Morton codes",
            "n_points": n,
            "unique_codes": u,
            "octree_nodes": o,
            "seed": args.seed,
        },
        "roofline": {
            "bound": "hbm",
            "kernel": dom,
            "achieved": d["GBps"],
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": d["roofline_frac"],
            "traffic": measured_traffic() if (args.log2n == 26 and dom == "binning") else None,
            "alg_bytes_per_launch": d["alg_bytes"],
            "avg_launch_ms": d["avg_ms"],
        },
        "achievable": None,
        "build_roofline": {
            "alg_bytes": build_bytes,
            "GBps": round(build_bytes / step_s / 1e9, 1),
            "frac": round(build_bytes / step_s / 1e9 / HBM_PEAK_GBS, 4),
        },
        "stages": stages,
        "sort_shape": list(tc.tune.shape(ctx, 32, False, n)),
        "sort_tuning": tuning,
    }
    tree.free()
    ach = achievable_bandwidth(ctx)
    out["achievable"] = dict(ach, kernel="tc_traffic_reference: copies of 256 and 512 MiB (bytes read + written) "
                                          "and a read of 1 GiB; non-temporal 16-byte loads, 16-byte stores",
                             binning_frac_of_copy=round(d["GBps"] / ach["copy_256MiB_GBps"], 4),
                             build_frac_of_copy=round(out["build_roofline"]["GBps"] / ach["copy_512MiB_GBps"], 4))
    if not args.no_extras:
        # The standalone sort of the same codes first, as the build's own passes are timed.
        codes = torch.empty(n, dtype=torch.int32, device="cuda")
        ctx.morton30(pts, codes, 0.0, 1024.0)
        out["sort"] = bench_sort_u32(ctx, codes, 20)
        del codes
        # Same build with the point -> leaf map (index payload through the sort,
        # run starts from unique).
        itree = tc.Tree(ctx, n, 30, point_index=True)
        ims = timed(lambda: itree.build(pts, n, 0.0, 1024.0), max(2, args.steps // 2), 1) * 1e3
        itree.free()
        out["indexed_build"] = {
            "workload": "the headline build plus the point -> leaf map (u32 point index sorted with the codes, "
                        "run starts from unique)",
            "ms": round(ims, 4),
            "Mpoints_per_s": round(n / (ims * 1e-3) / 1e6, 2),
        }
        out["pcie_inclusive"] = bench_pcie(ctx, pts_h)
        out["config1"] = bench_config1(ctx)
        del pts
        torch.cuda.empty_cache()
        out["sort_pairs_u64"] = bench_sort_pairs_u64(ctx, 27, 2)
        torch.cuda.empty_cache()
        out["batched"] = bench_batched(ctx, 1024, max(4, args.steps // 2))
        torch.cuda.empty_cache()
        out["served"] = bench_served(ctx)
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = run_cpu_baseline(pts_h)
    ctx.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
