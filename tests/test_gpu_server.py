"""Serving on the GPU: every request's result equals the oracle's single-cloud
build of the same points (bit for bit, every array; the per-cloud contract of
src/main.rs:60-170), alone and in a mix of concurrent requests, up to the
served workload's request sizes (2^12 .. 2^22 points); admission limits and
budget accounting."""
import threading
import time

import numpy as np
import pytest

from gpu_util import dev

pytestmark = [pytest.mark.gpu, pytest.mark.timeout(90)]

ARRAYS = ("unique", "prefix_n", "has_leaf_left", "has_leaf_right", "left_child", "parent", "edge_count", "offsets")


def plain_build(ctx, pts, code_bits=30, rng=1024.0):
    import tc

    n = len(pts)
    tree = tc.Tree(ctx, max(n, 1), code_bits)
    try:
        tree.build(dev(pts), n, 0.0, rng)
        return tree.download()
    finally:
        tree.free()


def oracle_tree(oracle, pts, code_bits=30, rng=1024.0):
    w = oracle.build_tree(pts, 0.0, rng, code_bits)
    w["n_unique"] = len(w["unique"])
    w["n_nodes"] = len(w["nodes"])
    return w


def assert_same_tree(got, want, tag=""):
    assert got["n_unique"] == want["n_unique"], tag
    assert got["n_nodes"] == want["n_nodes"], tag
    for k in ARRAYS:
        assert np.array_equal(got[k], want[k]), (tag, k)
    assert np.array_equal(got["nodes"].view(np.uint8).reshape(-1), want["nodes"].view(np.uint8).reshape(-1)), \
        (tag, "nodes")


def make_clouds(tclib, k, seed, lo=2 ** 8, hi=2 ** 17):
    r = np.random.default_rng(seed)
    sizes = np.round(np.exp(r.uniform(np.log(lo), np.log(hi), k))).astype(int)
    sizes[:4] = [0, 1, 2, 3]
    clouds = []
    for i, s in enumerate(sizes):
        if i % 2:
            clouds.append(tclib.gen_gaussian_mixture(int(s), seed * 1000 + i, 16, 4.0))
        else:
            clouds.append(tclib.gen_uniform(int(s), seed * 1000 + i))
    return clouds


def collect(srv, ids, timeout_s=60.0):
    out = {}
    t_end = time.time() + timeout_s
    while len(out) < len(ids):
        assert time.time() < t_end, "server did not finish"
        for res in srv.poll(64, wait_ms=50):
            out[res["id"]] = res
    return out


def test_served_results_equal_oracle_builds(gpu_ctx, oracle, tclib):
    import tc

    clouds = make_clouds(tclib, 48, seed=7)
    srv = tc.Server(gpu_ctx, mem_budget_bytes=16 << 30, n_streams=2, max_batch_clouds=16,
                    max_batch_points=1 << 20)
    try:
        d_pts = [dev(c) for c in clouds]
        ids = [srv.submit(p, len(c)) for p, c in zip(d_pts, clouds)]
        got = collect(srv, ids)
        st = srv.stats()
        assert st["completed"] == len(clouds) and st["batches"] < len(clouds)
        for i, (rid, c) in enumerate(zip(ids, clouds)):
            res = got[rid]
            assert res["status"] == 0 and res["n_points"] == len(c)
            assert res["t_submit"] <= res["t_start"] <= res["t_done"]
            assert_same_tree(srv.download(res), oracle_tree(oracle, c), f"cloud {i} n={len(c)}")
            srv.release(rid)
        assert srv.stats()["reserved_bytes"] == 0
    finally:
        srv.close()


def test_result_same_alone_and_in_a_concurrent_mix(gpu_ctx, oracle, tclib):
    import tc

    probe = tclib.gen_gaussian_mixture(150_000, 99, 16, 4.0)
    others = make_clouds(tclib, 40, seed=8, hi=2 ** 18)
    srv = tc.Server(gpu_ctx, mem_budget_bytes=16 << 30, n_streams=3, max_batch_clouds=8,
                    max_batch_points=1 << 21)
    try:
        d_probe = dev(probe)
        rid = srv.submit(d_probe)
        alone = collect(srv, [rid])[rid]
        alone_tree = srv.download(alone)
        srv.release(rid)

        d_others = [dev(c) for c in others]
        ids = []
        index_of = {}
        lock = threading.Lock()

        def client(lo, hi, seed):
            r = np.random.default_rng(seed)
            for j in range(lo, hi):
                time.sleep(float(r.uniform(0, 2e-4)))
                with lock:
                    rid = srv.submit(d_others[j], len(others[j]))
                    ids.append(rid)
                    index_of[rid] = j

        threads = [threading.Thread(target=client, args=(k * 10, k * 10 + 10, k)) for k in range(4)]
        for t in threads:
            t.start()
        time.sleep(5e-4)
        probe_id = srv.submit(d_probe)
        for t in threads:
            t.join()
        got = collect(srv, ids + [probe_id])
        assert_same_tree(srv.download(got[probe_id]), alone_tree, "probe in mix")
        assert_same_tree(alone_tree, oracle_tree(oracle, probe), "probe alone vs oracle")
        assert_same_tree(alone_tree, plain_build(gpu_ctx, probe), "probe alone vs plain")
        for j, rid in enumerate(ids):
            c = others[index_of[rid]]
            assert_same_tree(srv.download(got[rid]), oracle_tree(oracle, c), f"mix request {j} n={len(c)}")
        for rid in got:
            srv.release(rid)
    finally:
        srv.close()


def test_large_batches_overlapping_on_two_streams(gpu_ctx, oracle, tclib):
    """Two slots whose batches of 2^19 .. 2^21-point clouds run at once: the
    kernels of the two streams share the CUs, so a workgroup may start before
    a lower-numbered one of its own launch, and the sort and scan look-backs
    then compute the stalled predecessor tiles themselves.  Every result equals
    a plain build of the same cloud (and the oracle's, for the smallest)."""
    import tc

    clouds = make_clouds(tclib, 24, seed=19, lo=2 ** 19, hi=2 ** 21)[4:]
    srv = tc.Server(gpu_ctx, mem_budget_bytes=24 << 30, n_streams=2, max_batch_clouds=32,
                    max_batch_points=1 << 23)
    try:
        d_pts = [dev(c) for c in clouds]
        ids = [srv.submit(p, len(c)) for p, c in zip(d_pts, clouds)]
        got = collect(srv, ids)
        assert srv.stats()["batches"] >= 3
        small = int(np.argmin([len(c) for c in clouds]))
        for i, (rid, c) in enumerate(zip(ids, clouds)):
            res = got[rid]
            assert res["status"] == 0, i
            tree = srv.download(res)
            assert_same_tree(tree, plain_build(gpu_ctx, c), f"cloud {i} n={len(c)}")
            if i == small:
                assert_same_tree(tree, oracle_tree(oracle, c), f"cloud {i} vs oracle")
            srv.release(rid)
    finally:
        srv.close()


def test_two_slots_under_overload_keep_their_throughput(gpu_ctx, tclib):
    """Two slots with batches of up to 2^25 points under a 50k req/s Poisson
    trace of 2^12 .. 2^22-point requests.  While the look-back kernels could
    wait on workgroups that the other slot's kernel kept from starting, this
    ran at 2.7-6.7 Gpoints/s (launches of 50-310 ms); with the waiting
    workgroup computing the stalled tile it runs at about 17.  The floor of 8
    Gpoints/s fails on that regression, far from timing noise."""
    import tc

    rng = np.random.default_rng(0xC0FFEE05)
    sizes = np.round(np.exp(rng.uniform(np.log(2 ** 12), np.log(2 ** 22), 96))).astype(np.uint32)
    pts = [dev(tclib.gen_uniform(int(s), 500 + i)) for i, s in enumerate(sizes)]
    srv = tc.Server(gpu_ctx, n_streams=2, max_batch_points=1 << 25)
    try:
        srv.run_trace(np.arange(8) * 1e-4, pts[:8], sizes[:8])
        k = 1500
        t = np.cumsum(rng.exponential(1.0 / 50_000, k))
        t -= t[0]
        idx = rng.integers(0, len(pts), k)
        lat = srv.run_trace(t, [pts[i] for i in idx], sizes[idx])
        gpts = float(sizes[idx].sum()) / float(np.max(t + lat)) / 1e9
        assert gpts > 8.0, f"{gpts:.2f} Gpoints/s"
    finally:
        srv.close()


def test_63bit_served(gpu_ctx, oracle, tclib):
    import tc

    clouds = [tclib.gen_uniform(n, 60 + n, 2097152.0) for n in (5000, 70_000, 1)]
    srv = tc.Server(gpu_ctx, mem_budget_bytes=8 << 30, max_batch_clouds=8, max_batch_points=1 << 18, code_bits=63)
    try:
        d = [dev(c) for c in clouds]
        ids = [srv.submit(p, len(c), 0.0, 2097152.0) for p, c in zip(d, clouds)]
        got = collect(srv, ids)
        for rid, c in zip(ids, clouds):
            assert_same_tree(srv.download(got[rid]), oracle_tree(oracle, c, 63, 2097152.0), f"n={len(c)}")
            srv.release(rid)
    finally:
        srv.close()


def test_admission_limits(gpu_ctx, tclib):
    import tc

    srv = tc.Server(gpu_ctx, mem_budget_bytes=4 << 30, max_batch_clouds=4, max_batch_points=1 << 16)
    try:
        big = dev(tclib.gen_uniform(1 << 17, 5))
        with pytest.raises(tc.TcError) as e:
            srv.submit(big)
        assert e.value.status == -6  # TC_ERR_CAPACITY
        assert srv.stats()["rejected"] == 1
        with pytest.raises(tc.TcError):
            srv.release(12345)  # unknown id
    finally:
        srv.close()
    # A budget that leaves no room for results cannot be created.
    ws = int(tc.lib().tc_batch_memory_bytes(4, 1 << 16, 30))
    with pytest.raises(tc.TcError) as e:
        tc.Server(gpu_ctx, mem_budget_bytes=ws, n_streams=1, max_batch_clouds=4, max_batch_points=1 << 16)
    assert e.value.status == -5  # TC_ERR_OVER_BUDGET


def test_budget_holds_back_admission_until_release(gpu_ctx, tclib):
    import tc

    n = 1 << 16
    ws = int(tc.lib().tc_batch_memory_bytes(4, 1 << 17, 30))
    # Room for one worst-case result: after the first build its result is cut
    # back to its actual size, which still leaves too little for the next.
    srv = tc.Server(gpu_ctx, mem_budget_bytes=ws + tc.result_bytes(n) + 4096, n_streams=1, max_batch_clouds=4,
                    max_batch_points=1 << 17)
    try:
        d = [dev(tclib.gen_uniform(n, 300 + i)) for i in range(3)]
        ids = [srv.submit(p) for p in d]
        for k in range(3):
            got = collect(srv, ids[k:k + 1])
            time.sleep(0.05)
            assert srv.poll(8) == []  # the next one waits for memory
            st = srv.stats()
            assert 0 < st["reserved_bytes"] <= st["result_budget_bytes"]
            srv.release(ids[k])
            assert got[ids[k]]["status"] == 0
        assert srv.stats()["reserved_bytes"] == 0
    finally:
        srv.close()


def test_trace_runners(gpu_ctx, tclib):
    import tc

    r = np.random.default_rng(3)
    k = 60
    t = np.cumsum(r.exponential(1 / 2000.0, k))
    sizes = np.round(np.exp(r.uniform(np.log(2 ** 12), np.log(2 ** 18), k))).astype(np.uint32)
    pts = [dev(tclib.gen_uniform(int(s), 500 + i)) for i, s in enumerate(sizes)]
    srv = tc.Server(gpu_ctx, mem_budget_bytes=8 << 30, max_batch_clouds=32, max_batch_points=1 << 21)
    try:
        lat = srv.run_trace(t, pts, sizes)
        assert np.all(lat > 0) and np.all(lat < 5.0)
        assert srv.stats()["completed"] == k
    finally:
        srv.close()
    lat_p = tc.plain_run_trace(gpu_ctx, t, pts, sizes)
    assert np.all(lat_p > 0) and np.all(lat_p < 5.0)


def test_served_workload_sizes_poisson_match_oracle(gpu_ctx, oracle, tclib):
    """The served workload's request mix: sizes log-uniform in [2^12, 2^22],
    uniform xyz, Poisson arrivals at 50k req/s from a submitting thread, the
    server's default configuration (64 GiB budget, 2 streams).  The arrivals
    outpace the builds, so requests are batched (mean >= 8 per batch); every
    result is compared with the oracle's build of its points."""
    from concurrent.futures import ThreadPoolExecutor

    import tc

    r = np.random.default_rng(0xC0FFEE05)
    k = 128
    sizes = np.round(np.exp(r.uniform(np.log(2 ** 12), np.log(2 ** 22), k))).astype(np.uint32)
    sizes[0] = 1 << 22  # the largest request of the range
    clouds = [tclib.gen_uniform(int(s), 0xC0FFEE05 + i) for i, s in enumerate(sizes)]
    t_arr = np.cumsum(r.exponential(1 / 50_000.0, k))
    t_arr -= t_arr[0]
    d = [dev(c) for c in clouds]
    srv = tc.Server(gpu_ctx)
    try:
        ids = [0] * k
        t0 = time.perf_counter()
        for i in range(k):
            while time.perf_counter() - t0 < t_arr[i]:
                pass
            ids[i] = srv.submit(d[i], int(sizes[i]))
        got = collect(srv, ids, timeout_s=120.0)
        st = srv.stats()
        print(f"served workload: {k} requests, {int(sizes.sum())} points, {st['batches']} batches, "
              f"{st['mean_clouds_per_batch']:.1f} clouds per batch, p99 {st['latency_p99_ms']:.2f} ms", flush=True)
        assert st["completed"] == k
        assert st["mean_clouds_per_batch"] >= 8.0
        with ThreadPoolExecutor(16) as ex:
            want = list(ex.map(lambda c: oracle_tree(oracle, c), clouds))
        for i, rid in enumerate(ids):
            res = got[rid]
            assert res["status"] == 0 and res["n_points"] == sizes[i]
            assert_same_tree(srv.download(res), want[i], f"request {i} n={sizes[i]}")
            srv.release(rid)
        assert srv.stats()["reserved_bytes"] == 0
    finally:
        srv.close()


def test_backpressure_and_cancellation(gpu_ctx, oracle, tclib):
    """The live server's bounded admission queue and cancellation: with one
    stream busy on a large build, submissions beyond max_queued waiting
    requests get TC_ERR_BUSY; a waiting request can be cancelled (it comes back
    with TC_ERR_CANCELLED and no arrays), and the others still equal the oracle."""
    import tc

    big = dev(tclib.gen_uniform(1 << 22, 901))
    small = [tclib.gen_uniform(5000 + 100 * i, 910 + i) for i in range(10)]
    d_small = [dev(c) for c in small]
    srv = tc.Server(gpu_ctx, mem_budget_bytes=16 << 30, n_streams=1, max_batch_clouds=1,
                    max_batch_points=1 << 22, max_queued=4)
    try:
        first = srv.submit(big)
        ids, busy = [], 0
        for i, d in enumerate(d_small):
            try:
                ids.append((i, srv.submit(d)))
            except tc.TcError as e:
                assert e.status == tc._lib.TC_ERR_BUSY
                busy += 1
        assert len(ids) <= 5 and busy >= 5  # at most the queue bound (+1 if the big build had not started)
        # Cancel the last accepted request: still waiting behind the big build.
        i_c, id_c = ids[-1]
        cancelled = True
        try:
            srv.cancel(id_c)
        except tc.TcError as e:  # it could only have started if the big build were already done
            assert e.status in (tc._lib.TC_ERR_BUSY, tc._lib.TC_ERR_NOT_FOUND)
            cancelled = False
        with pytest.raises(tc.TcError):
            srv.cancel(10 ** 9)  # unknown id
        got = collect(srv, [first] + [r for _, r in ids])
        st = srv.stats()
        if cancelled:
            assert got[id_c]["status"] == tc._lib.TC_ERR_CANCELLED and got[id_c]["n_nodes"] == 0
            assert st["cancelled"] == 1
        assert st["rejected"] == busy and st["queued"] == 0
        for i, rid in ids:
            if cancelled and rid == id_c:
                continue
            assert got[rid]["status"] == 0
            assert_same_tree(srv.download(got[rid]), oracle_tree(oracle, small[i]), f"request {i}")
            srv.release(rid)
        srv.release(first)
        assert srv.stats()["reserved_bytes"] == 0
    finally:
        srv.close()
