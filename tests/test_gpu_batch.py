"""Batched (segmented) builds: every cloud's outputs inside a batch equal the
oracle's single-cloud build, eagerly and under graph replay."""
import numpy as np
import pytest

from gpu_util import dev

pytestmark = pytest.mark.gpu

KEYS = ("sorted", "unique", "prefix_n", "has_leaf_left", "has_leaf_right", "left_child", "parent", "edge_count",
        "offsets")


def make_clouds(tclib, sizes, seed, extent=1024.0):
    clouds = []
    for i, n in enumerate(sizes):
        if i % 2 == 0:
            clouds.append(tclib.gen_uniform(n, seed + i, extent))
        else:
            clouds.append(tclib.gen_gaussian_mixture(n, seed + i, 16, 4.0 * extent / 1024.0, extent))
    return clouds


def check_cloud(oracle, got, pts, mn, rg, bits, tag):
    want = oracle.build_tree(pts, mn, rg, code_bits=bits)
    for k in KEYS:
        g, w = np.asarray(got[k]), np.asarray(want[k])
        assert g.shape == w.shape, (tag, k, g.shape, w.shape)
        assert np.array_equal(g, w), (tag, k)
    assert np.array_equal(got["nodes"].view(np.uint8).reshape(-1), want["nodes"].view(np.uint8).reshape(-1)), tag


def run_batch(batch, clouds, mins=None, ranges=None, graph=False, stream=None):
    import torch

    counts = np.array([len(c) for c in clouds], np.uint32)
    allp = np.concatenate(clouds) if len(clouds) else np.zeros((0, 4), np.float32)
    pts = dev(allp if len(allp) else np.zeros((1, 4), np.float32))
    batch.set(counts, mins, ranges, stream=stream)
    if graph:
        batch.capture(pts, stream=stream)
        batch.replay(stream=stream)
    else:
        batch.build(pts, stream=stream)
    torch.cuda.synchronize()
    return pts


def test_batch_matches_single_cloud_builds(gpu_ctx, oracle, tclib):
    import tc

    sizes = [1000, 1, 2, 0, 3, 4096, 4097, 8191, 12000, 777, 65536, 5, 30000, 4095, 2, 100_000, 1024]
    clouds = make_clouds(tclib, sizes, 500)
    b = tc.Batch(gpu_ctx, 64, 400_000, 30)
    run_batch(b, clouds)
    got = b.download()
    assert len(got) == len(clouds)
    for i, (g, pts) in enumerate(zip(got, clouds)):
        check_cloud(oracle, g, pts, 0.0, 1024.0, 30, f"cloud {i} n={len(pts)}")
    b.close()


def test_batch_per_cloud_min_range(gpu_ctx, oracle, tclib):
    import tc

    clouds = make_clouds(tclib, [5000, 7000, 300], 900, 1.0)
    mins = np.array([-3.0, 10.0, 0.5], np.float32)
    ranges = np.array([7.0, 1.0, 2.0], np.float32)
    clouds = [(c * np.array([r, r, r, 1], np.float32) + np.array([m, m, m, 0], np.float32)).astype(np.float32)
              for c, m, r in zip(clouds, mins, ranges)]
    b = tc.Batch(gpu_ctx, 8, 20_000, 30)
    run_batch(b, clouds, mins, ranges)
    for i, (g, pts) in enumerate(zip(b.download(), clouds)):
        check_cloud(oracle, g, pts, float(mins[i]), float(ranges[i]), 30, f"cloud {i}")
    b.close()


def test_batch_63bit(gpu_ctx, oracle, tclib):
    import tc

    clouds = make_clouds(tclib, [3000, 1, 20000, 4500], 321, 2097152.0)
    b = tc.Batch(gpu_ctx, 8, 40_000, 63)
    run_batch(b, clouds, None, np.full(4, 2097152.0, np.float32))
    for i, (g, pts) in enumerate(zip(b.download(), clouds)):
        check_cloud(oracle, g, pts, 0.0, 2097152.0, 63, f"cloud {i}")
    b.close()


def test_graph_replay_new_batches(gpu_ctx, oracle, tclib):
    """One captured graph replays for different batches (new plan uploads); every
    replay is checked, with eager work and synchronisations in between."""
    import torch

    import tc

    stream = torch.cuda.Stream()
    cap_points = 200_000
    b = tc.Batch(gpu_ctx, 32, cap_points, 30)
    buf = torch.zeros(cap_points * 4, dtype=torch.float32, device="cuda")
    first = make_clouds(tclib, [10_000, 20_000, 5], 11)
    counts = np.array([len(c) for c in first], np.uint32)
    with torch.cuda.stream(stream):
        allp = np.concatenate(first)
        buf[: allp.size].copy_(torch.from_numpy(allp.reshape(-1)))
        b.set(counts, stream=stream)
        b.capture(buf, stream=stream)
    for rep, sizes in enumerate([[10_000, 20_000, 5], [1, 50_000, 3000, 0, 64_000], [150_000], [7, 8, 9]]):
        clouds = make_clouds(tclib, sizes, 1000 + rep)
        allp = np.concatenate(clouds)
        with torch.cuda.stream(stream):
            buf[: allp.size].copy_(torch.from_numpy(allp.reshape(-1)))
            b.set(np.array(sizes, np.uint32), stream=stream)
            b.replay(stream=stream)
        stream.synchronize()
        _ = torch.ones(1000, device="cuda").sum().item()  # unrelated eager work between replays
        got = b.download(stream)
        for i, (g, pts) in enumerate(zip(got, clouds)):
            check_cloud(oracle, g, pts, 0.0, 1024.0, 30, f"replay {rep} cloud {i}")
    b.close()


def test_batch_rejects_over_capacity(gpu_ctx):
    import tc

    b = tc.Batch(gpu_ctx, 4, 1000, 30)
    with pytest.raises(tc.TcError):
        b.set(np.array([600, 600], np.uint32))
    with pytest.raises(tc.TcError):
        b.set(np.ones(5, np.uint32))
    b.close()


@pytest.mark.parametrize("limit", [0, 1])
def test_batch_lookback_fallback_path(gpu_ctx, oracle, tclib, limit):
    """Segmented sort and scans with predecessor tiles computed by the waiting
    workgroup (all of them at 0, most at 1): same trees as the oracle's."""
    import tc

    L = tc.lib()
    clouds = make_clouds(tclib, [20_000, 3, 50_000, 1, 9000, 130_000], 4242)
    assert L.tc_set_lookback_spin_limit(gpu_ctx.handle, limit) == 0
    try:
        b = tc.Batch(gpu_ctx, 8, 250_000, 30)
        run_batch(b, clouds)
        for i, (g, pts) in enumerate(zip(b.download(), clouds)):
            check_cloud(oracle, g, pts, 0.0, 1024.0, 30, f"cloud {i}")
        b.close()
    finally:
        assert L.tc_set_lookback_spin_limit(gpu_ctx.handle, tc._lib.LOOKBACK_SPIN_DEFAULT) == 0


def test_batch_clouds_at_unique_tile_edges(gpu_ctx, oracle, tclib):
    """Clouds whose starts fall on and around the segmented unique's 16384-key
    tile edges, empty clouds among them, and a stretch of 3000 clouds of 0-5
    points (hundreds of cloud starts in one tile), against the oracle."""
    import tc

    r = np.random.default_rng(99)
    tiny = r.integers(0, 6, size=3000).tolist()
    sizes = [16384, 16383, 1, 16385, 0, 8192, 8191, 2, 0, 0, 16384, 40_000] + tiny + [16384, 3]
    clouds = make_clouds(tclib, sizes, 777)
    b = tc.Batch(gpu_ctx, len(sizes), sum(sizes), 30)
    try:
        run_batch(b, clouds)
        for i, (g, pts) in enumerate(zip(b.download(), clouds)):
            check_cloud(oracle, g, pts, 0.0, 1024.0, 30, f"cloud {i} ({len(pts)} points)")
    finally:
        b.close()
