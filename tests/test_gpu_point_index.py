"""Point -> leaf map on the GPU (sort with the point index as payload, unique
with run starts), against the oracle and, at full size, against torch's
stable sort.

The map has no reference counterpart (the reference sorts codes only,
src/main.rs:85); the oracle's version is pinned in test_oracle.py by numpy
(np.argsort kind="stable", np.unique return_index).
"""
import numpy as np
import pytest

from gpu_util import dev, empty, host

pytestmark = pytest.mark.gpu

TREE_KEYS = ("sorted", "unique", "prefix_n", "has_leaf_left", "has_leaf_right", "left_child", "parent",
             "edge_count", "offsets")


def _points(tclib, kind, n, seed):
    if kind == "uniform":
        return tclib.gen_uniform(n, seed)
    return tclib.gen_gaussian_mixture(n, seed, 16, 0.5)  # many points per 30-bit cell


@pytest.mark.parametrize("bits", [30, 63])
@pytest.mark.parametrize("kind", ["uniform", "clustered"])
@pytest.mark.parametrize("n", [0, 1, 2, 3, 1000, 8193, 100_000])
def test_indexed_build_matches_oracle(gpu_ctx, oracle, tclib, bits, kind, n):
    import tc

    pts = _points(tclib, kind, n, 100 + n)
    want = oracle.build_tree(pts, 0.0, 1024.0, bits, point_index=True)
    tree = tc.Tree(gpu_ctx, max(n, 1), bits, point_index=True)
    plain = tc.Tree(gpu_ctx, max(n, 1), bits)
    try:
        d = dev(pts) if n else dev(np.zeros((1, 4), np.float32))
        tree.build(d, n, 0.0, 1024.0)
        got = tree.download()
        plain.build(d, n, 0.0, 1024.0)
        ref = plain.download()
    finally:
        tree.free()
        plain.free()
    for k in TREE_KEYS:
        assert np.array_equal(got[k], want[k]), k
        assert np.array_equal(got[k], ref[k]), k  # the payload changes nothing else
    assert np.array_equal(got["nodes"].view(np.uint8), want["nodes"].view(np.uint8))
    assert np.array_equal(got["point_index"], want["point_index"])
    assert np.array_equal(got["leaf_first"], want["leaf_first"])


@pytest.mark.parametrize("dt", [np.uint32, np.uint64])
@pytest.mark.parametrize("n", [1, 8191, 8192, 8193, 3 * 8192 + 5, 16383, 16384, 16385, 3 * 16384 + 5, 200_000])
def test_unique_runs_matches_oracle(gpu_ctx, oracle, dt, n):
    r = np.random.default_rng(n)
    keys = np.sort(r.integers(0, max(n // 7, 2), size=n, dtype=np.uint64).astype(dt))
    keys[n // 2:] = keys[n // 2:] << np.array(20, dtype=dt)  # long runs and wide gaps
    keys = np.sort(keys)
    u_want, f_want = oracle.unique_runs(keys)
    d = dev(keys)
    out = empty(n, dt)
    first = empty(n + 1, np.uint32)
    cnt = empty(1, np.uint32)
    cnt.zero_()
    gpu_ctx.unique_runs(d, out, first, cnt, n=n)
    u = int(host(cnt, np.uint32)[0])
    assert u == len(u_want)
    assert np.array_equal(host(out, dt, u), u_want)
    assert np.array_equal(host(first, np.uint32, u + 1), f_want)


@pytest.mark.parametrize("bits", [30, 63])
def test_morton_sort_pairs_matches_oracle(gpu_ctx, oracle, tclib, bits):
    n = 123_457
    pts = tclib.gen_gaussian_mixture(n, 7, 4, 2.0)
    kdt = np.uint32 if bits == 30 else np.uint64
    codes = oracle.compute_morton(pts, 0.0, 1024.0) if bits == 30 else oracle.compute_morton63(pts, 0.0, 1024.0)
    want_k, want_v = oracle.radix_sort(codes, np.arange(n, dtype=np.uint32))
    d = dev(pts)
    k, v, ka, va = empty(n, kdt), empty(n, np.uint32), empty(n, kdt), empty(n, np.uint32)
    gpu_ctx.morton_sort_pairs(d, k, v, ka, va, 0.0, 1024.0)
    assert np.array_equal(host(k, kdt, n), want_k)
    assert np.array_equal(host(v, np.uint32, n), want_v)


def test_indexed_full_size_properties(gpu_ctx, oracle, tclib):
    """2^26 Gaussian-mixture points: the permutation equals torch's stable
    argsort of the encoder's codes and the leaf runs equal the oracle's run
    starts of the sorted codes."""
    import torch

    import tc

    n = 1 << 26
    pts = tclib.gen_gaussian_mixture(n, 0xC0FFEE02, 64, 8.0)
    d_pts = dev(pts)
    tree = tc.Tree(gpu_ctx, n, 30, point_index=True)
    try:
        tree.build(d_pts, n, 0.0, 1024.0)
        got = tree.download()
    finally:
        tree.free()
    codes = torch.empty(n, dtype=torch.int32, device="cuda")
    gpu_ctx.morton30(d_pts, codes, 0.0, 1024.0)
    ref = torch.sort(codes, stable=True)  # 30-bit codes: signed order == unsigned
    assert np.array_equal(got["sorted"], ref.values.cpu().numpy().view(np.uint32))
    assert np.array_equal(got["point_index"], ref.indices.to(torch.int32).cpu().numpy().view(np.uint32))
    u_want, f_want = oracle.unique_runs(got["sorted"])
    assert np.array_equal(got["unique"], u_want)
    assert np.array_equal(got["leaf_first"], f_want)
