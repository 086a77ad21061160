"""GPU parity: every stage's HIP output is bit-exact against the CPU oracle on
the same seeded inputs, through the C ABI.  Sizes stay where the oracle
finishes in seconds; full-size runs are covered by property checks in
test_gpu_scale.py."""
import os

import numpy as np
import pytest

from conftest import GOLDEN
from gpu_util import dev, empty, host, run_stages

pytestmark = pytest.mark.gpu

STAGE_KEYS = ("codes", "hist", "sorted", "unique", "prefix_n", "has_leaf_left", "has_leaf_right", "left_child",
              "parent", "edge_count", "offsets")


def oracle_all(oracle, pts, mn, rg, bits):
    r = oracle.build_tree(pts, mn, rg, code_bits=bits)
    r["hist"] = oracle.histogram(r["codes"])
    r["nodes"] = r["nodes"].view(np.uint8).reshape(-1)
    return r


def assert_same(got, want, keys=STAGE_KEYS + ("nodes",)):
    for k in keys:
        g, w = np.asarray(got[k]), np.asarray(want[k])
        assert g.shape == w.shape, (k, g.shape, w.shape)
        if not np.array_equal(g, w):
            bad = np.flatnonzero(g != w)
            raise AssertionError(f"{k}: {len(bad)} mismatches, first at {bad[:5]}: got {g[bad[:5]]} want {w[bad[:5]]}")


@pytest.mark.parametrize("name", ["uniform_2048_seed11.npz", "gauss_4096_seed12.npz", "uniform63_1024_seed13.npz"])
def test_golden_fixture_stage_by_stage(gpu_ctx, name):
    g = np.load(os.path.join(GOLDEN, name))
    got = run_stages(gpu_ctx, g["points"], float(g["min_coord"]), float(g["range"]), int(g["code_bits"]))
    assert_same(got, {k: g[k] for k in STAGE_KEYS + ("nodes",)})


def test_golden_pairs_u64(gpu_ctx):
    g = np.load(os.path.join(GOLDEN, "pairs_u64_3000_seed14.npz"))
    n = len(g["keys"])
    k, v = dev(g["keys"]), dev(g["vals"])
    ka, va = empty(n, np.uint64), empty(n, np.uint32)
    gpu_ctx.sort(k, ka, v, va)
    assert np.array_equal(host(k, np.uint64), g["sorted_keys"])
    assert np.array_equal(host(v, np.uint32), g["sorted_vals"])
    h = empty(2048, np.uint32)
    gpu_ctx.global_histogram(dev(g["keys"]), h)
    assert np.array_equal(host(h, np.uint32), g["hist"])


def test_config1_uniform_2p20(gpu_ctx, oracle, tclib):
    """Plain pipeline, single cloud: 2^20 points uniform in [0,1024)^3, every
    array bit-exact (each stage through its own entry point)."""
    pts = tclib.gen_uniform(1 << 20, 0xC0FFEE01)
    got = run_stages(gpu_ctx, pts, 0.0, 1024.0, 30)
    want = oracle_all(oracle, pts, 0.0, 1024.0, 30)
    assert_same(got, want)


def test_gaussian_mixture_heavy_duplicates(gpu_ctx, oracle, tclib):
    pts = tclib.gen_gaussian_mixture(1 << 19, 0xC0FFEE02, 64, 8.0)
    got = run_stages(gpu_ctx, pts, 0.0, 1024.0, 30)
    want = oracle_all(oracle, pts, 0.0, 1024.0, 30)
    assert len(want["unique"]) < len(pts)
    assert_same(got, want)


def test_63bit_tree(gpu_ctx, oracle, tclib):
    pts = tclib.gen_uniform(1 << 16, 77, 2097152.0)
    got = run_stages(gpu_ctx, pts, 0.0, 2097152.0, 63)
    want = oracle_all(oracle, pts, 0.0, 2097152.0, 63)
    assert_same(got, want)


@pytest.mark.parametrize("n", [1, 2, 3, 63, 64, 65, 4095, 4096, 4097, 7679, 7680, 7681, 12345])
def test_edge_sizes(gpu_ctx, oracle, tclib, n):
    pts = tclib.gen_uniform(n, 1000 + n)
    got = run_stages(gpu_ctx, pts, 0.0, 1024.0, 30)
    assert_same(got, oracle_all(oracle, pts, 0.0, 1024.0, 30))


def test_degenerate_distributions(gpu_ctx, oracle):
    cases = {}
    cases["all_same_point"] = np.tile(np.array([[3.5, 700.25, 12.0, 1.0]], np.float32), (5000, 1))
    p = np.zeros((4000, 4), np.float32)
    p[:, 0] = np.arange(4000) % 2  # codes 0 and 1 only (differ in bit 0)
    p[:, 3] = 1
    cases["bit0_only"] = p
    p = np.zeros((3000, 4), np.float32)
    p[:, :3] = np.random.default_rng(5).integers(0, 4, size=(3000, 3))  # one occupied corner cell block
    p[:, 3] = 1
    cases["tiny_cell_block"] = p
    p = np.zeros((2, 4), np.float32)
    p[1, :3] = 1023.5
    cases["two_extremes"] = p
    for name, pts in cases.items():
        got = run_stages(gpu_ctx, pts, 0.0, 1024.0, 30)
        assert_same(got, oracle_all(oracle, pts, 0.0, 1024.0, 30))


def test_out_of_range_inputs_match_oracle(gpu_ctx, oracle):
    """NaN, negative and >= range coordinates quantise by the defined rule on both sides."""
    pts = np.array([[np.nan, -1.0, 1024.0, 1.0], [1e30, -1e30, 0.0, 1.0], [-0.0, 1023.999, 2047.0, 1.0],
                    [5.0, np.inf, -np.inf, 1.0]], np.float32)
    codes = empty(4, np.uint32)
    gpu_ctx.morton30(dev(pts), codes, 0.0, 1024.0)
    assert np.array_equal(host(codes, np.uint32), oracle.compute_morton(pts, 0.0, 1024.0))
    c63 = empty(4, np.uint64)
    gpu_ctx.morton63(dev(pts), c63, 0.0, 1024.0)
    assert np.array_equal(host(c63, np.uint64), oracle.compute_morton63(pts, 0.0, 1024.0))


@pytest.mark.parametrize("min_coord,range_", [(0.0, 2.0), (-3.25, 0.5), (100.0, 2097152.0), (-1e-3, 1.0 / 1024),
                                              (7.0, 4096.0), (0.0, 3.0)])
def test_pow2_ranges_match_division(gpu_ctx, oracle, min_coord, range_):
    """With range a power of two the kernels multiply by scale / range; the
    oracle divides.  Random points, exact grid boundaries, their fp32
    neighbours and out-of-range values must quantise identically."""
    rng = np.random.default_rng(17)
    n = 20_000
    span = np.float32(range_)
    lo = np.float32(min_coord)
    x = (rng.random((n, 3), dtype=np.float32) * np.float32(1.2) - np.float32(0.1)) * span + lo
    g = np.arange(0, 1025, dtype=np.float32) / np.float32(1024) * span + lo  # grid boundaries
    edges = np.concatenate([g, np.nextafter(g, np.float32(-np.inf)), np.nextafter(g, np.float32(np.inf)),
                            np.array([np.nan, np.inf, -np.inf, 1e-45, -1e-45, 1e38], np.float32)])
    e = np.resize(edges, (len(edges) // 3 + 1) * 3).reshape(-1, 3)
    pts = np.concatenate([x, e]).astype(np.float32)
    pts = np.concatenate([pts, np.ones((len(pts), 1), np.float32)], axis=1)
    codes = empty(len(pts), np.uint32)
    gpu_ctx.morton30(dev(pts), codes, min_coord, range_)
    assert np.array_equal(host(codes, np.uint32), oracle.compute_morton(pts, min_coord, range_))
    c63 = empty(len(pts), np.uint64)
    gpu_ctx.morton63(dev(pts), c63, min_coord, range_)
    assert np.array_equal(host(c63, np.uint64), oracle.compute_morton63(pts, min_coord, range_))


def test_nonunit_min_range(gpu_ctx, oracle, tclib):
    pts = tclib.gen_uniform(50_000, 31, 1.0)
    pts[:, :3] = pts[:, :3] * np.float32(37.0) - np.float32(11.0)
    got = run_stages(gpu_ctx, pts, -11.0, 37.0, 30)
    assert_same(got, oracle_all(oracle, pts, -11.0, 37.0, 30))


def test_decode_kernel(gpu_ctx, oracle):
    codes = np.random.default_rng(6).integers(0, 1 << 30, size=10_000, dtype=np.uint64).astype(np.uint32)
    out = empty(4 * len(codes), np.float32)
    gpu_ctx.morton30_decode(dev(codes), out, -3.0, 17.0)
    got = host(out, np.float32).reshape(-1, 4)
    want = np.stack([oracle.morton32_to_xyz(int(c), -3.0, 17.0) for c in codes])
    assert np.array_equal(got, want)


SORT_PLANS = [(256, 8, 8), (512, 8, 8), (512, 12, 8), (512, 16, 8), (512, 20, 8), (512, 24, 8), (512, 28, 8),
              (512, 32, 8), (512, 36, 8), (1024, 16, 8), (1024, 8, 11), (1024, 16, 11)]


@pytest.mark.parametrize("block,kpt,bits", SORT_PLANS)
def test_sort_launch_shapes(gpu_ctx, oracle, tclib, block, kpt, bits):
    """Every binning launch shape and digit plan the tuner may pick gives the
    stable order: u32 / u64 keys with and without payload (plans with an odd
    number of passes copy the result back), and Morton sorts, whose 30- and
    63-bit codes get their own digit widths (3 x 10 and 3 x 11 + 3 x 10 bits
    in the wide plans)."""
    r = np.random.default_rng(block * 100 + kpt + bits)
    n = 300_001
    keys = r.integers(0, 1 << 32, size=n, dtype=np.uint64).astype(np.uint32)
    keys[::7] = keys[3]
    vals = np.arange(n, dtype=np.uint32)
    prev = gpu_ctx.sort_plan()
    gpu_ctx.set_sort_plan(block, kpt, bits)
    try:
        k, v = dev(keys), dev(vals)
        gpu_ctx.sort(k, empty(n, np.uint32), v, empty(n, np.uint32))
        sk, sv = oracle.radix_sort(keys, vals)
        assert np.array_equal(host(k, np.uint32), sk) and np.array_equal(host(v, np.uint32), sv)
        k = dev(keys)
        gpu_ctx.sort(k, empty(n, np.uint32))
        assert np.array_equal(host(k, np.uint32), sk)
        k64 = (keys.astype(np.uint64) << np.uint64(31)) | r.integers(0, 1 << 31, size=n, dtype=np.uint64)
        k, v = dev(k64), dev(vals)
        gpu_ctx.sort(k, empty(n, np.uint64), v, empty(n, np.uint32))
        sk, sv = oracle.radix_sort(k64, vals)
        assert np.array_equal(host(k, np.uint64), sk) and np.array_equal(host(v, np.uint32), sv)
        pts = tclib.gen_gaussian_mixture(n, block + kpt + bits, 8, 3.0)
        d = dev(pts)
        for cb, kdt in ((30, np.uint32), (63, np.uint64)):
            codes = oracle.compute_morton(pts, 0.0, 1024.0) if cb == 30 else oracle.compute_morton63(pts, 0.0, 1024.0)
            want_k, want_v = oracle.radix_sort(codes, vals)
            mk, mv, ka, va = empty(n, kdt), empty(n, np.uint32), empty(n, kdt), empty(n, np.uint32)
            gpu_ctx.morton_sort_pairs(d, mk, mv, ka, va, 0.0, 1024.0)
            assert np.array_equal(host(mk, kdt, n), want_k), cb
            assert np.array_equal(host(mv, np.uint32, n), want_v), cb
            mk = empty(n, kdt)
            gpu_ctx.morton_sort(d, mk, empty(n, kdt), 0.0, 1024.0)
            assert np.array_equal(host(mk, kdt, n), want_k), cb
    finally:
        gpu_ctx.set_sort_plan(*prev)


@pytest.mark.parametrize("n", [1, 2, 1000, 300_001, 1 << 20])
def test_morton_sort_fused_histogram(gpu_ctx, oracle, tclib, n):
    """tc_morton_sort_* (histogram of pass 0 fused into the encoder) == encode, then sort."""
    pts = tclib.gen_gaussian_mixture(n, 700 + n, 16, 4.0)
    pts[::11, :3] = -5.0  # clamped codes, many ties
    want30 = oracle.radix_sort(oracle.compute_morton(pts, 0.0, 1024.0))
    k = empty(n, np.uint32)
    gpu_ctx.morton_sort(dev(pts), k, empty(n, np.uint32), 0.0, 1024.0)
    assert np.array_equal(host(k, np.uint32), want30)
    want63 = oracle.radix_sort(oracle.compute_morton63(pts, 0.0, 1024.0))
    k = empty(n, np.uint64)
    gpu_ctx.morton_sort(dev(pts), k, empty(n, np.uint64), 0.0, 1024.0)
    assert np.array_equal(host(k, np.uint64), want63)


def test_binning_pass_api(gpu_ctx, oracle):
    r = np.random.default_rng(9)
    keys = r.integers(0, 1 << 32, size=100_000, dtype=np.uint64).astype(np.uint32)
    hist = oracle.histogram(keys)
    for shift in (0, 8, 16, 24):
        dst = empty(len(keys), np.uint32)
        gpu_ctx.binning_pass(dev(keys), dst, None, None, dev(hist), shift)
        want, _ = oracle.binning_pass(keys, None, hist[32 * shift:32 * shift + 256], shift)
        assert np.array_equal(host(dst, np.uint32), want)


def test_repeat_runs_identical(gpu_ctx, tclib):
    """Same input K times -> identical outputs (catches look-back and race bugs)."""
    pts = tclib.gen_gaussian_mixture(1 << 18, 4242, 16, 4.0)
    first = run_stages(gpu_ctx, pts, 0.0, 1024.0, 30)
    for _ in range(4):
        again = run_stages(gpu_ctx, pts, 0.0, 1024.0, 30)
        assert_same(again, first)


def test_fused_pipeline_matches_oracle(gpu_ctx, oracle, tclib):
    """tc_build_tree (fused edge-count+scan, device-side counts, no host sync)."""
    import tc

    for n, seed, gauss in ((1 << 20, 0xC0FFEE01, False), (300_000, 55, True), (5, 1, False), (1, 2, False)):
        pts = tclib.gen_gaussian_mixture(n, seed) if gauss else tclib.gen_uniform(n, seed)
        tree = tc.Tree(gpu_ctx, n, 30)
        tree.build(dev(pts), n, 0.0, 1024.0)
        got = tree.download()
        want = oracle_all(oracle, pts, 0.0, 1024.0, 30)
        assert got["n_unique"] == len(want["unique"])
        assert_same(got, want, ("sorted", "unique", "prefix_n", "has_leaf_left", "has_leaf_right", "left_child",
                                "parent", "edge_count", "offsets"))
        assert np.array_equal(got["nodes"].view(np.uint8).reshape(-1), want["nodes"])
        tree.free()


def test_builds_on_two_streams_at_once(oracle, tclib):
    """Two contexts build on two streams at once, four builds each per round.
    The two launches share the CUs, so a workgroup may start before a
    lower-numbered one of its own launch. The sort and scan look-backs then
    compute the stalled predecessor tiles themselves. Every round's trees must
    equal the oracle's."""
    import tc
    import torch

    clouds = [tclib.gen_gaussian_mixture(1 << 22, 91, 32, 6.0), tclib.gen_uniform(1 << 22, 92)]
    wants = [oracle_all(oracle, c, 0.0, 1024.0, 30) for c in clouds]
    ctxs = [tc.Context(0) for _ in clouds]
    streams = [torch.cuda.Stream() for _ in clouds]
    trees = [tc.Tree(x, len(c), 30) for x, c in zip(ctxs, clouds)]
    d_pts = [dev(c) for c in clouds]
    torch.cuda.synchronize()
    try:
        for rnd in range(3):
            for _ in range(4):
                for t, p, s, c in zip(trees, d_pts, streams, clouds):
                    t.build(p, len(c), 0.0, 1024.0, stream=s)
            torch.cuda.synchronize()
            for i, (t, w) in enumerate(zip(trees, wants)):
                got = t.download()
                assert got["n_unique"] == len(w["unique"]), (rnd, i)
                assert_same(got, w, ("sorted", "unique", "prefix_n", "has_leaf_left", "has_leaf_right",
                                     "left_child", "parent", "edge_count", "offsets"))
                assert np.array_equal(got["nodes"].view(np.uint8).reshape(-1), w["nodes"]), (rnd, i)
    finally:
        for t in trees:
            t.free()
        for x in ctxs:
            x.close()


def test_tree_counts_after_empty_and_refill(gpu_ctx, oracle, tclib):
    """The device counters are zeroed only for an empty cloud (unique writes
    the count otherwise): a build of 0 points after a full one reads back as
    empty, and the next non-empty build reads back its own counts."""
    import tc

    tree = tc.Tree(gpu_ctx, 4096, 30)
    try:
        for n, seed in ((4096, 31), (0, 0), (7, 32), (0, 0), (1, 33), (2500, 34)):
            pts = tclib.gen_uniform(max(n, 1), seed)
            tree.build(dev(pts), n, 0.0, 1024.0)
            got = tree.read_counts()
            if n == 0:
                assert (got["n_unique"], got["n_brt"], got["n_nodes"]) == (0, 0, 0)
                continue
            want = oracle_all(oracle, pts[:n], 0.0, 1024.0, 30)
            assert got["n_unique"] == len(want["unique"])
            full = tree.download()
            assert_same(full, want, ("sorted", "unique", "offsets"))
    finally:
        tree.free()


def read_device_counts(tree):
    import ctypes as C

    import torch

    from tc.api import _hip

    torch.cuda.synchronize()
    out = np.zeros(4, np.uint32)
    assert _hip().hipMemcpy(C.c_void_p(out.ctypes.data), C.c_void_p(tree.t.d_counts), C.c_size_t(16), 2) == 0
    return out


def test_device_counts_words_after_shrinking_builds(gpu_ctx, oracle, tclib):
    """tc_tree.d_counts holds n_unique, n_brt and the octree node count on the
    device after every build, also when a small build follows a larger one."""
    import tc

    tree = tc.Tree(gpu_ctx, 300_000, 30)
    try:
        for n, seed, gauss in ((300_000, 41, True), (20_000, 42, False), (1, 43, False), (2, 44, False),
                               (0, 0, False), (5000, 45, True)):
            pts = tclib.gen_gaussian_mixture(max(n, 1), seed) if gauss else tclib.gen_uniform(max(n, 1), seed)
            tree.build(dev(pts), n, 0.0, 1024.0)
            w = read_device_counts(tree)
            if n == 0:
                assert list(w[:3]) == [0, 0, 0]
                continue
            want = oracle.build_tree(pts[:n], 0.0, 1024.0, 30)
            u = len(want["unique"])
            assert list(w[:3]) == [u, max(u - 1, 0), len(want["nodes"])], (n, w)
    finally:
        tree.free()


@pytest.mark.parametrize("bits", [30, 63])
def test_unlinked_octree_matches_oracle(gpu_ctx, oracle, tclib, bits):
    """link_leaves=0 (the reference's default: k_LinkLeafNodes disabled,
    shader/octree.comp:229-236) gives the oracle's unlinked nodes, plain and
    batched."""
    import torch

    import tc

    rg = 1024.0 if bits == 30 else 2097152.0
    clouds = [tclib.gen_gaussian_mixture(200_000, 71, 16, 4.0 * rg / 1024.0, rg), tclib.gen_uniform(70_000, 72, rg),
              tclib.gen_uniform(3, 73, rg)]
    tree = tc.Tree(gpu_ctx, 200_000, bits)
    try:
        wants = []
        for pts in clouds:
            tree.build(dev(pts), len(pts), 0.0, rg, link_leaves=False)
            got = tree.download()
            w_linked = oracle.build_tree(pts, 0.0, rg, bits, link_leaves=True)["nodes"]
            want = oracle.build_tree(pts, 0.0, rg, bits, link_leaves=False)
            wants.append(want)
            assert_same(got, want, ("sorted", "unique", "prefix_n", "left_child", "parent", "edge_count", "offsets"))
            assert np.array_equal(got["nodes"].view(np.uint8).reshape(-1), want["nodes"].view(np.uint8).reshape(-1))
            if len(want["nodes"]) > 1:
                assert not np.array_equal(want["nodes"].view(np.uint8), w_linked.view(np.uint8))
    finally:
        tree.free()
    b = tc.Batch(gpu_ctx, 4, 300_000, bits)
    try:
        b.set(np.array([len(c) for c in clouds], np.uint32), None, np.full(len(clouds), rg, np.float32))
        b.build(dev(np.concatenate(clouds)), link_leaves=False)
        torch.cuda.synchronize()
        for i, (g, w) in enumerate(zip(b.download(), wants)):
            assert np.array_equal(g["nodes"].view(np.uint8).reshape(-1), w["nodes"].view(np.uint8).reshape(-1)), i
    finally:
        b.close()


def test_fused_pipeline_63bit(gpu_ctx, oracle, tclib):
    import tc

    pts = tclib.gen_uniform(100_000, 63, 2097152.0)
    tree = tc.Tree(gpu_ctx, len(pts), 63)
    tree.build(dev(pts), len(pts), 0.0, 2097152.0)
    got = tree.download()
    want = oracle_all(oracle, pts, 0.0, 2097152.0, 63)
    assert_same(got, want, ("sorted", "unique", "prefix_n", "left_child", "parent", "edge_count", "offsets"))
    assert np.array_equal(got["nodes"].view(np.uint8).reshape(-1), want["nodes"])
    tree.free()


def test_host_mirror_functions(gpu_ctx, oracle, tclib):
    """The reference-shaped host functions (src/main.rs:60-170 call sequence)."""
    import tc

    n = 100_000
    pts = tclib.gen_uniform(n, 99)
    codes = tc.compute_morton(pts, n, 0.0, 1024.0, 256, ctx=gpu_ctx)
    assert np.array_equal(codes, oracle.compute_morton(pts, 0.0, 1024.0))
    assert np.array_equal(tc.histogram(n, codes, ctx=gpu_ctx), oracle.histogram(codes))
    s = tc.radix_sort(codes, n, 256, ctx=gpu_ctx)
    assert np.array_equal(s, np.sort(codes))
    u_sort = s.copy()
    nu = tc.unique(u_sort, n, ctx=gpu_ctx)
    uniq = oracle.unique(s)
    assert nu == len(uniq) and np.array_equal(u_sort[:nu], uniq)
    nb = nu - 1
    pn, hll, hlr = np.zeros(nb, np.uint8), np.zeros(nb, np.uint8), np.zeros(nb, np.uint8)
    left, par = np.zeros(nb, np.int32), np.zeros(nb, np.int32)
    tc.build_radix_tree(nb, u_sort[:nu], pn, hll, hlr, left, par, 256, ctx=gpu_ctx)
    t = oracle.build_radix_tree(uniq)
    for k, v in dict(prefix_n=pn, has_leaf_left=hll, has_leaf_right=hlr, left_child=left, parent=par).items():
        assert np.array_equal(v, t[k]), k
    ec = np.zeros(nb, np.uint32)
    tc.edge_count(pn, par, ec, nb, 256, ctx=gpu_ctx)
    assert np.array_equal(ec, oracle.edge_count(pn, par))
    offs = np.zeros(nu, np.uint32)
    tc.partial_sum(ec, 0, nb, offs, ctx=gpu_ctx)
    assert np.array_equal(offs, oracle.exclusive_scan(ec))
    nodes = np.zeros(int(offs[nb]), tc.OCT_NODE_DTYPE)
    tc.build_octree(nodes, offs, ec, u_sort[:nu], pn, par, 0.0, 1024.0, nb, hll, hlr, left, 256, ctx=gpu_ctx)
    want = oracle.build_octree(offs, ec, uniq, t, 0.0, 1024.0)
    assert np.array_equal(nodes.view(np.uint8), want.view(np.uint8))


def test_invalid_arguments_rejected(gpu_ctx):
    import ctypes

    import tc

    L = tc.lib()
    h = gpu_ctx.handle
    assert L.tc_sort_keys_u32(h, None, None, 10, None) == tc._lib.TC_ERR_INVALID
    assert L.tc_sort_keys_u32(h, None, None, (1 << 30) + 1, None) in (tc._lib.TC_ERR_INVALID, tc._lib.TC_ERR_TOO_LARGE)
    assert L.tc_set_sort_params(h, 128, 16) == tc._lib.TC_ERR_INVALID
    assert L.tc_morton30(h, ctypes.c_void_p(8), ctypes.c_void_p(16), 4, 0.0, 1.0, None) == tc._lib.TC_ERR_INVALID
    k = empty(16, np.uint32)
    assert L.tc_sort_keys_u32(h, k.data_ptr(), k.data_ptr(), (1 << 30) + 1, None) == tc._lib.TC_ERR_TOO_LARGE
    # Scans and unique read their input again when a look-back computes a
    # stalled predecessor tile: outputs overlapping the input are refused.
    buf, cnt = empty(64, np.uint32), empty(1, np.uint32)
    p = buf.data_ptr()
    assert L.tc_exclusive_scan_u32(h, p, p, 16, None, None) == tc._lib.TC_ERR_INVALID
    assert L.tc_exclusive_scan_u32(h, p, p + 4 * 16, 16, None, None) == 0  # adjacent, not overlapping
    assert L.tc_unique_u32(h, p, p + 4 * 8, 16, cnt.data_ptr(), None) == tc._lib.TC_ERR_INVALID
    assert L.tc_unique_runs_u32(h, p, p + 4 * 32, p + 4 * 15, 16, cnt.data_ptr(), None) == tc._lib.TC_ERR_INVALID
    assert L.tc_sort_keys_u32(h, p, p + 4 * 8, 16, None) == tc._lib.TC_ERR_INVALID
    import torch

    torch.cuda.synchronize()


@pytest.mark.parametrize("limit", [0, 1])
def test_lookback_fallback_path(gpu_ctx, oracle, tclib, limit):
    """With the look-back wait bound at 0 every workgroup computes each
    predecessor tile's counts from the input itself; at 1 most predecessors
    are computed and the rest read, so the walk resumes after each computed
    tile.  Results must be identical to the look-back path's."""
    import tc

    L = tc.lib()
    pts = tclib.gen_gaussian_mixture(300_000, 77, 32, 6.0)
    want = oracle_all(oracle, pts, 0.0, 1024.0, 30)
    assert L.tc_set_lookback_spin_limit(gpu_ctx.handle, limit) == 0
    try:
        got = run_stages(gpu_ctx, pts, 0.0, 1024.0, 30)
        assert_same(got, want)
        r = np.random.default_rng(5)
        k64 = r.integers(0, 1 << 63, size=70_000, dtype=np.uint64)
        k64[::5] = k64[1]
        vals = np.arange(len(k64), dtype=np.uint32)
        k, v = dev(k64), dev(vals)
        gpu_ctx.sort(k, empty(len(k64), np.uint64), v, empty(len(k64), np.uint32))
        sk, sv = oracle.radix_sort(k64, vals)
        assert np.array_equal(host(k, np.uint64), sk) and np.array_equal(host(v, np.uint32), sv)
    finally:
        assert L.tc_set_lookback_spin_limit(gpu_ctx.handle, tc._lib.LOOKBACK_SPIN_DEFAULT) == 0


@pytest.mark.parametrize("limit", [0, 1])
@pytest.mark.parametrize("block,kpt,bits", [(1024, 16, 11), (1024, 8, 11), (256, 8, 8)])
def test_lookback_fallback_sort_plans(gpu_ctx, oracle, limit, block, kpt, bits):
    """Sorts whose look-backs compute predecessor tiles themselves (every one
    at limit 0, most at 1) under other plans than the tuned one: 11-bit digits
    with two digits per look-back thread (1024-thread blocks, 2,048 digits),
    and small 256 x 8 tiles (98 tiles, so long walks).  u32 pairs, u64 keys
    and u64 pairs against the oracle."""
    import tc

    L = tc.lib()
    n = 200_003
    r = np.random.default_rng(limit * 31 + kpt + bits)
    keys = r.integers(0, 1 << 32, size=n, dtype=np.uint64).astype(np.uint32)
    keys[::7] = keys[3]
    k64 = r.integers(0, 1 << 64, size=n, dtype=np.uint64)
    vals = np.arange(n, dtype=np.uint32)
    prev = gpu_ctx.sort_plan()
    gpu_ctx.set_sort_plan(block, kpt, bits)
    assert L.tc_set_lookback_spin_limit(gpu_ctx.handle, limit) == 0
    try:
        sk, sv = oracle.radix_sort(keys, vals)
        k, v = dev(keys), dev(vals)
        gpu_ctx.sort(k, empty(n, np.uint32), v, empty(n, np.uint32))
        assert np.array_equal(host(k, np.uint32, n), sk) and np.array_equal(host(v, np.uint32, n), sv)
        k = dev(k64)
        gpu_ctx.sort(k, empty(n, np.uint64))
        assert np.array_equal(host(k, np.uint64, n), np.sort(k64))
        sk, sv = oracle.radix_sort(k64, vals)
        k, v = dev(k64), dev(vals)
        gpu_ctx.sort(k, empty(n, np.uint64), v, empty(n, np.uint32))
        assert np.array_equal(host(k, np.uint64, n), sk) and np.array_equal(host(v, np.uint32, n), sv)
    finally:
        assert L.tc_set_lookback_spin_limit(gpu_ctx.handle, tc._lib.LOOKBACK_SPIN_DEFAULT) == 0
        gpu_ctx.set_sort_plan(*prev)


def test_reference_sort_input_10M_descending(gpu_ctx, oracle):
    """The reference's own sort input, keys 10^7 ... 1 (src/main.rs:189-190):
    global histogram against the oracle and its closed form (digit 0, bin 1 =
    39,062; bin 255 = 9,960,938; digit 3, bins >= 1 = 10^7), each binning pass
    against the oracle's, and the full sort against 1 ... 10^7."""
    n = 10_000_000
    keys = (n - np.arange(n, dtype=np.int64)).astype(np.uint32)
    hist = empty(1024, np.uint32)
    gpu_ctx.global_histogram(dev(keys), hist)
    h = host(hist, np.uint32)
    assert np.array_equal(h, oracle.histogram(keys))
    assert h[1] == 39_062 and h[255] == 9_960_938 and np.all(h[768 + 1:] == n)
    cur = keys
    for p in range(4):
        dst = empty(n, np.uint32)
        gpu_ctx.binning_pass(dev(cur), dst, None, None, dev(h), 8 * p)
        want, _ = oracle.binning_pass(cur, None, h[256 * p:256 * p + 256], 8 * p)
        got = host(dst, np.uint32, n)
        assert np.array_equal(got, want), p
        cur = got
    k = dev(keys)
    gpu_ctx.sort(k, empty(n, np.uint32))
    assert np.array_equal(host(k, np.uint32, n), np.arange(1, n + 1, dtype=np.uint32))


def test_markers_do_not_change_results(gpu_ctx, oracle, tclib):
    """roctx ranges on (the ROCm marker library is part of the image): the
    build is unchanged."""
    import tc

    pts = tclib.gen_gaussian_mixture(60_000, 77, 8, 3.0)
    gpu_ctx.markers(True)
    try:
        tree = tc.Tree(gpu_ctx, len(pts), 30)
        tree.build(dev(pts), len(pts), 0.0, 1024.0)
        got = tree.download()
        tree.free()
    finally:
        gpu_ctx.markers(False)
    want = oracle.build_tree(pts, 0.0, 1024.0, 30)
    for k in STAGE_KEYS:
        if k in want and k in got:
            assert np.array_equal(got[k], want[k]), k
    assert np.array_equal(got["nodes"].view(np.uint8), want["nodes"].view(np.uint8))


@pytest.mark.parametrize("n", [1, 2, 8191, 8192, 8193, 16383, 16384, 16385, 32767, 32768, 32769, 3 * 16384 + 5,
                               300_001])
def test_scan_modes_at_tile_edges(gpu_ctx, oracle, n):
    """Unique (u32, u64), exclusive scan and the fused edge-count scan at sizes
    around multiples of the scan kernels' 8192-element tile and of unique's
    16384-key (u32) and 8192-key (u64) tiles, against the oracle."""
    check_scan_modes(gpu_ctx, oracle, n)


@pytest.mark.parametrize("limit", [0, 1])
@pytest.mark.parametrize("n", [16385, 3 * 16384 + 5, 300_001])
def test_scan_modes_fallback(gpu_ctx, oracle, n, limit):
    """The same scan modes with predecessor tiles reduced by the waiting
    workgroup (every one at look-back limit 0, most at 1), over 2 to 37 tiles."""
    import tc

    L = tc.lib()
    assert L.tc_set_lookback_spin_limit(gpu_ctx.handle, limit) == 0
    try:
        check_scan_modes(gpu_ctx, oracle, n)
    finally:
        assert L.tc_set_lookback_spin_limit(gpu_ctx.handle, tc._lib.LOOKBACK_SPIN_DEFAULT) == 0


def check_scan_modes(gpu_ctx, oracle, n):
    r = np.random.default_rng(n + 17)
    for dt in (np.uint32, np.uint64):
        keys = np.sort(r.integers(0, max(n // 3, 2), size=n, dtype=np.uint64).astype(dt))
        out, cnt = empty(n, dt), empty(1, np.uint32)
        gpu_ctx.unique(dev(keys), out, cnt)
        want = oracle.unique(keys)
        got_n = int(host(cnt, np.uint32)[0])
        assert got_n == len(want) and np.array_equal(host(out, dt)[:got_n], want)
    vals = r.integers(0, 1000, size=n, dtype=np.uint64).astype(np.uint32)
    out = empty(n + 1, np.uint32)
    gpu_ctx.exclusive_scan(dev(vals), out)
    assert np.array_equal(host(out, np.uint32), oracle.exclusive_scan(vals))
    if n >= 2:
        codes = np.unique(r.integers(0, 1 << 30, size=n + 64, dtype=np.uint64).astype(np.uint32))[:n]
        t = oracle.build_radix_tree(codes)
        nb = len(codes) - 1
        if nb >= 1:
            ec, offs = empty(nb, np.uint32), empty(nb + 1, np.uint32)
            gpu_ctx.edge_count_scan(dev(t["prefix_n"]), dev(t["parent"]), ec, offs, nb)
            want_ec = oracle.edge_count(t["prefix_n"], t["parent"])
            assert np.array_equal(host(ec, np.uint32), want_ec)
            assert np.array_equal(host(offs, np.uint32), oracle.exclusive_scan(want_ec))


@pytest.mark.parametrize("dt", [np.uint32, np.uint64])
@pytest.mark.parametrize("pattern", ["one_run", "distinct", "pairs", "runs_across_tiles"])
def test_unique_key_patterns_and_unaligned_input(gpu_ctx, oracle, dt, pattern):
    """Unique over 5 tiles and a bit of keys in patterns that put run starts at
    every slot, at none, and across tile and row edges; the same keys starting
    one element into the buffer (not 16-byte aligned, so every tile takes the
    per-key load path) must give the same result."""
    n = 5 * 16384 + 77
    r = np.random.default_rng(31)
    if pattern == "one_run":
        keys = np.full(n, 12345, dt)
    elif pattern == "distinct":
        keys = np.sort(r.choice(np.arange(1 << 28, dtype=np.uint64), size=n, replace=False)).astype(dt)
    elif pattern == "pairs":
        keys = (np.arange(n, dtype=np.uint64) // 2).astype(dt)
    else:  # runs of 3000-20000 keys, longer than a row and than a tile
        lens = r.integers(3000, 20_000, size=n // 3000 + 2)
        keys = np.repeat(np.arange(len(lens), dtype=np.uint64) * 7, lens)[:n].astype(dt)
    want, want_first = oracle.unique_runs(keys)
    for offset in (0, 1):
        buf = dev(np.concatenate([np.zeros(offset, dt), keys]))
        src = buf[offset:]
        out, first, cnt = empty(n, dt), empty(n + 1, np.uint32), empty(1, np.uint32)
        gpu_ctx.unique(src, out, cnt, n=n)
        u = int(host(cnt, np.uint32)[0])
        assert u == len(want) and np.array_equal(host(out, dt, u), want), offset
        cnt.zero_()
        gpu_ctx.unique_runs(src, out, first, cnt, n=n)
        assert int(host(cnt, np.uint32)[0]) == len(want)
        assert np.array_equal(host(out, dt, u), want) and np.array_equal(host(first, np.uint32, u + 1), want_first)


@pytest.mark.parametrize("block,kpt", [(1024, 8), (1024, 16)])
@pytest.mark.parametrize("n", [1, 2, 3, 8191, 8192, 8193, 16384, 16385, 49153, 1 << 20])
def test_wide_digit_plans_at_tile_edges(gpu_ctx, oracle, tclib, block, kpt, n):
    """The wide digit plans (three passes of <= 11 bits for u32 keys, six for
    u64 keys) at sizes around their tile edges, against the oracle: u32 and u64
    key sorts with and without payload (odd pass counts copy back), and the
    30- and 63-bit Morton sorts (3 x 10 and 3 x 11 + 3 x 10 bits) with the
    point index, and a whole tree build under the wide plan."""
    import tc

    r = np.random.default_rng(n * 7 + kpt)
    keys = r.integers(0, 1 << 32, size=n, dtype=np.uint64).astype(np.uint32)
    keys[::5] = keys[0]
    vals = np.arange(n, dtype=np.uint32)
    k64 = r.integers(0, 1 << 64, size=n, dtype=np.uint64)
    prev = gpu_ctx.sort_plan()
    gpu_ctx.set_sort_plan(block, kpt, 11)
    try:
        sk, sv = oracle.radix_sort(keys, vals)
        k, v = dev(keys), dev(vals)
        gpu_ctx.sort(k, empty(n, np.uint32), v, empty(n, np.uint32))
        assert np.array_equal(host(k, np.uint32, n), sk) and np.array_equal(host(v, np.uint32, n), sv)
        k = dev(k64)
        gpu_ctx.sort(k, empty(n, np.uint64))
        assert np.array_equal(host(k, np.uint64, n), np.sort(k64))
        k, v = dev(k64), dev(vals)
        gpu_ctx.sort(k, empty(n, np.uint64), v, empty(n, np.uint32))
        sk, sv = oracle.radix_sort(k64, vals)
        assert np.array_equal(host(k, np.uint64, n), sk) and np.array_equal(host(v, np.uint32, n), sv)
        pts = tclib.gen_gaussian_mixture(n, n + 3, 8, 3.0)
        d = dev(pts)
        for cb, kdt in ((30, np.uint32), (63, np.uint64)):
            codes = oracle.compute_morton(pts, 0.0, 1024.0) if cb == 30 else oracle.compute_morton63(pts, 0.0, 1024.0)
            want_k, want_v = oracle.radix_sort(codes, vals)
            mk, mv = empty(n, kdt), empty(n, np.uint32)
            gpu_ctx.morton_sort_pairs(d, mk, mv, empty(n, kdt), empty(n, np.uint32), 0.0, 1024.0)
            assert np.array_equal(host(mk, kdt, n), want_k), cb
            assert np.array_equal(host(mv, np.uint32, n), want_v), cb
        if n >= 2:
            tree = tc.Tree(gpu_ctx, n, 30)
            tree.build(d, n, 0.0, 1024.0)
            got = tree.download()
            tree.free()
            want = oracle_all(oracle, pts, 0.0, 1024.0, 30)
            assert_same(got, want, ("sorted", "unique", "prefix_n", "left_child", "parent", "edge_count", "offsets"))
            assert np.array_equal(got["nodes"].view(np.uint8).reshape(-1), want["nodes"])
    finally:
        gpu_ctx.set_sort_plan(*prev)


def test_traffic_reference_modes(gpu_ctx, tclib):
    """The measurement aid's kernels: the read modes fold every word into the
    sink, the write mode fills, the copy mode copies the first half onto the
    second; bad sizes are rejected."""
    import ctypes as C

    import torch

    L = tclib.lib()
    words = 1 << 20
    buf = torch.arange(words, dtype=torch.int32, device="cuda")
    sink = torch.zeros(2048, dtype=torch.int32, device="cuda")
    ptr, nbytes = C.c_void_p(buf.data_ptr()), words * 4
    for mode in (0, 1):
        sink.zero_()
        assert L.tc_traffic_reference(gpu_ctx.handle, ptr, nbytes, mode, C.c_void_p(sink.data_ptr()), None) == 0
        folded = np.bitwise_xor.reduce(sink.cpu().numpy().astype(np.uint32))
        assert folded == np.bitwise_xor.reduce(np.arange(words, dtype=np.uint32))
    assert L.tc_traffic_reference(gpu_ctx.handle, ptr, nbytes, 3, None, None) == 0
    h = buf.cpu().numpy()
    assert np.array_equal(h[words // 2:], np.arange(words // 2, dtype=np.int32))
    assert L.tc_traffic_reference(gpu_ctx.handle, ptr, nbytes, 2, None, None) == 0
    assert np.all(buf.cpu().numpy().view(np.uint32) == 0x5A5A5A5A)
    assert L.tc_traffic_reference(gpu_ctx.handle, ptr, nbytes - 16, 3, None, None) != 0  # copy needs 32-byte units
    assert L.tc_traffic_reference(gpu_ctx.handle, ptr, nbytes, 4, None, None) != 0
