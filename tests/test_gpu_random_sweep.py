"""Seeded random sweep of whole builds against the oracle.

Each case draws its size log-uniformly from [1, 2^18] and one of several cloud
shapes -- uniform, tight clusters, points snapped to a coarse grid (many equal
codes), a flat slab, coordinates on cell boundaries -- with 30- or 63-bit
codes, a coordinate range that is or is not a power of two, and leaf links on
or off.  The fused build (tc_build_tree, the library's tuned sort plans) must
reproduce every array of the oracle's per-cloud pipeline (src/main.rs:60-170)
bit for bit.  One tree per code width is reused, so builds also follow larger
ones in the same buffers."""
import numpy as np
import pytest

from gpu_util import dev

pytestmark = pytest.mark.gpu

KEYS = ("sorted", "unique", "prefix_n", "has_leaf_left", "has_leaf_right", "left_child", "parent", "edge_count",
        "offsets")
KINDS = ("uniform", "clusters", "grid", "slab", "boundaries")
CASES = 48
MAX_N = 1 << 18


def make_case(i):
    r = np.random.default_rng(0x5EED0000 + i)
    n = int(np.exp(r.uniform(0.0, np.log(MAX_N))))
    kind = KINDS[i % len(KINDS)]
    bits = 30 if i % 3 else 63
    mn, rg = (0.0, 1024.0) if i % 4 else (-5.0, 37.5)
    link = bool(i % 5)
    if kind == "uniform":
        xyz = r.uniform(mn, mn + rg, (n, 3))
    elif kind == "clusters":
        c = r.uniform(mn, mn + rg, (int(r.integers(1, 9)), 3))
        xyz = c[r.integers(0, len(c), n)] + r.normal(0.0, rg / 512.0, (n, 3))
    elif kind == "grid":
        step = rg / float(r.integers(2, 33))
        xyz = mn + np.floor(r.uniform(0.0, rg, (n, 3)) / step) * step
    elif kind == "slab":
        xyz = r.uniform(mn, mn + rg, (n, 3))
        xyz[:, int(r.integers(0, 3))] = mn + rg * r.uniform()
    else:
        cells = float(1 << (10 if bits == 30 else 21))
        xyz = mn + r.integers(0, int(cells), (n, 3)) * (rg / cells)
    xyz = np.clip(xyz, mn, np.nextafter(np.float32(mn + rg), np.float32(mn)))
    pts = np.ones((n, 4), np.float32)
    pts[:, :3] = xyz
    return pts, bits, mn, rg, link, kind


@pytest.fixture(scope="module")
def trees(gpu_ctx, tclib):
    made = {b: tclib.Tree(gpu_ctx, MAX_N, b) for b in (30, 63)}
    yield made
    for t in made.values():
        t.free()


@pytest.mark.parametrize("case", range(CASES))
def test_random_build_matches_oracle(trees, oracle, case):
    pts, bits, mn, rg, link, kind = make_case(case)
    tree = trees[bits]
    tree.build(dev(pts), len(pts), mn, rg, link_leaves=link)
    got = tree.download()
    want = oracle.build_tree(pts, mn, rg, bits, link_leaves=link)
    where = f"case {case}: n={len(pts)} {kind} bits={bits} min={mn} range={rg} link={link}"
    for k in KEYS:
        g, w = np.asarray(got[k]), np.asarray(want[k])
        assert g.shape == w.shape and np.array_equal(g, w), f"{where}: {k} differs"
    assert np.array_equal(got["nodes"].view(np.uint8).reshape(-1), want["nodes"].view(np.uint8).reshape(-1)), \
        f"{where}: octree nodes differ"
    ok, msg = tree.validate()
    assert ok, f"{where}: {msg}"
