"""Full-size GPU runs of the benchmark workloads.

* 2^26 points, 64-centre Gaussian mixture (sigma 8): the full plain build is
  compared array by array with the oracle (the C oracle builds 2^26 points in
  about a second), and checked against properties that do not depend on the
  oracle: the sorted codes equal torch's sort of the encoder's output, unique
  codes are strictly increasing, offsets end at the node count, every octree
  child link points inside the tree.
* 2^27 63-bit keys with u32 index payloads: the pair sort equals torch's
  stable sort (keys below 2^63 sort the same signed and unsigned).
"""
import numpy as np
import pytest

from gpu_util import dev

pytestmark = pytest.mark.gpu


def test_full_size_gaussian_build_matches_oracle(gpu_ctx, oracle, tclib):
    import torch

    import tc

    n = 1 << 26
    pts = tclib.gen_gaussian_mixture(n, 0xC0FFEE02, 64, 8.0)
    d_pts = dev(pts)
    tree = tc.Tree(gpu_ctx, n, 30)
    try:
        tree.build(d_pts, n, 0.0, 1024.0)
        got = tree.download()
    finally:
        tree.free()

    # Oracle-independent properties.
    codes = torch.empty(n, dtype=torch.int32, device="cuda")
    gpu_ctx.morton30(d_pts, codes, 0.0, 1024.0)
    ref_sorted = torch.sort(codes).values.cpu().numpy().view(np.uint32)  # 30-bit codes: signed order == unsigned
    assert np.array_equal(got["sorted"], ref_sorted)
    u = got["unique"]
    assert np.all(u[1:] > u[:-1])
    assert len(u) == len(np.unique(ref_sorted))
    assert got["offsets"][-1] == got["n_nodes"]
    kids = got["nodes"]["children"]
    node_mask = got["nodes"]["child_node_mask"]
    for slot in range(8):
        is_node = (node_mask >> slot) & 1 == 1
        assert np.all((kids[is_node, slot] > 0) & (kids[is_node, slot] < got["n_nodes"]))

    want = oracle.build_tree(pts, 0.0, 1024.0, 30)
    for k in ("sorted", "unique", "prefix_n", "has_leaf_left", "has_leaf_right", "left_child", "parent",
              "edge_count", "offsets"):
        assert np.array_equal(got[k], want[k]), k
    assert np.array_equal(got["nodes"].view(np.uint8), want["nodes"].view(np.uint8))


def test_full_size_u64_pairs_sort(gpu_ctx, tclib):
    import torch

    n = 1 << 27
    pts = torch.from_numpy(tclib.gen_uniform(n, 0xC0FFEE03, 2097152.0)).cuda()
    keys = torch.empty(n, dtype=torch.int64, device="cuda")
    gpu_ctx.morton63(pts, keys, 0.0, 2097152.0)
    del pts
    vals = torch.arange(n, dtype=torch.int32, device="cuda")
    ref = torch.sort(keys, stable=True)
    k, v = keys.clone(), vals.clone()
    gpu_ctx.sort(k, torch.empty_like(k), v, torch.empty_like(v))
    torch.cuda.synchronize()
    assert torch.equal(k, ref.values)
    assert torch.equal(v.long(), ref.indices)
