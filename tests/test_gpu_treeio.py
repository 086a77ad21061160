"""Device trees through the invariant checker and the tree file (GPU builds)."""
import os

import numpy as np
import pytest

from gpu_util import dev

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("bits", [30, 63])
@pytest.mark.parametrize("index", [False, True])
@pytest.mark.parametrize("link", [True, False])
def test_gpu_tree_validates_and_exports(gpu_ctx, tclib, tmp_path, bits, index, link):
    import tc

    n = 77_777
    pts = tclib.gen_gaussian_mixture(n, 9, 12, 1.0)
    tree = tc.Tree(gpu_ctx, n, bits, point_index=index)
    try:
        tree.build(dev(pts), n, 0.0, 1024.0, link_leaves=link)
        ok, msg = tree.validate()
        assert ok, msg
        got = tree.download()
        path = os.path.join(tmp_path, "t.tct")
        tree.export(path)
    finally:
        tree.free()
    back = tc.load_tree(path)
    assert back["link_leaves"] == int(link) and back["code_bits"] == bits and back["n_points"] == n
    for k in ("sorted", "unique", "prefix_n", "left_child", "parent", "edge_count", "offsets") + \
            (("point_index", "leaf_first") if index else ()):
        assert np.array_equal(back[k], got[k]), k
    assert np.array_equal(back["nodes"].view(np.uint8), got["nodes"].view(np.uint8))
    assert tc.validate_tree(back)[0]


def test_full_size_tree_validates(gpu_ctx, tclib):
    """The 2^26 Gaussian-mixture build, checked against every invariant."""
    import tc

    n = 1 << 26
    pts = tclib.gen_gaussian_mixture(n, 0xC0FFEE02, 64, 8.0)
    tree = tc.Tree(gpu_ctx, n, 30, point_index=True)
    try:
        tree.build(dev(pts), n, 0.0, 1024.0)
        ok, msg = tree.validate()
    finally:
        tree.free()
    assert ok, msg
