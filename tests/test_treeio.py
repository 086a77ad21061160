"""Tree validity checker and tree files (host side, no GPU).

The checker is run on oracle trees (valid by the oracle's own pinning) and on
copies with one planted mistake each, which it must name; files written by the
library are read back by the Python loader array for array.
"""
import os

import numpy as np
import pytest


def oracle_tree(oracle, tclib, bits, link_leaves=True, point_index=False, n=20_000, seed=3):
    pts = tclib.gen_gaussian_mixture(n, seed, 8, 1.5)
    t = oracle.build_tree(pts, 0.0, 1024.0, bits, link_leaves=link_leaves, point_index=point_index)
    t.update(code_bits=bits, min_coord=0.0, range=1024.0, link_leaves=int(link_leaves), n_points=n)
    return t


@pytest.mark.parametrize("bits", [30, 63])
@pytest.mark.parametrize("link", [True, False])
@pytest.mark.parametrize("index", [False, True])
def test_oracle_trees_are_valid(oracle, tclib, bits, link, index):
    ok, msg = tclib.validate_tree(oracle_tree(oracle, tclib, bits, link, index))
    assert ok, msg


def test_small_and_degenerate_trees_are_valid(oracle, tclib):
    for n in (0, 1, 2, 3, 9):
        pts = tclib.gen_uniform(n, 40 + n)
        t = oracle.build_tree(pts, 0.0, 1024.0, 30, point_index=True)
        t.update(code_bits=30, min_coord=0.0, range=1024.0, link_leaves=1, n_points=n)
        ok, msg = tclib.validate_tree(t)
        assert ok, (n, msg)
    pts = np.tile(np.array([[5.0, 5.0, 5.0, 1.0], [900.0, 3.0, 70.0, 1.0]], np.float32), (50, 1))
    t = oracle.build_tree(pts, 0.0, 1024.0, 30, point_index=True)
    t.update(code_bits=30, min_coord=0.0, range=1024.0, link_leaves=1, n_points=len(pts))
    assert tclib.validate_tree(t)[0]


def _mut(t, name, fn):
    u = dict(t)
    u[name] = np.array(t[name], copy=True)
    fn(u[name])
    return u


MUTATIONS = {
    "unique order": ("unique", lambda a: a.__setitem__(slice(3, 5), a[3:5][::-1].copy()), "unique"),
    "prefix_n": ("prefix_n", lambda a: a.__setitem__(7, a[7] + 1), "prefix_n"),
    "parent": ("parent", lambda a: a.__setitem__(len(a) // 2, (a[len(a) // 2] + 1) % len(a)), "parent"),
    "left_child": ("left_child", lambda a: a.__setitem__(0, a[0] + 1), ""),
    "leaf flag": ("has_leaf_left", lambda a: a.__setitem__(np.argmax(a == 1), 0), ""),
    "edge_count": ("edge_count", lambda a: a.__setitem__(11, a[11] + 1), "edge_count"),
    "offsets": ("offsets", lambda a: a.__setitem__(20, a[20] + 1), "offsets"),
    "node corner": ("nodes", lambda a: a["corner"].__setitem__((5, 1), a["corner"][5, 1] + 1.0), "corner"),
    "node size": ("nodes", lambda a: a["cell_size"].__setitem__(9, a["cell_size"][9] * 2), "cell size"),
    "child link": ("nodes", lambda a: a["children"].__setitem__(
        (0, int(np.argmax((a["child_node_mask"][0] >> np.arange(8)) & 1))), 0), ""),
    "leaf link": ("nodes", lambda a: a["children"].__setitem__(
        tuple(np.argwhere(((a["child_node_mask"][:, None] >> np.arange(8)) & 1 == 0) & (a["children"] > 0))[0]), 0),
        "leaf"),
    "point index": ("point_index", lambda a: a.__setitem__(1, a[0]), "point_index"),
    "leaf_first": ("leaf_first", lambda a: a.__setitem__(2, a[3]), "leaf_first"),
    "sorted": ("sorted", lambda a: a.__setitem__(100, a[100] + 1), "sorted"),
}


@pytest.mark.parametrize("bits", [30, 63])
@pytest.mark.parametrize("name", sorted(MUTATIONS))
def test_planted_mistakes_are_found(oracle, tclib, bits, name):
    t = oracle_tree(oracle, tclib, bits, point_index=True)
    field, fn, word = MUTATIONS[name]
    ok, msg = tclib.validate_tree(_mut(t, field, fn))
    assert not ok, name
    assert word in msg, (name, msg)


@pytest.mark.parametrize("bits", [30, 63])
def test_file_roundtrip(oracle, tclib, tmp_path, bits):
    t = oracle_tree(oracle, tclib, bits, point_index=True)
    path = os.path.join(tmp_path, "tree.tct")
    tclib.write_tree(t, path)
    back = tclib.load_tree(path)
    for k in ("sorted", "unique", "prefix_n", "has_leaf_left", "has_leaf_right", "left_child", "parent",
              "edge_count", "offsets", "point_index", "leaf_first"):
        assert np.array_equal(back[k], t[k]), k
    assert np.array_equal(back["nodes"].view(np.uint8), t["nodes"].view(np.uint8))
    assert (back["code_bits"], back["n_points"], back["n_unique"], back["n_nodes"]) == \
        (bits, t["n_points"], len(t["unique"]), len(t["nodes"]))
    assert np.float32(back["min_coord"]) == np.float32(0.0) and np.float32(back["range"]) == np.float32(1024.0)
    with open(path, "rb") as f:
        head = f.read(16)
    assert head[:8] == b"TCTREE\x00\x01"
    ok, msg = tclib.validate_tree(back)
    assert ok, msg
