"""Host-side tree files and checks.

* ``load_tree(path)`` reads a tc-tree file (written by ``Tree.export`` or
  ``write_tree``) into numpy arrays.
* ``validate_tree(tree)`` runs the library's invariant checker
  (tc_tree_validate_host) on a dict of host arrays, e.g. ``Tree.download()``
  plus the build parameters, or a loaded file.
* ``write_tree(tree, path)`` writes such a dict (tc_tree_write_host).

File layout: 8-byte magic ``TCTREE\\0\\1``, a little-endian u64 header length,
a JSON header (parameters and, per array, its name, dtype, count and byte
offset), then the arrays, each 64-byte aligned.  Octree nodes are 64-byte
``tc_oct_node`` records (``OCT_NODE_DTYPE``).
"""
from __future__ import annotations

import ctypes as C
import json

import numpy as np

from ._lib import OCT_NODE_DTYPE, TcTreeHost, check, lib

MAGIC = b"TCTREE\x00\x01"

_ARRAYS = ("sorted", "unique", "prefix_n", "has_leaf_left", "has_leaf_right", "left_child", "parent", "edge_count",
           "offsets", "nodes", "point_index", "leaf_first")
_DTYPES = {"prefix_n": np.uint8, "has_leaf_left": np.uint8, "has_leaf_right": np.uint8, "left_child": np.int32,
           "parent": np.int32, "edge_count": np.uint32, "offsets": np.uint32, "point_index": np.uint32,
           "leaf_first": np.uint32}


def load_tree(path: str) -> dict:
    """Arrays and parameters of a tc-tree file (numpy arrays, copies)."""
    with open(path, "rb") as f:
        blob = f.read()
    if blob[:8] != MAGIC:
        raise ValueError(f"{path}: not a tc-tree file")
    hlen = int.from_bytes(blob[8:16], "little")
    head = json.loads(blob[16:16 + hlen].decode())
    out = {k: head[k] for k in ("code_bits", "min_coord", "range", "link_leaves", "n_points", "n_unique", "n_nodes")}
    for a in head["arrays"]:
        dt = OCT_NODE_DTYPE if a["dtype"] == "tc_oct_node" else np.dtype(a["dtype"])
        out[a["name"]] = np.frombuffer(blob, dtype=dt, count=a["count"], offset=a["offset"]).copy()
    return out


def _host_view(t: dict):
    """A TcTreeHost over the dict's arrays (kept alive by the returned list)."""
    bits = int(t["code_bits"])
    kdt = np.uint32 if bits == 30 else np.uint64
    keep = []
    v = TcTreeHost()
    v.code_bits = bits
    v.min_coord = float(t["min_coord"])
    v.range = float(t["range"])
    v.link_leaves = int(t.get("link_leaves", 1))
    v.n_unique = len(t["unique"])
    v.n_nodes = len(t["nodes"])
    n_pts = t.get("n_points")
    if n_pts is None:
        for name in ("sorted", "point_index"):
            if t.get(name) is not None:
                n_pts = len(t[name])
                break
    v.n_points = int(n_pts or 0)
    for name in _ARRAYS:
        a = t.get(name)
        if a is None:
            continue
        if name in ("sorted", "unique"):
            a = np.ascontiguousarray(a, dtype=kdt)
        elif name == "nodes":
            a = np.ascontiguousarray(a).view(np.uint8)
        else:
            a = np.ascontiguousarray(a, dtype=_DTYPES[name])
        keep.append(a)
        setattr(v, name, a.ctypes.data if a.size else None)
    return v, keep


def validate_tree(t: dict):
    """(ok, message) from the library's invariant checker."""
    v, keep = _host_view(t)
    buf = C.create_string_buffer(512)
    st = lib().tc_tree_validate_host(C.byref(v), buf, len(buf))
    if st not in (0, -10):
        check(st, "tc_tree_validate_host")
    del keep
    return st == 0, buf.value.decode()


def write_tree(t: dict, path: str):
    v, keep = _host_view(t)
    check(lib().tc_tree_write_host(C.byref(v), str(path).encode()), "tc_tree_write_host")
    del keep
