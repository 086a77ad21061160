"""ctypes front end of the CPU oracle (``liboracle.so``).

TEST INFRASTRUCTURE ONLY: the checker that the HIP path is compared against.
Only tests, the smoke check and the benchmark's CPU-baseline leg import this
module.  Every function mirrors one stage of the reference pipeline; the C
source (``tc_oracle.c``) cites the reference shader each one restates.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

OCT_NODE_DTYPE = np.dtype(
    [
        ("children", "<i4", (8,)),
        ("corner", "<f4", (4,)),
        ("cell_size", "<f4"),
        ("child_node_mask", "<i4"),
        ("child_leaf_mask", "<i4"),
        ("pad_", "<i4"),
    ]
)
assert OCT_NODE_DTYPE.itemsize == 64


def build(force: bool = False) -> str:
    """Compile liboracle.so with the committed Makefile."""
    args = ["make", "-C", _HERE] + (["-B"] if force else [])
    subprocess.run(args, check=True, stdout=subprocess.DEVNULL)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        P = C.c_void_p
        sz, u32, i32, f32, u64 = C.c_size_t, C.c_uint32, C.c_int, C.c_float, C.c_uint64
        sig = {
            "tco_quantize": (u32, [f32, f32, f32, f32]),
            "tco_morton30_point": (u32, [f32, f32, f32, f32, f32]),
            "tco_morton63_point": (u64, [f32, f32, f32, f32, f32]),
            "tco_morton30": (None, [P, P, sz, f32, f32]),
            "tco_morton63": (None, [P, P, sz, f32, f32]),
            "tco_morton30_decode": (None, [u32, f32, f32, P]),
            "tco_morton63_decode": (None, [u64, f32, f32, P]),
            "tco_global_histogram_u32": (None, [P, sz, P]),
            "tco_global_histogram_u64": (None, [P, sz, P]),
            "tco_binning_pass_u32": (None, [P, P, P, P, sz, P, i32]),
            "tco_binning_pass_u64": (None, [P, P, P, P, sz, P, i32]),
            "tco_sort_u32": (None, [P, P, sz, P, P]),
            "tco_sort_u64": (None, [P, P, sz, P, P]),
            "tco_unique_u32": (u32, [P, P, sz]),
            "tco_unique_u64": (u32, [P, P, sz]),
            "tco_unique_runs_u32": (u32, [P, P, P, sz]),
            "tco_unique_runs_u64": (u32, [P, P, P, sz]),
            "tco_build_radix_tree_u32": (None, [P, u32, P, P, P, P, P]),
            "tco_build_radix_tree_u64": (None, [P, u32, P, P, P, P, P]),
            "tco_edge_count": (None, [P, P, P, u32, i32]),
            "tco_exclusive_scan_u32": (u32, [P, P, sz]),
            "tco_build_octree_u32": (u32, [P, P, P, P, P, P, P, P, P, u32, f32, f32, i32]),
            "tco_build_octree_u64": (u32, [P, P, P, P, P, P, P, P, P, u32, f32, f32, i32]),
            "tco_octree_node_bound": (u64, [u64, i32]),
            "tco_mt_build_tree_u32": (u32, [P, u32, f32, f32, i32] + [P] * 13 + [u64, i32, P]),
            "tco_mt_build_tree_u64": (u32, [P, u32, f32, f32, i32] + [P] * 13 + [u64, i32, P]),
            "tco_std_sort_ms": (C.c_double, [P, u32, P]),
            "tco_mt_sort_pairs_u64_ms": (C.c_double, [P, P, u32, i32]),
            "tco_std_stable_sort_pairs_u64_ms": (C.c_double, [P, P, u32]),
        }
        for name, (res, argt) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = argt
        _lib = L
    return _lib


def _p(a: np.ndarray | None):
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data


def _pts(points: np.ndarray) -> np.ndarray:
    points = np.ascontiguousarray(points, dtype=np.float32)
    assert points.ndim == 2 and points.shape[1] == 4
    return points


# ---------------------------------------------------------------- Morton
def compute_morton(points: np.ndarray, min_coord: float, range_: float) -> np.ndarray:
    pts = _pts(points)
    out = np.empty(len(pts), np.uint32)
    lib().tco_morton30(_p(pts), _p(out), len(pts), min_coord, range_)
    return out


def compute_morton63(points: np.ndarray, min_coord: float, range_: float) -> np.ndarray:
    pts = _pts(points)
    out = np.empty(len(pts), np.uint64)
    lib().tco_morton63(_p(pts), _p(out), len(pts), min_coord, range_)
    return out


def morton32_to_xyz(code: int, min_coord: float, range_: float) -> np.ndarray:
    out = np.empty(4, np.float32)
    lib().tco_morton30_decode(int(code), min_coord, range_, _p(out))
    return out


def morton64_to_xyz(code: int, min_coord: float, range_: float) -> np.ndarray:
    out = np.empty(4, np.float32)
    lib().tco_morton63_decode(int(code), min_coord, range_, _p(out))
    return out


# ------------------------------------------------------------- Histogram
def histogram(keys: np.ndarray) -> np.ndarray:
    keys = np.ascontiguousarray(keys)
    if keys.dtype == np.uint32:
        out = np.empty(4 * 256, np.uint32)
        lib().tco_global_histogram_u32(_p(keys), len(keys), _p(out))
    elif keys.dtype == np.uint64:
        out = np.empty(8 * 256, np.uint32)
        lib().tco_global_histogram_u64(_p(keys), len(keys), _p(out))
    else:
        raise TypeError(keys.dtype)
    return out


def binning_pass(keys: np.ndarray, values: np.ndarray | None, digit_prefix: np.ndarray, radix_shift: int):
    keys = np.ascontiguousarray(keys)
    dk = np.empty_like(keys)
    dv = None if values is None else np.empty_like(values)
    pre = np.ascontiguousarray(digit_prefix, dtype=np.uint32)
    fn = lib().tco_binning_pass_u32 if keys.dtype == np.uint32 else lib().tco_binning_pass_u64
    fn(_p(keys), _p(dk), _p(values), _p(dv), len(keys), _p(pre), radix_shift)
    return dk, dv


def radix_sort(keys: np.ndarray, values: np.ndarray | None = None):
    k = np.array(keys, copy=True)
    v = None if values is None else np.array(values, dtype=np.uint32, copy=True)
    tk = np.empty_like(k)
    tv = None if v is None else np.empty_like(v)
    fn = lib().tco_sort_u32 if k.dtype == np.uint32 else lib().tco_sort_u64
    fn(_p(k), _p(v), len(k), _p(tk), _p(tv))
    return (k, v) if values is not None else k


# ---------------------------------------------------------------- Unique
def unique(sorted_keys: np.ndarray) -> np.ndarray:
    s = np.ascontiguousarray(sorted_keys)
    out = np.empty_like(s)
    fn = lib().tco_unique_u32 if s.dtype == np.uint32 else lib().tco_unique_u64
    n = fn(_p(s), _p(out), len(s))
    return out[:n].copy()


def unique_runs(sorted_keys: np.ndarray):
    """(unique keys, run starts with the total appended)."""
    s = np.ascontiguousarray(sorted_keys)
    out = np.empty_like(s)
    first = np.empty(len(s) + 1, np.uint32)
    fn = lib().tco_unique_runs_u32 if s.dtype == np.uint32 else lib().tco_unique_runs_u64
    n = fn(_p(s), _p(out), _p(first), len(s))
    return out[:n].copy(), first[:n + 1].copy()


# ------------------------------------------------------------ Radix tree
def build_radix_tree(codes: np.ndarray):
    c = np.ascontiguousarray(codes)
    nk = len(c)
    nb = max(nk - 1, 0)
    prefix_n = np.zeros(nb, np.uint8)
    hll = np.zeros(nb, np.uint8)
    hlr = np.zeros(nb, np.uint8)
    left = np.zeros(nb, np.int32)
    parent = np.zeros(nb, np.int32)
    fn = lib().tco_build_radix_tree_u32 if c.dtype == np.uint32 else lib().tco_build_radix_tree_u64
    fn(_p(c), nk, _p(prefix_n), _p(hll), _p(hlr), _p(left), _p(parent))
    return dict(prefix_n=prefix_n, has_leaf_left=hll, has_leaf_right=hlr, left_child=left, parent=parent)


def edge_count(prefix_n: np.ndarray, parent: np.ndarray, code_bits: int = 30) -> np.ndarray:
    nb = len(prefix_n)
    out = np.empty(nb, np.uint32)
    lib().tco_edge_count(_p(np.ascontiguousarray(prefix_n)), _p(np.ascontiguousarray(parent)), _p(out), nb, code_bits)
    return out


def exclusive_scan(values: np.ndarray) -> np.ndarray:
    v = np.ascontiguousarray(values, dtype=np.uint32)
    out = np.empty(len(v) + 1, np.uint32)
    lib().tco_exclusive_scan_u32(_p(v), _p(out), len(v))
    return out


def build_octree(offsets, edge_counts, codes, tree, min_coord, range_, link_leaves=True):
    codes = np.ascontiguousarray(codes)
    nb = len(edge_counts)
    total = int(offsets[nb]) if nb > 0 else 0
    nodes = np.zeros(max(total, 1), OCT_NODE_DTYPE)
    fn = lib().tco_build_octree_u32 if codes.dtype == np.uint32 else lib().tco_build_octree_u64
    n = fn(
        _p(nodes), _p(np.ascontiguousarray(offsets, dtype=np.uint32)),
        _p(np.ascontiguousarray(edge_counts, dtype=np.uint32)), _p(codes),
        _p(tree["prefix_n"]), _p(tree["parent"]), _p(tree["left_child"]),
        _p(tree["has_leaf_left"]), _p(tree["has_leaf_right"]), nb, min_coord, range_, int(link_leaves),
    )
    return nodes[:n]


def octree_node_bound(n_keys: int, code_bits: int) -> int:
    return int(lib().tco_octree_node_bound(n_keys, code_bits))


# ------------------------------------------------------- Whole pipeline
def build_tree(points: np.ndarray, min_coord: float, range_: float, code_bits: int = 30, link_leaves=True,
               point_index=False):
    """morton -> sort -> unique -> radix tree -> edge count -> scan -> octree,
    the stage order of src/main.rs:60-170.  point_index=True sorts each
    point's index along (stable pairs sort) and keeps the run starts of unique:
    the point -> leaf map (point_index, leaf_first)."""
    codes = compute_morton(points, min_coord, range_) if code_bits == 30 else compute_morton63(points, min_coord, range_)
    extra = {}
    if point_index:
        sorted_codes, perm = radix_sort(codes, np.arange(len(codes), dtype=np.uint32))
        uniq, first = unique_runs(sorted_codes)
        extra = dict(point_index=perm, leaf_first=first)
    else:
        sorted_codes = radix_sort(codes)
        uniq = unique(sorted_codes)
    tree = build_radix_tree(uniq)
    ec = edge_count(tree["prefix_n"], tree["parent"], code_bits)
    offsets = exclusive_scan(ec)
    nodes = build_octree(offsets, ec, uniq, tree, min_coord, range_, link_leaves)
    return dict(codes=codes, sorted=sorted_codes, unique=uniq, edge_count=ec, offsets=offsets, nodes=nodes, **tree,
                **extra)


# ------------------------------------------------- multi-threaded CPU baseline
MT_STAGES = ("morton", "sort", "unique", "radix_tree", "edge_count", "scan", "octree")


def build_tree_mt(points: np.ndarray, min_coord: float, range_: float, code_bits: int = 30, threads: int = 0,
                  link_leaves=True, timings: dict | None = None):
    """The whole pipeline on `threads` host threads (cpu_mt.cpp; 0 = all
    cores): same outputs as build_tree.  `timings` receives per-stage ms."""
    pts = _pts(points)
    n = len(pts)
    threads = threads or (os.cpu_count() or 1)
    kdt = np.uint32 if code_bits == 30 else np.uint64
    codes, srt, tmp, uniq = (np.empty(max(n, 1), kdt) for _ in range(4))
    nu = C.c_uint32(0)
    nb = max(n - 1, 1)
    pn, hll, hlr = (np.zeros(nb, np.uint8) for _ in range(3))
    left, par = np.zeros(nb, np.int32), np.zeros(nb, np.int32)
    ec, offs = np.zeros(nb, np.uint32), np.zeros(nb + 1, np.uint32)
    cap = int(lib().tco_octree_node_bound(max(n, 1), code_bits))
    nodes = np.zeros(max(cap, 1), OCT_NODE_DTYPE)
    st = np.zeros(7, np.float64)
    fn = lib().tco_mt_build_tree_u32 if code_bits == 30 else lib().tco_mt_build_tree_u64
    total = fn(_p(pts), n, min_coord, range_, threads, _p(codes), _p(srt), _p(tmp), _p(uniq), C.byref(nu), _p(pn),
               _p(hll), _p(hlr), _p(left), _p(par), _p(ec), _p(offs), _p(nodes), cap, int(bool(link_leaves)), _p(st))
    if total == 0xFFFFFFFF:
        raise RuntimeError("octree node bound exceeded")
    u = nu.value
    b = max(u - 1, 0)
    if timings is not None:
        timings.update(dict(zip(MT_STAGES, st.tolist())))
    return dict(codes=codes[:n], sorted=srt[:n], unique=uniq[:u], prefix_n=pn[:b], has_leaf_left=hll[:b],
                has_leaf_right=hlr[:b], left_child=left[:b], parent=par[:b], edge_count=ec[:b],
                offsets=offs[:b + 1] if b else np.zeros(1, np.uint32), nodes=nodes[:total].copy())


def std_sort_ms(codes: np.ndarray) -> float:
    c = np.ascontiguousarray(codes, dtype=np.uint32)
    work = np.empty_like(c)
    return float(lib().tco_std_sort_ms(_p(c), len(c), _p(work)))


def sort_pairs_u64_mt(keys: np.ndarray, vals: np.ndarray, threads: int = 0):
    """Stable LSD radix sort of u64 keys with u32 values on `threads` host
    threads (0 = all cores); returns (keys, vals, milliseconds)."""
    k = np.array(keys, dtype=np.uint64, copy=True)
    v = np.array(vals, dtype=np.uint32, copy=True)
    ms = lib().tco_mt_sort_pairs_u64_ms(_p(k), _p(v), len(k), threads or (os.cpu_count() or 1))
    return k, v, float(ms)


def std_stable_sort_pairs_u64(keys: np.ndarray, vals: np.ndarray):
    """std::stable_sort of (key, value) pairs by key; returns (keys, vals,
    milliseconds)."""
    k = np.array(keys, dtype=np.uint64, copy=True)
    v = np.array(vals, dtype=np.uint32, copy=True)
    ms = lib().tco_std_stable_sort_pairs_u64_ms(_p(k), _p(v), len(k))
    return k, v, float(ms)
