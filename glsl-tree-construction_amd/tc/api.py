"""Python host side of libtc.

Two layers, both calling the C ABI (include/tc.h):

* ``Context`` -- device-level stage calls on torch CUDA (HIP) tensors, queued
  on the current torch stream; the building block of the pipeline, batching and
  serving code.
* module-level functions with the reference's names (``compute_morton``,
  ``histogram``, ``radix_sort``, ``unique``, ``build_radix_tree``,
  ``edge_count``, ``partial_sum``, ``build_octree``, ``morton32_to_xyz``) --
  numpy-in / numpy-out mirrors of the reference's per-stage Rust functions
  (src/morton.rs:76, src/histogram.rs:38, src/radixsort.rs:36,
  src/unique.rs:37, src/build_radix_tree.rs:34, src/edge_count.rs:41,
  src/partial_sum.rs:43, src/octree.rs:73, src/morton.rs:43).  Where the
  reference panics (``.unwrap()``) these raise ``TcError``.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import OCT_NODE_DTYPE, TREE_POINT_INDEX, TcError, TcTree, check, lib

_default_ctx = None


def _np_ptr(a: np.ndarray | None):
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"], "arrays must be C-contiguous"
    return a.ctypes.data


def _t_ptr(t):
    if t is None:
        return None
    assert t.is_cuda and t.is_contiguous(), "device tensors must be contiguous CUDA tensors"
    return t.data_ptr()


def _stream(stream):
    import torch

    if stream is None:
        stream = torch.cuda.current_stream()
    return stream.cuda_stream


class Context:
    """Owns a tc_ctx (device scratch); externally synchronised, one stream at a time."""

    def __init__(self, device: int = 0):
        self._h = C.c_void_p()
        check(lib().tc_create(C.byref(self._h), device), "tc_create")
        self.device = device

    @property
    def handle(self):
        return self._h

    def close(self):
        if self._h:
            lib().tc_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def reserve(self, max_n: int):
        check(lib().tc_reserve(self._h, max_n), "tc_reserve")

    def set_sort_params(self, block: int, kpt: int):
        check(lib().tc_set_sort_params(self._h, block, kpt), "tc_set_sort_params")

    def set_sort_plan(self, block: int, kpt: int, digit_bits: int = 8):
        """Fixed sort launch shape and digit plan (8: 8-bit digits; 11: digits
        of at most 11 bits, fewer passes)."""
        check(lib().tc_set_sort_plan(self._h, block, kpt, digit_bits), "tc_set_sort_plan")

    def sort_plan(self):
        b, k, d = C.c_int(), C.c_int(), C.c_int()
        check(lib().tc_get_sort_plan(self._h, C.byref(b), C.byref(k), C.byref(d)), "tc_get_sort_plan")
        return b.value, k.value, d.value

    def profile(self, on: bool = True):
        check(lib().tc_profile_enable(self._h, int(bool(on))), "tc_profile_enable")

    def profile_reset(self):
        check(lib().tc_profile_reset(self._h), "tc_profile_reset")

    def profile_report(self) -> dict:
        """{kernel name: (total ms, launches)} of the launches profiled so far."""
        L = lib()
        check(L.tc_profile_collect(self._h), "tc_profile_collect")
        out = {}
        for kid in range(13):
            ms, cnt = C.c_double(), C.c_uint64()
            check(L.tc_profile_get(self._h, kid, C.byref(ms), C.byref(cnt)), "tc_profile_get")
            if cnt.value:
                out[L.tc_kernel_name(kid).decode()] = (ms.value, cnt.value)
        return out

    def markers(self, on: bool = True):
        """roctx ranges around every launch and build call (tc_set_markers)."""
        check(lib().tc_set_markers(self._h, int(bool(on))), "tc_set_markers")

    def sort_params(self):
        b, k = C.c_int(), C.c_int()
        check(lib().tc_get_sort_params(self._h, C.byref(b), C.byref(k)), "tc_get_sort_params")
        return b.value, k.value

    # ---- stages on device tensors ------------------------------------
    def morton30(self, pts, codes, min_coord, range_, stream=None):
        check(lib().tc_morton30(self._h, _t_ptr(pts), _t_ptr(codes), pts.shape[0], min_coord, range_,
                                _stream(stream)), "tc_morton30")

    def morton63(self, pts, codes, min_coord, range_, stream=None):
        check(lib().tc_morton63(self._h, _t_ptr(pts), _t_ptr(codes), pts.shape[0], min_coord, range_,
                                _stream(stream)), "tc_morton63")

    def morton_sort(self, pts, keys, alt, min_coord, range_, n=None, stream=None):
        """Morton codes of ``pts`` (30-bit into int32 ``keys``, 63-bit into int64) sorted ascending;
        the histogram of the first sort pass is fused into the encoding kernel."""
        n = pts.shape[0] if n is None else n
        fn = lib().tc_morton_sort_u64 if keys.element_size() == 8 else lib().tc_morton_sort_u32
        check(fn(self._h, _t_ptr(pts), _t_ptr(keys), _t_ptr(alt), n, min_coord, range_, _stream(stream)),
              "tc_morton_sort")

    def morton30_decode(self, codes, corners, min_coord, range_, stream=None):
        check(lib().tc_morton30_decode(self._h, _t_ptr(codes), _t_ptr(corners), codes.shape[0], min_coord,
                                       range_, _stream(stream)), "tc_morton30_decode")

    def global_histogram(self, keys, hist, n=None, stream=None):
        n = keys.shape[0] if n is None else n
        fn = lib().tc_global_histogram_u64 if keys.element_size() == 8 else lib().tc_global_histogram_u32
        check(fn(self._h, _t_ptr(keys), n, _t_ptr(hist), _stream(stream)), "tc_global_histogram")

    def binning_pass(self, src_k, dst_k, src_v, dst_v, hist, radix_shift, n=None, stream=None):
        n = src_k.shape[0] if n is None else n
        fn = lib().tc_binning_pass_u64 if src_k.element_size() == 8 else lib().tc_binning_pass_u32
        check(fn(self._h, _t_ptr(src_k), _t_ptr(dst_k), _t_ptr(src_v), _t_ptr(dst_v), n, _t_ptr(hist),
                 radix_shift, _stream(stream)), "tc_binning_pass")

    def sort(self, keys, alt, vals=None, vals_alt=None, n=None, stream=None):
        n = keys.shape[0] if n is None else n
        L = lib()
        if keys.element_size() == 8:
            if vals is None:
                st = L.tc_sort_keys_u64(self._h, _t_ptr(keys), _t_ptr(alt), n, _stream(stream))
            else:
                st = L.tc_sort_pairs_u64(self._h, _t_ptr(keys), _t_ptr(vals), _t_ptr(alt), _t_ptr(vals_alt), n,
                                         _stream(stream))
        else:
            if vals is None:
                st = L.tc_sort_keys_u32(self._h, _t_ptr(keys), _t_ptr(alt), n, _stream(stream))
            else:
                st = L.tc_sort_pairs_u32(self._h, _t_ptr(keys), _t_ptr(vals), _t_ptr(alt), _t_ptr(vals_alt), n,
                                         _stream(stream))
        check(st, "tc_sort")

    def morton_sort_pairs(self, pts, keys, index, keys_alt, index_alt, min_coord, range_, n=None, stream=None):
        """As morton_sort, with each point's input index sorted along (int32 ``index``)."""
        n = pts.shape[0] if n is None else n
        fn = lib().tc_morton_sort_pairs_u64 if keys.element_size() == 8 else lib().tc_morton_sort_pairs_u32
        check(fn(self._h, _t_ptr(pts), _t_ptr(keys), _t_ptr(index), _t_ptr(keys_alt), _t_ptr(index_alt), n,
                 min_coord, range_, _stream(stream)), "tc_morton_sort_pairs")

    def unique(self, sorted_keys, out, d_count, n=None, stream=None):
        n = sorted_keys.shape[0] if n is None else n
        fn = lib().tc_unique_u64 if sorted_keys.element_size() == 8 else lib().tc_unique_u32
        check(fn(self._h, _t_ptr(sorted_keys), _t_ptr(out), n, _t_ptr(d_count), _stream(stream)), "tc_unique")

    def unique_runs(self, sorted_keys, out, run_first, d_count, n=None, stream=None):
        """unique plus run starts: run_first[j] = position of run j's first key, run_first[U] = n."""
        n = sorted_keys.shape[0] if n is None else n
        fn = lib().tc_unique_runs_u64 if sorted_keys.element_size() == 8 else lib().tc_unique_runs_u32
        check(fn(self._h, _t_ptr(sorted_keys), _t_ptr(out), _t_ptr(run_first), n, _t_ptr(d_count),
                 _stream(stream)), "tc_unique_runs")

    def build_radix_tree(self, codes, n_keys, prefix_n, hll, hlr, left, parent, d_n_keys=None, stream=None):
        fn = lib().tc_build_radix_tree_u64 if codes.element_size() == 8 else lib().tc_build_radix_tree_u32
        check(fn(self._h, _t_ptr(codes), n_keys, _t_ptr(d_n_keys), _t_ptr(prefix_n), _t_ptr(hll), _t_ptr(hlr),
                 _t_ptr(left), _t_ptr(parent), _stream(stream)), "tc_build_radix_tree")

    def edge_count(self, prefix_n, parent, ec, n_brt, d_n_keys=None, stream=None):
        check(lib().tc_edge_count(self._h, _t_ptr(prefix_n), _t_ptr(parent), _t_ptr(ec), n_brt, _t_ptr(d_n_keys),
                                  _stream(stream)), "tc_edge_count")

    def exclusive_scan(self, values, out, n=None, d_n_keys=None, stream=None):
        n = values.shape[0] if n is None else n
        check(lib().tc_exclusive_scan_u32(self._h, _t_ptr(values), _t_ptr(out), n, _t_ptr(d_n_keys),
                                          _stream(stream)), "tc_exclusive_scan_u32")

    def edge_count_scan(self, prefix_n, parent, ec, offsets, n_brt, d_n_keys=None, stream=None):
        check(lib().tc_edge_count_scan(self._h, _t_ptr(prefix_n), _t_ptr(parent), _t_ptr(ec), _t_ptr(offsets),
                                       n_brt, _t_ptr(d_n_keys), _stream(stream)), "tc_edge_count_scan")

    def build_octree(self, nodes, node_cap, offsets, ec, codes, prefix_n, parent, left, hll, hlr, n_brt,
                     min_coord, range_, link_leaves=True, d_n_keys=None, stream=None):
        fn = lib().tc_build_octree_u64 if codes.element_size() == 8 else lib().tc_build_octree_u32
        check(fn(self._h, _t_ptr(nodes), node_cap, _t_ptr(offsets), _t_ptr(ec), _t_ptr(codes), _t_ptr(prefix_n),
                 _t_ptr(parent), _t_ptr(left), _t_ptr(hll), _t_ptr(hlr), n_brt, _t_ptr(d_n_keys), min_coord,
                 range_, int(bool(link_leaves)), _stream(stream)), "tc_build_octree")


class Tree:
    """Device-resident buffers of one plain build (tc_tree)."""

    def __init__(self, ctx: Context, capacity: int, code_bits: int = 30, point_index: bool = False):
        """point_index=True also builds the point -> leaf map on every build:
        ``point_index`` (sorted position -> input point) and ``leaf_first``
        (leaf j holds point_index[leaf_first[j]:leaf_first[j+1]])."""
        self.ctx = ctx
        self.t = TcTree()
        flags = TREE_POINT_INDEX if point_index else 0
        check(lib().tc_tree_alloc_ex(ctx.handle, capacity, code_bits, flags, C.byref(self.t)), "tc_tree_alloc_ex")

    def free(self):
        if self.t.capacity or self.t.sorted:
            lib().tc_tree_free(self.ctx.handle, C.byref(self.t))

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    def build(self, pts, n, min_coord, range_, link_leaves=True, stream=None):
        check(lib().tc_build_tree(self.ctx.handle, _t_ptr(pts), n, min_coord, range_, int(bool(link_leaves)),
                                  C.byref(self.t), _stream(stream)), "tc_build_tree")

    def validate(self, stream=None):
        """Checks the last build's invariants (tc_tree_validate); returns (ok, message)."""
        buf = C.create_string_buffer(512)
        st = lib().tc_tree_validate(self.ctx.handle, C.byref(self.t), buf, len(buf), _stream(stream))
        if st not in (0, -10):
            check(st, "tc_tree_validate")
        return st == 0, buf.value.decode()

    def export(self, path: str, stream=None):
        """Writes the last build to a tc-tree file (tc_tree_export); read it with tc.load_tree."""
        check(lib().tc_tree_export(self.ctx.handle, C.byref(self.t), str(path).encode(), _stream(stream)),
              "tc_tree_export")

    def read_counts(self, stream=None):
        check(lib().tc_tree_read_counts(self.ctx.handle, C.byref(self.t), _stream(stream)), "tc_tree_read_counts")
        return dict(n_points=self.t.n_points, n_unique=self.t.n_unique, n_brt=self.t.n_brt, n_nodes=self.t.n_nodes)

    def download(self, stream=None) -> dict:
        """Copy every output array of the last build to host numpy arrays."""
        cnt = self.read_counts(stream)
        u, nb, no = cnt["n_unique"], cnt["n_brt"], cnt["n_nodes"]
        n = self.t.n_points
        kdt = np.uint32 if self.t.code_bits == 30 else np.uint64

        def dl(ptr, count, dtype):
            out = np.empty(count, dtype)
            if count:
                hip = _hip()
                st = hip.hipMemcpy(C.c_void_p(out.ctypes.data), C.c_void_p(ptr), C.c_size_t(out.nbytes), 2)
                if st != 0:
                    raise TcError(-4, "hipMemcpy")
            return out

        res = dict(cnt)
        res["sorted"] = dl(self.t.sorted, n, kdt)
        res["unique"] = dl(self.t.unique, u, kdt)
        res["prefix_n"] = dl(self.t.prefix_n, nb, np.uint8)
        res["has_leaf_left"] = dl(self.t.has_leaf_left, nb, np.uint8)
        res["has_leaf_right"] = dl(self.t.has_leaf_right, nb, np.uint8)
        res["left_child"] = dl(self.t.left_child, nb, np.int32)
        res["parent"] = dl(self.t.parent, nb, np.int32)
        res["edge_count"] = dl(self.t.edge_count, nb, np.uint32)
        res["offsets"] = dl(self.t.offsets, nb + 1, np.uint32) if n else np.zeros(1, np.uint32)
        raw = dl(self.t.nodes, no * 64, np.uint8)
        res["nodes"] = raw.view(OCT_NODE_DTYPE)
        if self.t.point_index:
            res["point_index"] = dl(self.t.point_index, n, np.uint32)
            res["leaf_first"] = dl(self.t.leaf_first, u + 1, np.uint32) if n else np.zeros(1, np.uint32)
        return res


_hip_lib = None


def _hip():
    global _hip_lib
    if _hip_lib is None:
        _hip_lib = C.CDLL("libamdhip64.so")
        _hip_lib.hipMemcpy.restype = C.c_int
        _hip_lib.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    return _hip_lib


def default_context() -> Context:
    global _default_ctx
    if _default_ctx is None:
        _default_ctx = Context(0)
    return _default_ctx


# ------------------------------------------------------------------------------
# Host-array mirrors of the reference's per-stage functions.
# ------------------------------------------------------------------------------
def _pts(points) -> np.ndarray:
    p = np.ascontiguousarray(points, dtype=np.float32)
    if p.ndim != 2 or p.shape[1] != 4:
        raise ValueError("points must be float32[n][4]")
    return p


def compute_morton(input_points, n: int, min_coord: float, range_: float, group_size: int = 256, ctx=None):
    """compute_morton, src/morton.rs:76-83: 30-bit codes of float[n][4] points."""
    ctx = ctx or default_context()
    pts = _pts(input_points)
    if n > len(pts):
        raise ValueError("n exceeds the number of points")
    out = np.empty(n, np.uint32)
    check(lib().tc_host_compute_morton(ctx.handle, _np_ptr(pts), _np_ptr(out), n, min_coord, range_, group_size),
          "compute_morton")
    return out


def histogram(input_size: int, input_data, ctx=None):
    """histogram, src/histogram.rs:38: u32[4*256] exclusive digit prefixes."""
    ctx = ctx or default_context()
    keys = np.ascontiguousarray(input_data, dtype=np.uint32)
    out = np.empty(1024, np.uint32)
    check(lib().tc_host_histogram(ctx.handle, input_size, _np_ptr(keys), _np_ptr(out)), "histogram")
    return out


def radix_sort(in_data, n: int, group_size: int = 256, ctx=None):
    """radix_sort, src/radixsort.rs:36 (contract of src/main.rs:85: ascending)."""
    ctx = ctx or default_context()
    keys = np.ascontiguousarray(in_data, dtype=np.uint32)
    out = np.empty(n, np.uint32)
    check(lib().tc_host_radix_sort(ctx.handle, _np_ptr(keys), n, _np_ptr(out), group_size), "radix_sort")
    return out


def unique(u_sort: np.ndarray, n: int, ctx=None) -> int:
    """unique, src/unique.rs:37: in place on u_sort, returns the unique count."""
    ctx = ctx or default_context()
    if u_sort.dtype != np.uint32 or not u_sort.flags["C_CONTIGUOUS"]:
        raise ValueError("u_sort must be a contiguous uint32 array")
    cnt = C.c_uint32()
    check(lib().tc_host_unique(ctx.handle, _np_ptr(u_sort), n, C.byref(cnt)), "unique")
    return cnt.value


def build_radix_tree(n: int, codes, prefix_n, has_leaf_left, has_leaf_right, left_child, parent,
                     group_size: int = 256, ctx=None):
    """build_radix_tree, src/build_radix_tree.rs:34-43: n = node count (n_unique - 1)."""
    ctx = ctx or default_context()
    c = np.ascontiguousarray(codes, dtype=np.uint32)
    check(lib().tc_host_build_radix_tree(ctx.handle, n, _np_ptr(c), _np_ptr(prefix_n), _np_ptr(has_leaf_left),
                                         _np_ptr(has_leaf_right), _np_ptr(left_child), _np_ptr(parent), group_size),
          "build_radix_tree")


def edge_count(prefix_n, parent, u_edge_count, n_brt_nodes: int, group_size: int = 256, ctx=None):
    """edge_count, src/edge_count.rs:41-47."""
    ctx = ctx or default_context()
    check(lib().tc_host_edge_count(ctx.handle, _np_ptr(np.ascontiguousarray(prefix_n, np.uint8)),
                                   _np_ptr(np.ascontiguousarray(parent, np.int32)), _np_ptr(u_edge_count),
                                   n_brt_nodes, group_size), "edge_count")


def partial_sum(u_edge_count, first: int, last: int, d_first, ctx=None):
    """partial_sum, src/partial_sum.rs:43 (exclusive; d_first[last-first] = total)."""
    ctx = ctx or default_context()
    src = np.ascontiguousarray(u_edge_count, np.uint32)
    if len(d_first) < last - first + 1:
        raise ValueError("d_first needs last - first + 1 entries")
    check(lib().tc_host_partial_sum(ctx.handle, _np_ptr(src), first, last, _np_ptr(d_first)), "partial_sum")


def build_octree(oct_nodes, node_offsets, rt_node_counts, codes, rt_prefix_n, rt_parents, min_coord, range_, n,
                 rt_has_leaf_left, rt_has_leaf_right, rt_left_child, group_size: int = 256, ctx=None):
    """build_octree, src/octree.rs:73-87: n radix nodes; oct_nodes (OCT_NODE_DTYPE)
    must hold node_offsets[n] entries."""
    ctx = ctx or default_context()
    offs = np.ascontiguousarray(node_offsets, np.uint32)
    if n > 0 and len(oct_nodes) < int(offs[n]):
        raise ValueError("oct_nodes too small")
    check(lib().tc_host_build_octree(
        ctx.handle, _np_ptr(oct_nodes), _np_ptr(offs), _np_ptr(np.ascontiguousarray(rt_node_counts, np.uint32)),
        _np_ptr(np.ascontiguousarray(codes, np.uint32)), _np_ptr(np.ascontiguousarray(rt_prefix_n, np.uint8)),
        _np_ptr(np.ascontiguousarray(rt_parents, np.int32)), min_coord, range_, n,
        _np_ptr(np.ascontiguousarray(rt_has_leaf_left, np.uint8)),
        _np_ptr(np.ascontiguousarray(rt_has_leaf_right, np.uint8)),
        _np_ptr(np.ascontiguousarray(rt_left_child, np.int32)), group_size), "build_octree")


def morton32_to_xyz(code: int, min_coord: float, range_: float) -> np.ndarray:
    """morton32_to_xyz, src/morton.rs:43 (host only)."""
    out = np.empty(4, np.float32)
    lib().tc_morton32_to_xyz(_np_ptr(out), int(code), min_coord, range_)
    return out


def morton64_to_xyz(code: int, min_coord: float, range_: float) -> np.ndarray:
    out = np.empty(4, np.float32)
    lib().tc_morton64_to_xyz(_np_ptr(out), int(code), min_coord, range_)
    return out


def octree_node_bound(n_keys: int, code_bits: int = 30) -> int:
    return int(lib().tc_octree_node_bound(n_keys, code_bits))


# ---------------------------------------------------------------- generators
def gen_uniform(n: int, seed: int, extent: float = 1024.0) -> np.ndarray:
    out = np.empty((n, 4), np.float32)
    lib().tc_gen_uniform(_np_ptr(out), n, seed, extent)
    return out


def gen_gaussian_mixture(n: int, seed: int, n_centres: int = 64, sigma: float = 8.0,
                         extent: float = 1024.0) -> np.ndarray:
    out = np.empty((n, 4), np.float32)
    lib().tc_gen_gaussian_mixture(_np_ptr(out), n, seed, n_centres, sigma, extent)
    return out
