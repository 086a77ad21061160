"""ctypes bindings of libtc.so (the C ABI declared in include/tc.h).

The library is built in-tree (``make`` in the package directory, or
``__graft_entry__.build()``).  Loading fails loudly when it is missing: there is
no CPU fallback for any stage.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_DIR, "libtc.so")

TC_OK = 0
LOOKBACK_SPIN_DEFAULT = 64  # TC_LOOKBACK_SPIN_DEFAULT in tc.h
STATUS_NAMES = {
    0: "TC_OK", -1: "TC_ERR_INVALID", -2: "TC_ERR_NOMEM", -3: "TC_ERR_TOO_LARGE", -4: "TC_ERR_HIP",
    -5: "TC_ERR_OVER_BUDGET", -6: "TC_ERR_CAPACITY", -7: "TC_ERR_BUSY", -8: "TC_ERR_NOT_FOUND",
    -9: "TC_ERR_NO_DEVICE", -10: "TC_ERR_BAD_TREE", -11: "TC_ERR_IO", -12: "TC_ERR_CANCELLED",
}
TC_ERR_INVALID, TC_ERR_NOMEM, TC_ERR_TOO_LARGE, TC_ERR_HIP = -1, -2, -3, -4
TC_ERR_OVER_BUDGET, TC_ERR_CAPACITY, TC_ERR_BUSY, TC_ERR_NOT_FOUND, TC_ERR_NO_DEVICE = -5, -6, -7, -8, -9
TC_ERR_BAD_TREE, TC_ERR_IO, TC_ERR_CANCELLED = -10, -11, -12

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


class TcError(RuntimeError):
    def __init__(self, status: int, where: str = ""):
        self.status = status
        name = STATUS_NAMES.get(status, str(status))
        super().__init__(f"{where}: {name}" if where else name)


class TcTree(C.Structure):
    _fields_ = [
        ("capacity", C.c_uint32),
        ("code_bits", C.c_int),
        ("node_capacity", C.c_uint64),
        ("sorted", C.c_void_p),
        ("alt", C.c_void_p),
        ("unique", C.c_void_p),
        ("prefix_n", C.c_void_p),
        ("has_leaf_left", C.c_void_p),
        ("has_leaf_right", C.c_void_p),
        ("left_child", C.c_void_p),
        ("parent", C.c_void_p),
        ("edge_count", C.c_void_p),
        ("offsets", C.c_void_p),
        ("nodes", C.c_void_p),
        ("d_counts", C.c_void_p),
        ("n_points", C.c_uint32),
        ("n_unique", C.c_uint32),
        ("n_brt", C.c_uint32),
        ("n_nodes", C.c_uint32),
        ("point_index", C.c_void_p),
        ("index_alt", C.c_void_p),
        ("leaf_first", C.c_void_p),
        ("min_coord", C.c_float),
        ("range", C.c_float),
        ("link_leaves", C.c_int),
    ]


class TcTreeHost(C.Structure):
    """tc_tree_host: one tree in host memory (optional arrays may be NULL)."""
    _fields_ = [
        ("code_bits", C.c_int),
        ("min_coord", C.c_float),
        ("range", C.c_float),
        ("link_leaves", C.c_int),
        ("n_points", C.c_uint32),
        ("n_unique", C.c_uint32),
        ("n_nodes", C.c_uint32),
        ("sorted", C.c_void_p),
        ("unique", C.c_void_p),
        ("prefix_n", C.c_void_p),
        ("has_leaf_left", C.c_void_p),
        ("has_leaf_right", C.c_void_p),
        ("left_child", C.c_void_p),
        ("parent", C.c_void_p),
        ("edge_count", C.c_void_p),
        ("offsets", C.c_void_p),
        ("nodes", C.c_void_p),
        ("point_index", C.c_void_p),
        ("leaf_first", C.c_void_p),
    ]


TREE_POINT_INDEX = 1  # tc_tree_alloc_ex flag


CLOUD_INFO_DTYPE = np.dtype(
    [("point_offset", "<u4"), ("n_points", "<u4"), ("unique_offset", "<u4"), ("n_unique", "<u4"),
     ("node_offset", "<u4"), ("n_nodes", "<u4")]
)


class TcBatchView(C.Structure):
    _fields_ = [
        ("code_bits", C.c_int),
        ("n_clouds", C.c_uint32),
        ("n_points", C.c_uint32),
        ("sorted", C.c_void_p),
        ("unique", C.c_void_p),
        ("prefix_n", C.c_void_p),
        ("has_leaf_left", C.c_void_p),
        ("has_leaf_right", C.c_void_p),
        ("left_child", C.c_void_p),
        ("parent", C.c_void_p),
        ("edge_count", C.c_void_p),
        ("offsets", C.c_void_p),
        ("nodes", C.c_void_p),
        ("node_capacity", C.c_uint64),
    ]


class TcTreePtrs(C.Structure):
    _fields_ = [
        ("unique", C.c_void_p),
        ("prefix_n", C.c_void_p),
        ("has_leaf_left", C.c_void_p),
        ("has_leaf_right", C.c_void_p),
        ("left_child", C.c_void_p),
        ("parent", C.c_void_p),
        ("edge_count", C.c_void_p),
        ("offsets", C.c_void_p),
        ("nodes", C.c_void_p),
    ]


class TcServerConfig(C.Structure):
    _fields_ = [
        ("mem_budget_bytes", C.c_uint64),
        ("n_streams", C.c_uint32),
        ("max_batch_clouds", C.c_uint32),
        ("max_batch_points", C.c_uint32),
        ("code_bits", C.c_int),
        ("link_leaves", C.c_int),
        ("max_queued", C.c_uint32),
    ]


class TcServerResult(C.Structure):
    _fields_ = [
        ("id", C.c_uint64),
        ("status", C.c_int32),
        ("n_points", C.c_uint32),
        ("n_unique", C.c_uint32),
        ("n_nodes", C.c_uint32),
        ("tree", TcTreePtrs),
        ("t_submit", C.c_double),
        ("t_start", C.c_double),
        ("t_done", C.c_double),
    ]


class TcServerStats(C.Structure):
    _fields_ = [
        ("submitted", C.c_uint64),
        ("completed", C.c_uint64),
        ("rejected", C.c_uint64),
        ("batches", C.c_uint64),
        ("points_completed", C.c_uint64),
        ("mean_clouds_per_batch", C.c_double),
        ("latency_p50_ms", C.c_double),
        ("latency_p99_ms", C.c_double),
        ("latency_mean_ms", C.c_double),
        ("latency_max_ms", C.c_double),
        ("first_submit_s", C.c_double),
        ("last_done_s", C.c_double),
        ("reserved_bytes", C.c_uint64),
        ("peak_reserved_bytes", C.c_uint64),
        ("result_budget_bytes", C.c_uint64),
        ("cancelled", C.c_uint64),
        ("queued", C.c_uint64),
    ]


class TcSimConfig(C.Structure):
    _fields_ = [
        ("n_streams", C.c_uint32),
        ("max_batch_clouds", C.c_uint32),
        ("max_batch_points", C.c_uint32),
        ("result_budget_bytes", C.c_uint64),
        ("code_bits", C.c_int),
        ("launch_s", C.c_double),
        ("per_point_s", C.c_double),
        ("hold_s", C.c_double),
        ("max_queued", C.c_uint32),
        ("alloc_fail_after", C.c_uint32),
        ("alloc_fail_count", C.c_uint32),
    ]


P = C.c_void_p
I = C.c_int
U32 = C.c_uint32
U64 = C.c_uint64
F = C.c_float
S = C.c_void_p  # stream

# name -> (restype, argtypes)
SIGNATURES = {
    "tc_status_string": (C.c_char_p, [I]),
    "tc_version": (C.c_char_p, []),
    "tc_create": (I, [C.POINTER(C.c_void_p), I]),
    "tc_destroy": (I, [P]),
    "tc_reserve": (I, [P, U32]),
    "tc_set_sort_params": (I, [P, I, I]),
    "tc_set_lookback_spin_limit": (I, [P, U32]),
    "tc_profile_enable": (I, [P, I]),
    "tc_set_markers": (I, [P, I]),
    "tc_profile_collect": (I, [P]),
    "tc_profile_reset": (I, [P]),
    "tc_profile_get": (I, [P, I, C.POINTER(C.c_double), C.POINTER(U64)]),
    "tc_kernel_name": (C.c_char_p, [I]),
    "tc_get_sort_params": (I, [P, C.POINTER(I), C.POINTER(I)]),
    "tc_autotune_sort": (I, [P, U32]),
    "tc_tuned_sort_shape": (I, [P, I, I, U32, C.POINTER(I), C.POINTER(I)]),
    "tc_set_tuned_sort_shape": (I, [P, I, I, I, I, I]),
    "tc_clear_sort_tuning": (I, [P]),
    "tc_sort_shape_list": (I, [C.POINTER(I), C.POINTER(I), I]),
    "tc_sort_plan_list": (I, [C.POINTER(I), C.POINTER(I), C.POINTER(I), I]),
    "tc_set_sort_plan": (I, [P, I, I, I]),
    "tc_get_sort_plan": (I, [P, C.POINTER(I), C.POINTER(I), C.POINTER(I)]),
    "tc_tuned_sort_plan": (I, [P, I, I, U32, C.POINTER(I), C.POINTER(I), C.POINTER(I)]),
    "tc_set_tuned_sort_plan": (I, [P, I, I, I, I, I, I]),
    "tc_sort_keys_u32_noalloc": (I, [P, P, P, U32, S]),
    "tc_traffic_reference": (I, [P, P, C.c_size_t, I, P, S]),
    "tc_morton30": (I, [P, P, P, U32, F, F, S]),
    "tc_morton63": (I, [P, P, P, U32, F, F, S]),
    "tc_morton30_decode": (I, [P, P, P, U32, F, F, S]),
    "tc_morton32_to_xyz": (None, [P, U32, F, F]),
    "tc_morton64_to_xyz": (None, [P, U64, F, F]),
    "tc_global_histogram_u32": (I, [P, P, U32, P, S]),
    "tc_global_histogram_u64": (I, [P, P, U32, P, S]),
    "tc_binning_pass_u32": (I, [P, P, P, P, P, U32, P, U32, S]),
    "tc_binning_pass_u64": (I, [P, P, P, P, P, U32, P, U32, S]),
    "tc_sort_keys_u32": (I, [P, P, P, U32, S]),
    "tc_sort_pairs_u32": (I, [P, P, P, P, P, U32, S]),
    "tc_sort_keys_u64": (I, [P, P, P, U32, S]),
    "tc_morton_sort_u32": (I, [P, P, P, P, U32, F, F, S]),
    "tc_morton_sort_u64": (I, [P, P, P, P, U32, F, F, S]),
    "tc_sort_pairs_u64": (I, [P, P, P, P, P, U32, S]),
    "tc_morton_sort_pairs_u32": (I, [P, P, P, P, P, P, U32, F, F, S]),
    "tc_morton_sort_pairs_u64": (I, [P, P, P, P, P, P, U32, F, F, S]),
    "tc_unique_u32": (I, [P, P, P, U32, P, S]),
    "tc_unique_u64": (I, [P, P, P, U32, P, S]),
    "tc_unique_runs_u32": (I, [P, P, P, P, U32, P, S]),
    "tc_unique_runs_u64": (I, [P, P, P, P, U32, P, S]),
    "tc_build_radix_tree_u32": (I, [P, P, U32, P, P, P, P, P, P, S]),
    "tc_build_radix_tree_u64": (I, [P, P, U32, P, P, P, P, P, P, S]),
    "tc_build_radix_tree_ec_u32": (I, [P, P, U32, P, P, P, P, P, P, P, S]),
    "tc_build_radix_tree_ec_u64": (I, [P, P, U32, P, P, P, P, P, P, P, S]),
    "tc_edge_count": (I, [P, P, P, P, U32, P, S]),
    "tc_exclusive_scan_u32": (I, [P, P, P, U32, P, S]),
    "tc_edge_count_scan": (I, [P, P, P, P, P, U32, P, S]),
    "tc_build_octree_u32": (I, [P, P, U32, P, P, P, P, P, P, P, P, U32, P, F, F, I, S]),
    "tc_build_octree_u64": (I, [P, P, U32, P, P, P, P, P, P, P, P, U32, P, F, F, I, S]),
    "tc_octree_node_bound": (U64, [U64, I]),
    "tc_tree_alloc": (I, [P, U32, I, C.POINTER(TcTree)]),
    "tc_tree_alloc_ex": (I, [P, U32, I, I, C.POINTER(TcTree)]),
    "tc_tree_free": (I, [P, C.POINTER(TcTree)]),
    "tc_build_tree": (I, [P, P, U32, F, F, I, C.POINTER(TcTree), S]),
    "tc_tree_read_counts": (I, [P, C.POINTER(TcTree), S]),
    "tc_tree_validate_host": (I, [C.POINTER(TcTreeHost), C.c_char_p, C.c_size_t]),
    "tc_tree_write_host": (I, [C.POINTER(TcTreeHost), C.c_char_p]),
    "tc_tree_validate": (I, [P, C.POINTER(TcTree), C.c_char_p, C.c_size_t, S]),
    "tc_tree_export": (I, [P, C.POINTER(TcTree), C.c_char_p, S]),
    "tc_host_compute_morton": (I, [P, P, P, U32, F, F, U32]),
    "tc_host_histogram": (I, [P, U32, P, P]),
    "tc_host_radix_sort": (I, [P, P, U32, P, U32]),
    "tc_host_unique": (I, [P, P, U32, C.POINTER(U32)]),
    "tc_host_build_radix_tree": (I, [P, C.c_int32, P, P, P, P, P, P, U32]),
    "tc_host_edge_count": (I, [P, P, P, P, U32, U32]),
    "tc_host_partial_sum": (I, [P, P, U32, U32, P]),
    "tc_host_build_octree": (I, [P, P, P, P, P, P, P, F, F, C.c_int32, P, P, P, U32]),
    "tc_batch_create": (I, [P, U32, U32, I, C.POINTER(C.c_void_p)]),
    "tc_batch_destroy": (I, [P]),
    "tc_batch_set": (I, [P, U32, P, P, P, S]),
    "tc_batch_build": (I, [P, P, I, S]),
    "tc_batch_capture": (I, [P, P, I, S]),
    "tc_batch_replay": (I, [P, S]),
    "tc_batch_read_info": (I, [P, P, S]),
    "tc_batch_arrays": (I, [P, P]),
    "tc_batch_set_io": (I, [P, U32, P, P, P, P, P, S]),
    "tc_batch_read_info_async": (I, [P, P, P, S]),
    "tc_batch_memory_bytes": (U64, [U32, U32, I]),
    "tc_tree_result_bytes": (U64, [U32, I]),
    "tc_server_config_default": (None, [C.POINTER(TcServerConfig)]),
    "tc_server_create": (I, [P, C.POINTER(TcServerConfig), C.POINTER(C.c_void_p)]),
    "tc_server_destroy": (I, [P]),
    "tc_server_submit": (I, [P, P, U32, F, F, C.POINTER(U64)]),
    "tc_server_poll": (I, [P, C.POINTER(TcServerResult), U32, C.POINTER(U32), I]),
    "tc_server_release": (I, [P, U64]),
    "tc_server_drain": (I, [P]),
    "tc_server_stats_get": (I, [P, C.POINTER(TcServerStats)]),
    "tc_server_stats_reset": (I, [P]),
    "tc_server_run_trace": (I, [P, U32, P, P, P, F, F, P]),
    "tc_plain_run_trace": (I, [P, U32, P, P, P, F, F, I, P]),
    "tc_server_simulate": (I, [C.POINTER(TcSimConfig), U32, P, P, P, P, P, C.POINTER(U64)]),
    "tc_server_simulate_ex": (I, [C.POINTER(TcSimConfig), U32, P, P, P, P, P, P, P, C.POINTER(U64)]),
    "tc_server_cancel": (I, [P, U64]),
    "tc_gen_uniform": (None, [P, U64, U64, F]),
    "tc_gen_gaussian_mixture": (None, [P, U64, U64, U32, F, F]),
}

_lib = None


def lib():
    """Load libtc.so; raise if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"libtc.so not found at {LIB_PATH}: build it first (make -C {PKG_DIR}, or __graft_entry__.build())"
            )
        # libtc needs libamdhip64.so.7.  PyTorch ships its own copy with that
        # soname; loading torch first makes libtc bind to the same runtime, so
        # torch's streams and allocations are valid handles for libtc.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def extra_signatures(names_and_sigs: dict):
    """Register signatures of further exported functions (used by submodules)."""
    L = lib()
    for name, (res, args) in names_and_sigs.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
        SIGNATURES.setdefault(name, (res, args))


def check(status: int, where: str = "") -> None:
    if status != TC_OK:
        raise TcError(status, where)
