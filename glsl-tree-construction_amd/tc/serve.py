"""Serving: tree builds admitted as they arrive (tc_server_*).

A ``Server`` owns a scheduler thread and ``n_streams`` batch slots.  Requests
name device point tensors; whenever a slot is free the requests waiting at that
moment go out as one segmented, graph-replayed build (arrival order, within the
batch limits and the device-memory budget).  Each result stays in device memory
-- equal to a plain build of the same points -- until ``release``.

``simulate`` runs the same admission/batching policy against a modelled device
(no GPU), which is how the policy is tested on a CPU.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import (OCT_NODE_DTYPE, TcError, TcServerConfig, TcServerResult, TcServerStats, TcSimConfig, check,
                   lib)
from .api import Context, _hip, _t_ptr

RESULT_FIELDS = ("unique", "prefix_n", "has_leaf_left", "has_leaf_right", "left_child", "parent", "edge_count",
                 "offsets", "nodes")


def _dl(ptr, count, dtype):
    out = np.empty(count, dtype)
    if count:
        st = _hip().hipMemcpy(C.c_void_p(out.ctypes.data), C.c_void_p(ptr), C.c_size_t(out.nbytes), 2)
        if st != 0:
            raise TcError(-4, "hipMemcpy")
    return out


def result_bytes(n: int, code_bits: int = 30) -> int:
    return int(lib().tc_tree_result_bytes(n, code_bits))


class Server:
    def __init__(self, ctx: Context, mem_budget_bytes: int = 64 << 30, n_streams: int = 2,
                 max_batch_clouds: int = 256, max_batch_points: int = 1 << 24, code_bits: int = 30,
                 link_leaves: bool = True, max_queued: int = 1 << 16):
        cfg = TcServerConfig()
        lib().tc_server_config_default(C.byref(cfg))
        cfg.mem_budget_bytes = mem_budget_bytes
        cfg.n_streams = n_streams
        cfg.max_batch_clouds = max_batch_clouds
        cfg.max_batch_points = max_batch_points
        cfg.code_bits = code_bits
        cfg.link_leaves = int(bool(link_leaves))
        cfg.max_queued = max_queued
        self.ctx = ctx
        self.code_bits = code_bits
        self._h = C.c_void_p()
        check(lib().tc_server_create(ctx.handle, C.byref(cfg), C.byref(self._h)), "tc_server_create")

    def close(self):
        if self._h:
            lib().tc_server_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def submit(self, pts, n=None, min_coord=0.0, range_=1024.0) -> int:
        """Queue a build of the first n points of a device float32 [N, 4] tensor
        (which must stay alive until its result has been polled)."""
        n = pts.shape[0] if n is None else n
        rid = C.c_uint64()
        check(lib().tc_server_submit(self._h, _t_ptr(pts), n, min_coord, range_, C.byref(rid)), "tc_server_submit")
        return rid.value

    def poll(self, max_results: int = 64, wait_ms: int = 0) -> list:
        buf = (TcServerResult * max_results)()
        k = C.c_uint32()
        check(lib().tc_server_poll(self._h, buf, max_results, C.byref(k), wait_ms), "tc_server_poll")
        out = []
        for r in buf[:k.value]:
            out.append(dict(id=r.id, status=r.status, n_points=r.n_points, n_unique=r.n_unique, n_nodes=r.n_nodes,
                            tree={f: getattr(r.tree, f) for f in RESULT_FIELDS},
                            t_submit=r.t_submit, t_start=r.t_start, t_done=r.t_done))
        return out

    def cancel(self, rid: int):
        """Drop a request that has not started (TcError TC_ERR_BUSY if it is
        being built, TC_ERR_NOT_FOUND if unknown or finished); it then comes
        back from poll with status TC_ERR_CANCELLED."""
        check(lib().tc_server_cancel(self._h, rid), "tc_server_cancel")

    def release(self, rid: int):
        check(lib().tc_server_release(self._h, rid), "tc_server_release")

    def drain(self):
        check(lib().tc_server_drain(self._h), "tc_server_drain")

    def stats(self) -> dict:
        st = TcServerStats()
        check(lib().tc_server_stats_get(self._h, C.byref(st)), "tc_server_stats_get")
        return {f: getattr(st, f) for f, _ in TcServerStats._fields_}

    def reset_stats(self):
        check(lib().tc_server_stats_reset(self._h), "tc_server_stats_reset")

    def download(self, res: dict) -> dict:
        """Host copies of a polled result's arrays, in the layout of Tree.download."""
        u, o = res["n_unique"], res["n_nodes"]
        nb = max(u - 1, 0)
        kdt = np.uint32 if self.code_bits == 30 else np.uint64
        t = res["tree"]
        out = dict(n_points=res["n_points"], n_unique=u, n_brt=nb, n_nodes=o)
        out["unique"] = _dl(t["unique"], u, kdt)
        out["prefix_n"] = _dl(t["prefix_n"], nb, np.uint8)
        out["has_leaf_left"] = _dl(t["has_leaf_left"], nb, np.uint8)
        out["has_leaf_right"] = _dl(t["has_leaf_right"], nb, np.uint8)
        out["left_child"] = _dl(t["left_child"], nb, np.int32)
        out["parent"] = _dl(t["parent"], nb, np.int32)
        out["edge_count"] = _dl(t["edge_count"], nb, np.uint32)
        out["offsets"] = _dl(t["offsets"], nb + 1, np.uint32) if u else np.zeros(1, np.uint32)
        out["nodes"] = _dl(t["nodes"], o * 64, np.uint8).view(OCT_NODE_DTYPE)
        return out

    def run_trace(self, t_arrive, pts_list, n_list, min_coord=0.0, range_=1024.0) -> np.ndarray:
        """Submit request i at t_arrive[i] seconds (from a native submitting
        thread), release results as they come; returns done - arrival (s)."""
        t = np.ascontiguousarray(t_arrive, np.float64)
        ptrs = (C.c_void_p * len(t))(*[_t_ptr(p) for p in pts_list])
        n = np.ascontiguousarray(n_list, np.uint32)
        lat = np.zeros(len(t), np.float64)
        check(lib().tc_server_run_trace(self._h, len(t), t.ctypes.data, ptrs, n.ctypes.data, min_coord, range_,
                                        lat.ctypes.data), "tc_server_run_trace")
        return lat


def plain_run_trace(ctx: Context, t_arrive, pts_list, n_list, min_coord=0.0, range_=1024.0,
                    code_bits: int = 30) -> np.ndarray:
    """The same trace through a one-request-at-a-time plain loop."""
    t = np.ascontiguousarray(t_arrive, np.float64)
    ptrs = (C.c_void_p * len(t))(*[_t_ptr(p) for p in pts_list])
    n = np.ascontiguousarray(n_list, np.uint32)
    lat = np.zeros(len(t), np.float64)
    check(lib().tc_plain_run_trace(ctx.handle, len(t), t.ctypes.data, ptrs, n.ctypes.data, min_coord, range_,
                                   code_bits, lat.ctypes.data), "tc_plain_run_trace")
    return lat


def simulate(t_arrive, n, n_streams=2, max_batch_clouds=256, max_batch_points=1 << 24,
             result_budget_bytes=64 << 30, code_bits=30, launch_s=50e-6, per_point_s=1e-10, hold_s=0.0,
             max_queued=0, alloc_fail_after=0, alloc_fail_count=0, t_cancel=None, check_status=True) -> dict:
    """Simulated run of the server's policy (tc_server_simulate_ex).  Returns
    per-request t_start / t_done / batch_of / status and the peak reservation;
    with check_status=False a run that ends with requests stuck (status
    TC_ERR_NOMEM) returns instead of raising."""
    cfg = TcSimConfig(n_streams, max_batch_clouds, max_batch_points, result_budget_bytes, code_bits, launch_s,
                      per_point_s, hold_s, max_queued, alloc_fail_after, alloc_fail_count)
    t = np.ascontiguousarray(t_arrive, np.float64)
    nn = np.ascontiguousarray(n, np.uint32)
    k = len(t)
    tc_ = None if t_cancel is None else np.ascontiguousarray(t_cancel, np.float64)
    t_start = np.zeros(k, np.float64)
    t_done = np.zeros(k, np.float64)
    batch_of = np.zeros(k, np.uint32)
    status = np.zeros(k, np.int32)
    peak = C.c_uint64()
    st = lib().tc_server_simulate_ex(C.byref(cfg), k, t.ctypes.data, nn.ctypes.data,
                                     None if tc_ is None else tc_.ctypes.data, t_start.ctypes.data,
                                     t_done.ctypes.data, batch_of.ctypes.data, status.ctypes.data, C.byref(peak))
    if check_status or st != -2:
        check(st, "tc_server_simulate_ex")
    return dict(t_start=t_start, t_done=t_done, batch_of=batch_of, status=status, peak_reserved=peak.value,
                result=st)
