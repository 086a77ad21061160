"""Batched builds: many clouds in one segmented launch per stage (tc_batch_*).

Each cloud's outputs equal a plain single-cloud build of its points; the
arrays are concatenated with per-cloud offsets (``info()``).  ``capture()``
records the whole batched pipeline as a HIP graph; after ``set()`` uploads a
new plan (any batch within the capacity), ``replay()`` runs it again.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import CLOUD_INFO_DTYPE, OCT_NODE_DTYPE, TcBatchView, TcError, check, lib
from .api import Context, _hip, _stream, _t_ptr


class Batch:
    def __init__(self, ctx: Context, max_clouds: int, max_points: int, code_bits: int = 30):
        self.ctx = ctx
        self.code_bits = code_bits
        self._h = C.c_void_p()
        check(lib().tc_batch_create(ctx.handle, max_clouds, max_points, code_bits, C.byref(self._h)),
              "tc_batch_create")
        self.n_clouds = 0

    def close(self):
        if self._h:
            lib().tc_batch_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set(self, counts, mins=None, ranges=None, stream=None):
        counts = np.ascontiguousarray(counts, np.uint32)
        mn = None if mins is None else np.ascontiguousarray(mins, np.float32)
        rg = None if ranges is None else np.ascontiguousarray(ranges, np.float32)
        self._keep = (counts, mn, rg)  # the upload is asynchronous
        check(lib().tc_batch_set(self._h, len(counts), counts.ctypes.data,
                                 None if mn is None else mn.ctypes.data,
                                 None if rg is None else rg.ctypes.data, _stream(stream)), "tc_batch_set")
        self.n_clouds = len(counts)

    def build(self, pts, link_leaves=True, stream=None):
        check(lib().tc_batch_build(self._h, _t_ptr(pts), int(bool(link_leaves)), _stream(stream)), "tc_batch_build")

    def capture(self, pts, link_leaves=True, stream=None):
        check(lib().tc_batch_capture(self._h, _t_ptr(pts), int(bool(link_leaves)), _stream(stream)),
              "tc_batch_capture")

    def replay(self, stream=None):
        check(lib().tc_batch_replay(self._h, _stream(stream)), "tc_batch_replay")

    def info(self, stream=None) -> np.ndarray:
        out = np.zeros(max(self.n_clouds, 1), CLOUD_INFO_DTYPE)
        check(lib().tc_batch_read_info(self._h, out.ctypes.data, _stream(stream)), "tc_batch_read_info")
        return out[: self.n_clouds]

    def view(self) -> TcBatchView:
        v = TcBatchView()
        check(lib().tc_batch_arrays(self._h, C.byref(v)), "tc_batch_arrays")
        return v

    def download(self, stream=None) -> list:
        """Per-cloud host copies of every output array, in the layout of a plain build."""
        info = self.info(stream)
        v = self.view()
        kdt = np.uint32 if self.code_bits == 30 else np.uint64
        n_pts = int(info["point_offset"][-1] + info["n_points"][-1]) if len(info) else 0
        n_u = int(info["unique_offset"][-1] + info["n_unique"][-1]) if len(info) else 0
        n_o = int(info["node_offset"][-1] + info["n_nodes"][-1]) if len(info) else 0

        def dl(ptr, count, dtype):
            out = np.empty(count, dtype)
            if count:
                st = _hip().hipMemcpy(C.c_void_p(out.ctypes.data), C.c_void_p(ptr), C.c_size_t(out.nbytes), 2)
                if st != 0:
                    raise TcError(-4, "hipMemcpy")
            return out

        sorted_ = dl(v.sorted, n_pts, kdt)
        uniq = dl(v.unique, n_u, kdt)
        pn = dl(v.prefix_n, n_u, np.uint8)
        hll = dl(v.has_leaf_left, n_u, np.uint8)
        hlr = dl(v.has_leaf_right, n_u, np.uint8)
        left = dl(v.left_child, n_u, np.int32)
        par = dl(v.parent, n_u, np.int32)
        ec = dl(v.edge_count, n_u, np.uint32)
        offs = dl(v.offsets, n_u + 1, np.uint32)
        nodes = dl(v.nodes, n_o * 64, np.uint8).view(OCT_NODE_DTYPE)
        res = []
        for c in info:
            p0, n = int(c["point_offset"]), int(c["n_points"])
            u0, u = int(c["unique_offset"]), int(c["n_unique"])
            o0, o = int(c["node_offset"]), int(c["n_nodes"])
            nb = max(u - 1, 0)
            res.append(dict(
                n_points=n, n_unique=u, n_brt=nb, n_nodes=o,
                sorted=sorted_[p0:p0 + n], unique=uniq[u0:u0 + u],
                prefix_n=pn[u0:u0 + nb], has_leaf_left=hll[u0:u0 + nb], has_leaf_right=hlr[u0:u0 + nb],
                left_child=left[u0:u0 + nb], parent=par[u0:u0 + nb], edge_count=ec[u0:u0 + nb],
                offsets=(offs[u0:u0 + nb + 1] - np.uint32(o0)) if u > 0 else np.zeros(1, np.uint32),
                nodes=nodes[o0:o0 + o],
            ))
        return res
