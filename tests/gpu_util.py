"""Helpers for GPU tests: numpy <-> device tensors (unsigned arrays travel as
signed views of the same bytes) and a stage-by-stage runner on a tc.Context."""
import numpy as np

_VIEW = {np.dtype(np.uint32): np.int32, np.dtype(np.uint64): np.int64, np.dtype(np.uint8): np.uint8,
         np.dtype(np.int32): np.int32, np.dtype(np.float32): np.float32, np.dtype(np.int64): np.int64}


def dev(a: np.ndarray):
    import torch

    a = np.ascontiguousarray(a)
    return torch.from_numpy(a.view(_VIEW[a.dtype]).copy()).cuda()


def empty(n: int, dtype):
    import torch

    tdt = {np.uint32: torch.int32, np.int32: torch.int32, np.uint64: torch.int64, np.uint8: torch.uint8,
           np.float32: torch.float32}[dtype]
    return torch.empty(max(int(n), 1), dtype=tdt, device="cuda")


def host(t, dtype, n=None):
    import torch

    torch.cuda.synchronize()
    a = t.cpu().numpy()
    a = a.view(dtype)
    return a if n is None else a[:n].copy()


def run_stages(ctx, points: np.ndarray, min_coord: float, range_: float, code_bits: int = 30, link_leaves=True):
    """Run every stage through its own C-ABI entry point (no fused calls) and
    return host copies of all outputs."""
    import torch

    n = len(points)
    kdt = np.uint32 if code_bits == 30 else np.uint64
    pts = dev(points.astype(np.float32))
    codes = empty(n, kdt)
    (ctx.morton30 if code_bits == 30 else ctx.morton63)(pts, codes, min_coord, range_)
    out = {"codes": host(codes, kdt, n)}
    hist = empty(256 * (4 if code_bits == 30 else 8), np.uint32)
    ctx.global_histogram(codes, hist, n=n)
    out["hist"] = host(hist, np.uint32)
    keys = codes.clone()
    alt = empty(n, kdt)
    ctx.sort(keys, alt, n=n)
    out["sorted"] = host(keys, kdt, n)
    uniq = empty(n, kdt)
    cnt = empty(1, np.uint32)
    cnt.zero_()
    ctx.unique(keys, uniq, cnt, n=n)
    u = int(host(cnt, np.uint32)[0])
    out["unique"] = host(uniq, kdt, u)
    nb = max(u - 1, 0)
    pn, hll, hlr = empty(nb, np.uint8), empty(nb, np.uint8), empty(nb, np.uint8)
    left, par = empty(nb, np.int32), empty(nb, np.int32)
    ctx.build_radix_tree(uniq, u, pn, hll, hlr, left, par)
    out.update(prefix_n=host(pn, np.uint8, nb), has_leaf_left=host(hll, np.uint8, nb),
               has_leaf_right=host(hlr, np.uint8, nb), left_child=host(left, np.int32, nb),
               parent=host(par, np.int32, nb))
    ec = empty(nb, np.uint32)
    ctx.edge_count(pn, par, ec, nb)
    out["edge_count"] = host(ec, np.uint32, nb)
    offs = empty(nb + 1, np.uint32)
    ctx.exclusive_scan(ec, offs, n=nb)
    out["offsets"] = host(offs, np.uint32, nb + 1)
    total = int(out["offsets"][-1]) if nb else 0
    nodes = torch.zeros(max(total, 1) * 16, dtype=torch.int32, device="cuda")
    ctx.build_octree(nodes, total, offs, ec, uniq, pn, par, left, hll, hlr, nb, min_coord, range_, link_leaves)
    out["nodes"] = host(nodes, np.uint8)[: total * 64]
    torch.cuda.synchronize()
    return out
