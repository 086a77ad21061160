"""The library's host-only code -- the serving scheduler and arena
(sched.hpp) through the policy simulation, the tree validator and writer,
the input generators -- built with AddressSanitizer and UndefinedBehavior-
Sanitizer (g++, host code only: no HIP file is part of this build) and driven
on normal, saturated, cancelled, failing and malformed inputs in a child
process whose errors are fatal."""
import os
import shutil
import subprocess
import sys
import textwrap

import pytest

from conftest import PKG, ROOT

SOURCES = ("sched_sim.cpp", "treeio.cpp", "gen.cpp")


@pytest.fixture(scope="module")
def san_lib(tmp_path_factory):
    if not shutil.which("g++"):
        pytest.skip("g++ not available")
    out = tmp_path_factory.mktemp("san") / "libtc_host_san.so"
    srcs = [os.path.join(PKG, "csrc", s) for s in SOURCES]
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-fPIC", "-shared", "-fno-omit-frame-pointer",
           "-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-pthread", "-o", str(out)] + srcs
    subprocess.run(cmd, check=True)
    asan = subprocess.run(["g++", "-print-file-name=libasan.so"], check=True, capture_output=True,
                          text=True).stdout.strip()
    return str(out), asan


DRIVER = textwrap.dedent('''
    import ctypes as C, sys, os, tempfile
    import numpy as np
    sys.path.insert(0, sys.argv[2]); sys.path.insert(0, sys.argv[3])
    from tc import _lib as T
    import oracle
    L = C.CDLL(sys.argv[1])
    P = C.c_void_p
    L.tc_server_simulate_ex.argtypes = [C.POINTER(T.TcSimConfig), C.c_uint32] + [P] * 7 + [C.POINTER(C.c_uint64)]
    L.tc_tree_validate_host.argtypes = [P, C.c_char_p, C.c_size_t]
    L.tc_tree_write_host.argtypes = [P, C.c_char_p]
    L.tc_gen_uniform.argtypes = [P, C.c_uint64, C.c_uint64, C.c_float]
    L.tc_gen_gaussian_mixture.argtypes = [P, C.c_uint64, C.c_uint64, C.c_uint32, C.c_float, C.c_float]
    L.tc_tree_result_bytes.restype = C.c_uint64
    L.tc_tree_result_bytes.argtypes = [C.c_uint32, C.c_int]

    def sim(t, n, t_cancel=None, **kw):
        cfg = T.TcSimConfig(kw.get("streams", 2), kw.get("clouds", 8), kw.get("points", 1 << 22),
                            kw.get("budget", 1 << 34), 30, 1e-4, 1e-9, kw.get("hold", 0.0),
                            kw.get("max_queued", 0), kw.get("fail_after", 0), kw.get("fail_count", 0))
        k = len(t)
        t = np.ascontiguousarray(t, np.float64); n = np.ascontiguousarray(n, np.uint32)
        tc_ = None if t_cancel is None else np.ascontiguousarray(t_cancel, np.float64)
        a, b = np.zeros(k), np.zeros(k)
        bo, st = np.zeros(k, np.uint32), np.zeros(k, np.int32)
        peak = C.c_uint64()
        r = L.tc_server_simulate_ex(C.byref(cfg), k, t.ctypes.data, n.ctypes.data,
                                    None if tc_ is None else tc_.ctypes.data, a.ctypes.data, b.ctypes.data,
                                    bo.ctypes.data, st.ctypes.data, C.byref(peak))
        return r, st

    rng = np.random.default_rng(5)
    t = np.cumsum(rng.exponential(1e-5, 400)); n = rng.integers(1, 1 << 21, 400).astype(np.uint32)
    n[::50] = (1 << 23)  # above the batch limit: rejected
    assert sim(t, n)[0] == 0
    assert sim(t, n, max_queued=8)[0] == 0
    tcan = np.where(np.arange(400) % 3 == 0, t + 2e-5, np.inf)
    assert sim(t, n, tcan, max_queued=32)[0] == 0
    assert sim(t, n, fail_after=20, fail_count=50)[0] == 0
    assert sim(t, n, fail_after=20, fail_count=0xFFFFFFFF)[0] == T.TC_ERR_NOMEM
    assert sim(t, n, budget=int(L.tc_tree_result_bytes(1 << 21, 30)) * 3, hold=1e-3)[0] == 0
    assert sim(t[:0], n[:0])[0] == 0

    pts = np.zeros((20000, 4), np.float32)
    L.tc_gen_uniform(pts.ctypes.data, 20000, 3, 1024.0)
    L.tc_gen_gaussian_mixture(pts.ctypes.data, 20000, 4, 16, 4.0, 1024.0)
    w = oracle.build_tree(pts, 0.0, 1024.0, 30)
    u = len(w["unique"]); nb = u - 1
    arrs = {k: np.ascontiguousarray(w[k]) for k in ("sorted", "unique", "prefix_n", "has_leaf_left",
            "has_leaf_right", "left_child", "parent", "edge_count", "offsets", "nodes")}
    def host(a):
        h = T.TcTreeHost()
        h.code_bits, h.min_coord, h.range, h.link_leaves = 30, 0.0, 1024.0, 1
        h.n_points, h.n_unique, h.n_nodes = len(a["sorted"]), u, len(a["nodes"])
        for k, f in (("sorted", "sorted"), ("unique", "unique"), ("prefix_n", "prefix_n"),
                     ("has_leaf_left", "has_leaf_left"), ("has_leaf_right", "has_leaf_right"),
                     ("left_child", "left_child"), ("parent", "parent"), ("edge_count", "edge_count"),
                     ("offsets", "offsets"), ("nodes", "nodes")):
            setattr(h, f, a[k].ctypes.data)
        return h
    msg = C.create_string_buffer(256)
    h = host(arrs)
    assert L.tc_tree_validate_host(C.byref(h), msg, 256) == 0, msg.value
    with tempfile.TemporaryDirectory() as d:
        assert L.tc_tree_write_host(C.byref(h), os.path.join(d, "t.tct").encode()) == 0
    bad = 0
    for key, idx, val in (("left_child", 3, nb + 7), ("left_child", 5, -9), ("parent", 4, 1 << 30),
                          ("prefix_n", 2, 200), ("offsets", nb, 10 ** 6), ("edge_count", 1, 99)):
        a = {k: v.copy() for k, v in arrs.items()}
        a[key][idx] = val
        h = host(a)
        bad += L.tc_tree_validate_host(C.byref(h), msg, 256) != 0
    nodes = arrs["nodes"].copy()
    nodes["children"][0, :] = 2 ** 30; nodes["child_node_mask"][0] = 0xFF
    a = dict(arrs); a["nodes"] = nodes
    bad += L.tc_tree_validate_host(C.byref(host(a)), msg, 256) != 0
    nodes = arrs["nodes"].copy()
    nodes["child_node_mask"][:] = 0; nodes["children"][:] = 7  # leaf slots everywhere, levels too deep
    a = dict(arrs); a["nodes"] = nodes
    bad += L.tc_tree_validate_host(C.byref(host(a)), msg, 256) != 0
    assert bad == 8, bad
    print("sanitized host code ok")
''')


def test_host_code_under_asan_ubsan(san_lib):
    lib, asan = san_lib
    env = dict(os.environ, LD_PRELOAD=asan, ASAN_OPTIONS="detect_leaks=0:abort_on_error=1",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    r = subprocess.run([sys.executable, "-c", DRIVER, lib, PKG, os.path.join(ROOT, "oracle")], env=env,
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, (r.stdout[-3000:], r.stderr[-6000:])
    assert "sanitized host code ok" in r.stdout
