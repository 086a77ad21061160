"""CPU-side checks of the product library: it loads, exports every entry point
that include/tc.h declares, and its host-only functions (generators, scalar
decode, node bound) behave.  No kernel is launched here."""
import ctypes
import os
import re

import numpy as np
import pytest

from conftest import ROOT


def declared_functions():
    names = []
    for hdr in os.listdir(os.path.join(ROOT, "include")):
        if not hdr.endswith(".h"):
            continue
        text = open(os.path.join(ROOT, "include", hdr)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names += re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s+(tc_\w+)\s*\(", text, flags=re.M)
    return sorted(set(names))


def test_header_declares_the_stage_chain():
    names = declared_functions()
    for stage in ("tc_morton30", "tc_global_histogram_u32", "tc_binning_pass_u32", "tc_sort_keys_u32",
                  "tc_unique_u32", "tc_build_radix_tree_u32", "tc_edge_count", "tc_exclusive_scan_u32",
                  "tc_build_octree_u32", "tc_build_tree"):
        assert stage in names


def test_library_exports_every_declared_symbol(tclib):
    lib = ctypes.CDLL(tclib.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_python_bindings_cover_the_header(tclib):
    from tc import _lib

    tclib.lib()
    bound = set(_lib.SIGNATURES)
    missing = [n for n in declared_functions() if n not in bound]
    assert not missing, missing


def test_sort_plan_list(tclib):
    """The launch shapes and digit plans the tuner chooses among (and the GPU
    launch-shape tests cover)."""
    from test_gpu_parity import SORT_PLANS

    assert tclib.tune.built_shapes() == SORT_PLANS


def test_status_strings(tclib):
    L = tclib.lib()
    assert L.tc_status_string(0) == b"ok"
    assert L.tc_status_string(-3).startswith(b"size above")
    assert L.tc_version().startswith(b"tc ")


def test_create_without_gpu_reports_no_device(tclib):
    import torch

    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    h = ctypes.c_void_p()
    assert tclib.lib().tc_create(ctypes.byref(h), 0) == tclib._lib.TC_ERR_NO_DEVICE


def test_generators_deterministic_and_chunking_independent(tclib):
    a = tclib.gen_uniform(100_000, 7)
    b = tclib.gen_uniform(1000, 7)
    assert np.array_equal(a[:1000], b)
    assert np.all((a[:, :3] >= 0) & (a[:, :3] < 1024)) and np.all(a[:, 3] == 1)
    # 24-bit grid: exact multiples of 2^-14 in [0, 1024)
    assert np.all((a[:, :3] * 16384) == np.floor(a[:, :3] * 16384))
    g1 = tclib.gen_gaussian_mixture(200_000, 9, 64, 8.0)
    g2 = tclib.gen_gaussian_mixture(500, 9, 64, 8.0)
    assert np.array_equal(g1[:500], g2)
    assert np.all((g1[:, :3] >= 0) & (g1[:, :3] < 1024)) and np.all(g1[:, 3] == 1)
    assert not np.array_equal(tclib.gen_uniform(10, 1), tclib.gen_uniform(10, 2))


def test_gaussian_mixture_has_heavy_duplicates(tclib, oracle):
    """The 2^26 config's distribution (64 centres, sigma 8) gives ~7% unique codes;
    a scaled-down sample keeps clusters tight enough to collide."""
    g = tclib.gen_gaussian_mixture(1 << 20, 5, 64, 8.0)
    codes = oracle.compute_morton(g, 0.0, 1024.0)
    u = len(np.unique(codes))
    assert u < 0.9 * len(codes)


def test_host_decode_matches_oracle(tclib, oracle):
    r = np.random.default_rng(3)
    for c in r.integers(0, 1 << 30, size=300):
        assert np.array_equal(tclib.morton32_to_xyz(int(c), -7.5, 33.0), oracle.morton32_to_xyz(int(c), -7.5, 33.0))
    for c in r.integers(0, 1 << 63, size=300, dtype=np.uint64):
        assert np.array_equal(tclib.morton64_to_xyz(int(c), 1.0, 3.0), oracle.morton64_to_xyz(int(c), 1.0, 3.0))


def test_node_bound_matches_oracle_and_holds(tclib, oracle):
    for n in (0, 1, 2, 3, 1000, 1 << 20, 1 << 26):
        for bits in (30, 63):
            assert tclib.octree_node_bound(n, bits) == oracle.octree_node_bound(n, bits)
    g = np.load(os.path.join(ROOT, "tests", "golden", "gauss_4096_seed12.npz"))
    assert len(g["nodes"]) // 64 <= tclib.octree_node_bound(len(g["unique"]), 30)
