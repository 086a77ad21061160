"""tc_sort_keys_u32_noalloc (the sort whose scratch is a static device array)
against tc_sort_keys_u32 on the same keys, through the C ABI: bit-exact, on two
calls in a row with the second one smaller, at the size limit (2^24 keys) and
with a launch shape whose tiles are too small for the static regions."""
import ctypes as C

import numpy as np
import pytest

from gpu_util import dev, empty, host

pytestmark = pytest.mark.gpu


def sort_both(L, ctx, keys):
    import torch

    n = len(keys)
    a, b = dev(keys), dev(keys)
    alt_a, alt_b = empty(n, np.uint32), empty(n, np.uint32)
    assert L.tc_sort_keys_u32(ctx.handle, C.c_void_p(a.data_ptr()), C.c_void_p(alt_a.data_ptr()), n, None) == 0
    assert L.tc_sort_keys_u32_noalloc(ctx.handle, C.c_void_p(b.data_ptr()), C.c_void_p(alt_b.data_ptr()), n,
                                      None) == 0
    torch.cuda.synchronize()
    return host(a, np.uint32, n), host(b, np.uint32, n)


@pytest.mark.parametrize("shape", [None, (256, 8)])
def test_noalloc_sort_matches_allocating_sort(gpu_ctx, tclib, shape):
    import tc

    L = tc.lib()
    r = np.random.default_rng(77)
    saved = gpu_ctx.sort_params()
    try:
        if shape:
            gpu_ctx.set_sort_params(*shape)
        for n, hi in ((1 << 24, 1 << 32), (1_000_003, 1 << 30), (5, 1 << 32)):
            keys = r.integers(0, hi, size=n, dtype=np.uint64).astype(np.uint32)
            want, got = sort_both(L, gpu_ctx, keys)
            assert np.array_equal(got, want), (shape, n)
            assert np.array_equal(got, np.sort(keys)), (shape, n)
    finally:
        gpu_ctx.set_sort_params(*saved)


def test_noalloc_sort_rejects_sizes_above_its_limit(gpu_ctx):
    import tc

    L = tc.lib()
    n = (1 << 24) + 1
    k = empty(16, np.uint32)
    # Rejected before any device work (the buffers are not touched).
    assert L.tc_sort_keys_u32_noalloc(gpu_ctx.handle, C.c_void_p(k.data_ptr()), C.c_void_p(k.data_ptr()), n,
                                      None) == -3
    assert L.tc_sort_keys_u32_noalloc(gpu_ctx.handle, None, None, 0, None) == 0
