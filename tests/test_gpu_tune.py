"""Autotuner: it measures, the table it leaves is made of built shapes, and every
sort it steers (any size, keys or pairs, u32 or u64) still matches the oracle."""
import numpy as np
import pytest

from gpu_util import dev, empty, host

pytestmark = pytest.mark.gpu


def test_autotune_table_and_sorts(oracle, tclib, tmp_path):
    import tc

    ctx = tc.Context(0)
    try:
        shapes = set(tc.tune.built_shapes())
        t = tc.tune.autotune(ctx, 1 << 20)
        assert set(t) == {"u32_keys", "u32_pairs", "u64_keys", "u64_pairs"}
        for entries in t.values():
            assert entries and all(tuple(v) in shapes for v in entries.values())
        r = np.random.default_rng(21)
        for n in (1000, 70_000, 1 << 20):
            keys = r.integers(0, 1 << 32, size=n, dtype=np.uint64).astype(np.uint32)
            vals = np.arange(n, dtype=np.uint32)
            k, v = dev(keys), dev(vals)
            ctx.sort(k, empty(n, np.uint32), v, empty(n, np.uint32))
            sk, sv = oracle.radix_sort(keys, vals)
            assert np.array_equal(host(k, np.uint32), sk) and np.array_equal(host(v, np.uint32), sv)
            k64 = r.integers(0, 1 << 63, size=n, dtype=np.uint64)
            k = dev(k64)
            ctx.sort(k, empty(n, np.uint64))
            assert np.array_equal(host(k, np.uint64), np.sort(k64))
        # Save / clear / load round trip.
        path = str(tmp_path / "tune.json")
        tc.tune.save(ctx, path, 1 << 20)
        tc.tune.clear(ctx)
        assert tc.tune.shape(ctx, 32, False, 1 << 20) == ctx.sort_plan()
        tc.tune.load(ctx, path)
        assert tc.tune.table(ctx, 1 << 20) == t
        # An explicit shape replaces the tuned ones.
        ctx.set_sort_params(512, 28)
        assert tc.tune.shape(ctx, 64, True, 1 << 16) == (512, 28, 8)
        ctx.set_sort_plan(1024, 16, 11)
        assert tc.tune.shape(ctx, 32, False, 1 << 16) == (1024, 16, 11)
    finally:
        ctx.close()


def test_tuner_command_line_writes_a_loadable_table(tclib, tmp_path):
    """`python -m tc.tune --out PATH --max-log2n N`: the table it writes holds
    built shapes for every variant and size bucket up to 2^N and loads back."""
    import json

    import tc

    path = str(tmp_path / "cli.json")
    assert tc.tune.main(["--out", path, "--max-log2n", "14"]) == 0
    with open(path) as f:
        t = json.load(f)
    shapes = set(tc.tune.built_shapes())
    assert set(t) == {"u32_keys", "u32_pairs", "u64_keys", "u64_pairs"}
    for entries in t.values():
        assert set(entries) == {"12", "14"}
        assert all(tuple(v) in shapes for v in entries.values())
    ctx = tc.Context(0)
    try:
        tc.tune.load(ctx, path)
        assert {k: {str(lg): v for lg, v in e.items()} for k, e in tc.tune.table(ctx, 1 << 14).items()} == t
    finally:
        ctx.close()
