"""Autotuner of the sort's launch shape and digit plan (tc_autotune_sort).

The binning pass is built for twelve (workgroup size, keys per lane, digit
bits) combinations: 8-bit digits, or digits of at most 11 bits (fewer passes);
which is fastest depends on the key width, the payload and the size.
``autotune`` times them all on the device and the context then picks, for each
sort, the shape measured at the nearest size (batches: at their total point
capacity).  The table can be saved as JSON and restored into another context.
"""
from __future__ import annotations

import ctypes as C
import json

from ._lib import check, lib

VARIANTS = {"u32_keys": (32, 0), "u32_pairs": (32, 1), "u64_keys": (64, 0), "u64_pairs": (64, 1)}


def built_shapes() -> list:
    """The (block, keys per thread, digit bits) launch shapes and digit plans
    the library is built with."""
    k = lib().tc_sort_plan_list(None, None, None, 0)
    b, t, d = (C.c_int * k)(), (C.c_int * k)(), (C.c_int * k)()
    lib().tc_sort_plan_list(b, t, d, k)
    return list(zip(b, t, d))


def autotune(ctx, max_n: int = 1 << 24) -> dict:
    check(lib().tc_autotune_sort(ctx.handle, max_n), "tc_autotune_sort")
    return table(ctx, max_n)


def shape(ctx, key_bits: int, pairs: bool, n: int) -> tuple:
    """(block, keys per thread, digit bits) a sort of n keys uses now."""
    b, k, d = C.c_int(), C.c_int(), C.c_int()
    check(lib().tc_tuned_sort_plan(ctx.handle, key_bits, int(bool(pairs)), n, C.byref(b), C.byref(k), C.byref(d)),
          "tc_tuned_sort_plan")
    return b.value, k.value, d.value


def table(ctx, max_n: int = 1 << 30) -> dict:
    """Shape per variant and size bucket (log2 n, even buckets up to max_n)."""
    out = {}
    for name, (bits, pairs) in VARIANTS.items():
        out[name] = {lg: list(shape(ctx, bits, pairs, 1 << lg)) for lg in range(12, 31, 2) if (1 << lg) <= max_n}
    return out


def save(ctx, path: str, max_n: int = 1 << 30):
    with open(path, "w") as f:
        json.dump({k: {str(lg): v for lg, v in t.items()} for k, t in table(ctx, max_n).items()}, f, indent=1)


def load(ctx, path: str):
    with open(path) as f:
        t = json.load(f)
    for name, entries in t.items():
        bits, pairs = VARIANTS[name]
        for lg, entry in entries.items():
            block, kpt = entry[0], entry[1]
            digit_bits = entry[2] if len(entry) > 2 else 8
            check(lib().tc_set_tuned_sort_plan(ctx.handle, bits, pairs, int(lg), block, kpt, digit_bits),
                  "tc_set_tuned_sort_plan")


def clear(ctx):
    check(lib().tc_clear_sort_tuning(ctx.handle), "tc_clear_sort_tuning")


def main(argv=None) -> int:
    """Re-time every built shape on device 0 and write the table:
    ``python -m tc.tune --out tc/tuning_mi355x.json --max-log2n 26`` (run from
    the package directory)."""
    import argparse

    from .api import Context

    ap = argparse.ArgumentParser(description=main.__doc__)
    ap.add_argument("--out", required=True)
    ap.add_argument("--max-log2n", type=int, default=26)
    args = ap.parse_args(argv)
    ctx = Context(0)
    autotune(ctx, 1 << args.max_log2n)
    save(ctx, args.out, 1 << args.max_log2n)
    print(json.dumps(table(ctx, 1 << args.max_log2n)), flush=True)
    ctx.close()
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
