"""Writes the committed golden fixtures under tests/golden/.

Run from the repository root:  python tests/golden/make_golden.py

1. output_txt_global_hist.json -- data from the reference's only run log
   (vulkan/radixsort-vulkan/output.txt): the 1024 global-histogram words its
   GPU produced for the keys 10,000,000 ... 1 (lines 7684-8707), the first and
   last of the 7680 input keys it printed (lines 2-7681), and its timing line.
   Read only when /root/reference is present; the JSON is committed.
2. *.npz -- small seeded inputs with every stage output of the CPU oracle
   (oracle/liboracle.so), generated with the product's seeded generators.
"""
import json
import os
import re
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "glsl-tree-construction_amd"))

REF_LOG = "/root/reference/vulkan/radixsort-vulkan/output.txt"


def write_reference_log_fixture():
    if not os.path.exists(REF_LOG):
        print("reference log not present; keeping the committed JSON")
        return
    hist = {}
    first_keys = []
    timing = None
    with open(REF_LOG) as f:
        for lineno, line in enumerate(f, 1):
            m = re.match(r"after: b_global_hist\[(\d+)\]: (\d+)", line)
            if m:
                hist[int(m.group(1))] = int(m.group(2))
                continue
            m = re.match(r"coords: (\d+)", line)
            if m:
                first_keys.append(int(m.group(1)))
                continue
            m = re.match(r"Time taken: ([0-9.]+)ms", line)
            if m:
                timing = float(m.group(1))
    assert sorted(hist) == list(range(1024))
    out = {
        "source": "vulkan/radixsort-vulkan/output.txt",
        "lines_hist": "7684-8707",
        "lines_keys": "2-7681",
        "n_keys": 10_000_000,
        "keys_rule": "input[i] = 10000000 - i (src/main.rs:189-190)",
        "printed_keys_count": len(first_keys),
        "printed_keys_first": first_keys[0],
        "printed_keys_last": first_keys[-1],
        "time_ms_4_binning_passes": timing,
        "global_hist": [hist[i] for i in range(1024)],
    }
    with open(os.path.join(HERE, "output_txt_global_hist.json"), "w") as f:
        json.dump(out, f, indent=0)
    print("wrote output_txt_global_hist.json")


def pipeline_fixture(name, pts, min_coord, range_, code_bits):
    import oracle as O

    r = O.build_tree(pts, min_coord, range_, code_bits=code_bits)
    hist = O.histogram(r["codes"])
    np.savez_compressed(
        os.path.join(HERE, name),
        points=pts, min_coord=np.float32(min_coord), range=np.float32(range_), code_bits=np.int32(code_bits),
        codes=r["codes"], hist=hist, sorted=r["sorted"], unique=r["unique"], prefix_n=r["prefix_n"],
        has_leaf_left=r["has_leaf_left"], has_leaf_right=r["has_leaf_right"], left_child=r["left_child"],
        parent=r["parent"], edge_count=r["edge_count"], offsets=r["offsets"],
        nodes=r["nodes"].view(np.uint8).reshape(-1),
    )
    print(f"wrote {name}: n={len(pts)} unique={len(r['unique'])} nodes={len(r['nodes'])}")


def pairs_fixture(name, n, seed):
    import oracle as O

    rs = np.random.default_rng(seed)
    keys = rs.integers(0, 1 << 63, size=n, dtype=np.uint64)
    keys[rs.integers(0, n, size=n // 8)] = keys[0]  # duplicates exercise stability
    vals = np.arange(n, dtype=np.uint32)
    sk, sv = O.radix_sort(keys, vals)
    np.savez_compressed(os.path.join(HERE, name), keys=keys, vals=vals, sorted_keys=sk, sorted_vals=sv,
                        hist=O.histogram(keys))
    print(f"wrote {name}")


def main():
    import tc

    write_reference_log_fixture()
    pipeline_fixture("uniform_2048_seed11.npz", tc.gen_uniform(2048, 11), 0.0, 1024.0, 30)
    pipeline_fixture("gauss_4096_seed12.npz", tc.gen_gaussian_mixture(4096, 12, 8, 3.0), 0.0, 1024.0, 30)
    pipeline_fixture("uniform63_1024_seed13.npz", tc.gen_uniform(1024, 13, 2097152.0), 0.0, 2097152.0, 63)
    pairs_fixture("pairs_u64_3000_seed14.npz", 3000, 14)


if __name__ == "__main__":
    main()
