// result_layout.hpp -- layout of one cloud's output buffers (tc_tree_ptrs) in a
// single allocation: each array 256-byte aligned, the octree nodes last so the
// allocation can be cut back to the actual node count once it is known.
#pragma once

#include <cstdint>

#include "../../include/tc.h"

struct TreeResultLayout {
    uint64_t unique, prefix_n, hll, hlr, left, parent, ec, offsets, nodes, total;

    static uint64_t up(uint64_t x) { return (x + 255) & ~(uint64_t)255; }

    TreeResultLayout(uint32_t n, int code_bits, uint64_t n_nodes) {
        const uint64_t m = n ? n : 1;
        const uint64_t ks = code_bits == 30 ? 4 : 8;
        uint64_t o = 0;
        unique = o; o += up(m * ks);
        prefix_n = o; o += up(m);
        hll = o; o += up(m);
        hlr = o; o += up(m);
        left = o; o += up(m * 4);
        parent = o; o += up(m * 4);
        ec = o; o += up(m * 4);
        offsets = o; o += up(m * 4);
        nodes = o; o += up(n_nodes * sizeof(tc_oct_node));
        total = o;
    }

    tc_tree_ptrs at(char* base) const {
        tc_tree_ptrs t;
        t.unique = base + unique;
        t.prefix_n = reinterpret_cast<uint8_t*>(base + prefix_n);
        t.has_leaf_left = reinterpret_cast<uint8_t*>(base + hll);
        t.has_leaf_right = reinterpret_cast<uint8_t*>(base + hlr);
        t.left_child = reinterpret_cast<int32_t*>(base + left);
        t.parent = reinterpret_cast<int32_t*>(base + parent);
        t.edge_count = reinterpret_cast<uint32_t*>(base + ec);
        t.offsets = reinterpret_cast<uint32_t*>(base + offsets);
        t.nodes = reinterpret_cast<tc_oct_node*>(base + nodes);
        return t;
    }
};
