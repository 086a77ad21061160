// tree_core.hpp -- per-node device functions of the radix tree and octree
// stages, shared by the plain kernels (tree.hip) and the segmented batch
// kernels (batch.hip).  Every function works on one cloud's arrays (pointers
// already offset to the cloud) with cloud-local indices; `node_base` is the
// global index of the cloud's first octree node, subtracted from stored links
// so a cloud's octree is the same whether built alone or inside a batch.
#pragma once

#include "tc_internal.hpp"

template <typename K>
struct Delta {
    const K* __restrict__ c;
    int64_t nk;
    K ci;
    // Common-prefix length over the code width, -1 outside [0, nk-1].
    TC_DEV int operator()(int64_t j) const {
        if (j < 0 || j > nk - 1) return -1;
        return CodeTraits<K>::delta_bits(ci ^ c[j]);
    }
};

TC_DEV int log2_ceil_u32(uint32_t x) {
    const int fl = 31 - __clz(x);
    return fl + (((uint32_t)1 << fl) < x ? 1 : 0);
}

// k_BuildRadixTree, shader/build_radix_tree.comp:71-146 (Karras 2012), for
// internal node i of a cloud of nk strictly increasing codes.
//
// When ec is non-NULL the node also writes the edge count of its internal
// children (countEdges, shader/edge_count.comp:36-38): a child covering
// [lo, hi] has prefix length lcp(codes[lo], codes[hi]), so its count
// prefix_n[child]/3 - prefix_n[i]/3 is known here without reading the child's
// output; node 0 writes the root's count, 1.
template <typename K>
TC_DEV void brt_node(const K* __restrict__ codes, int64_t nk, int64_t i, uint8_t* __restrict__ prefix_n,
                     uint8_t* __restrict__ hll, uint8_t* __restrict__ hlr, int32_t* __restrict__ left_child,
                     int32_t* __restrict__ parent, uint32_t* __restrict__ ec = nullptr) {
    Delta<K> delta{codes, nk, codes[i]};
    // Direction (Karras 2012 eq. 1); ties (duplicate codes only) take +1.
    const int d = (delta(i + 1) - delta(i - 1)) >= 0 ? 1 : -1;
    const int dmin = delta(i - d);
    int64_t lmax = 2;
    while (delta(i + lmax * d) > dmin) lmax *= 2;
    int64_t l = 0;
    for (int64_t t = lmax / 2; t >= 1; t /= 2)
        if (delta(i + (l + t) * d) > dmin) l += t;
    const int64_t j = i + l * d;
    const int dnode = delta(j);
    prefix_n[i] = (uint8_t)dnode;
    int64_t s = 0;
    const uint32_t maxdiv = l <= 1 ? 1u : (uint32_t)1 << log2_ceil_u32((uint32_t)l);
    for (uint32_t div = 2, t = ((uint32_t)l + 1) / 2; div <= maxdiv; div <<= 1, t = ((uint32_t)l + div - 1) / div)
        if (delta(i + (s + (int64_t)t) * d) > dnode) s += t;
    const int64_t gamma = i + s * d + (d < 0 ? d : 0);
    left_child[i] = (int32_t)gamma;
    const bool leaf_l = (i < j ? i : j) == gamma;
    const bool leaf_r = (i > j ? i : j) == gamma + 1;
    hll[i] = leaf_l;
    hlr[i] = leaf_r;
    if (!leaf_l) parent[gamma] = (int32_t)i;
    if (!leaf_r) parent[gamma + 1] = (int32_t)i;
    if (i == 0) parent[0] = 0;  // the reference's zero-initialised slot (src/main.rs:103)
    if (ec) {
        const int64_t lo = i < j ? i : j, hi = i < j ? j : i;
        const int level = dnode / 3;
        if (!leaf_l) {
            Delta<K> dl{codes, nk, codes[lo]};
            ec[gamma] = (uint32_t)(dl(gamma) / 3 - level);
        }
        if (!leaf_r) {
            Delta<K> dr{codes, nk, codes[gamma + 1]};
            ec[gamma + 1] = (uint32_t)(dr(hi) / 3 - level);
        }
        if (i == 0) ec[0] = 1u;
    }
}

// Nearest radix ancestor-or-self with a non-zero edge count (at most three
// steps in a valid tree: along a run of zero counts the prefix length falls by
// one per step while its level stays the same; the root's count is 1).
// Bounded and range-checked so malformed input cannot loop or fault.
TC_DEV uint32_t contributing(uint32_t i, uint32_t nb, const uint32_t* __restrict__ ec,
                             const int32_t* __restrict__ parent) {
    for (int step = 0; step < 64 && ec[i] == 0; ++step) {
        const uint32_t p = (uint32_t)parent[i];
        i = p < nb ? p : 0u;
    }
    return i;
}

template <typename K>
TC_DEV uint32_t child_slot(K code, int level) {
    constexpr int BITS = CodeTraits<K>::BITS;
    return (uint32_t)((code >> (BITS - 3 * level)) & 7u);
}

// One radix node's share of the octree (k_MakeOctNodes + k_LinkLeafNodes,
// shader/octree.comp:126-222), written without atomics: radix node a with
// ec[a] > 0 owns the chain offsets[a] .. offsets[a]+ec[a]-1 (bottom, i.e.
// deepest, first).  Chain node j > 0 has one child, node j-1.  The bottom node's
// children are found by a bounded walk down the radix tree from a through
// nodes with ec == 0 (they share a's level): each internal node met with
// ec > 0 contributes its chain top, each radix leaf met contributes itself
// (SetLeaf: children[slot] = leaf, leaf mask left as cleared, i.e. 0).  Every
// child of the bottom node sits one level below it, so its slot is the next 3
// bits of its code.  The walk replaces the upward ancestor search of
// shader/octree.comp:157-165 and the racy mask updates (each node is written
// once, whole, by one lane).  offsets[] hold global node indices; stored links
// are local (minus node_base).
template <typename K>
TC_DEV void octree_build_node(tc_oct_node* __restrict__ nodes, uint32_t node_cap, uint32_t node_base,
                              const uint32_t* __restrict__ offsets, const uint32_t* __restrict__ ec,
                              const K* __restrict__ codes, const uint8_t* __restrict__ prefix_n,
                              const int32_t* __restrict__ left_child, const uint8_t* __restrict__ hll,
                              const uint8_t* __restrict__ hlr, uint32_t a, uint32_t nb, int root_level,
                              int link_leaves, float min_coord, float range) {
    constexpr int BITS = CodeTraits<K>::BITS;
    const uint32_t cnt = ec[a];
    if (cnt == 0) return;
    const uint32_t off = offsets[a];
    const K code = codes[a];
    const int bottom = prefix_n[a] / 3;

    // Children of the bottom node.
    int32_t bch[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    int32_t bmask = 0;
    {
        uint32_t stack[8];
        int sp = 0;
        stack[sp++] = a;
        int visits = 0;
        while (sp > 0 && visits < 16) {
            const uint32_t x = stack[--sp];
            ++visits;
            const int32_t lc = left_child[x];
            const uint8_t leaf[2] = {hll[x], hlr[x]};
#pragma unroll
            for (int side = 0; side < 2; ++side) {
                const uint32_t c = (uint32_t)lc + side;
                if (c > nb) continue;  // a valid tree's children are in [0, nb]
                if (leaf[side]) {
                    if (link_leaves) bch[child_slot<K>(codes[c], bottom + 1)] = (int32_t)c;
                } else if (c < nb) {
                    const uint32_t e = ec[c];
                    if (e > 0) {
                        const uint32_t slot = child_slot<K>(codes[c], bottom + 1);
                        bch[slot] = (int32_t)(offsets[c] + e - 1 - node_base);
                        bmask |= 1 << slot;
                    } else if (sp < 8) {
                        stack[sp++] = c;
                    }
                }
            }
        }
    }

    for (uint32_t j = 0; j < cnt; ++j) {
        const uint32_t idx = off + j;
        if (idx >= node_cap) break;
        const int level = bottom - (int)j;
        const int sh = BITS - 3 * level;
        float c4[4];
        CodeTraits<K>::decode((code >> sh) << sh, min_coord, range, c4);
        const float cell = __fdiv_rn(range, (float)(1u << (level - root_level)));
        int32_t ch[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        int32_t mask = 0;
        if (j == 0) {
#pragma unroll
            for (int s = 0; s < 8; ++s) ch[s] = bch[s];
            mask = bmask;
        } else {  // chain-internal child: the node one level deeper
            const uint32_t slot = child_slot<K>(code, level + 1);
            ch[slot] = (int32_t)(idx - 1 - node_base);
            mask = 1 << slot;
        }
        int4* dst = reinterpret_cast<int4*>(nodes + idx);
        dst[0] = make_int4(ch[0], ch[1], ch[2], ch[3]);
        dst[1] = make_int4(ch[4], ch[5], ch[6], ch[7]);
        dst[2] = make_int4(__float_as_int(c4[0]), __float_as_int(c4[1]), __float_as_int(c4[2]), __float_as_int(c4[3]));
        dst[3] = make_int4(__float_as_int(cell), mask, 0, 0);
    }
}
