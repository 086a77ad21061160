// tree.hip -- Karras binary radix tree, edge count and octree emission.
//
// Reference: k_BuildRadixTree, shader/build_radix_tree.comp:71-153;
// countEdges, shader/edge_count.comp:30-41; k_MakeOctNodes and
// k_LinkLeafNodes, shader/octree.comp:126-236; root set-up src/main.rs:140-153.
// Published algorithm: Karras, "Maximizing Parallelism in the Construction of
// BVHs, Octrees, and k-d Trees" (HPG 2012).
//
// gfx950 design: one lane per radix node in grid-stride loops over a fixed
// grid, so the node count may be read on the device (no host round trip after
// unique).  The octree is built by two kernels: `emit` writes every node of a
// radix node's chain as four 16 B stores (including the chain's internal child
// links), then `link` attaches each chain top and each leaf to the bottom node
// of the nearest contributing ancestor, with atomic-or on the child mask.  The
// two-kernel split replaces a clearing pass over the node array and the
// reference's racy read-modify-write of the masks.
#include "tc_internal.hpp"

namespace {

constexpr int TREE_BLOCK = 256;

TC_DEV uint32_t resolve_keys(uint32_t cap, const uint32_t* d_n_keys) {
    if (!d_n_keys) return cap;
    const uint32_t u = *d_n_keys;
    return u < cap ? u : cap;
}

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

template <typename K>
__global__ __launch_bounds__(TREE_BLOCK) void k_build_radix_tree(
    const K* __restrict__ codes, uint32_t n_keys_cap, const uint32_t* __restrict__ d_n_keys,
    uint8_t* __restrict__ prefix_n, uint8_t* __restrict__ hll, uint8_t* __restrict__ hlr,
    int32_t* __restrict__ left_child, int32_t* __restrict__ parent) {
    const uint32_t nk = resolve_keys(n_keys_cap, d_n_keys);
    const uint32_t nb = nk > 0 ? nk - 1 : 0;
    const uint32_t stride = gridDim.x * TREE_BLOCK;
    for (uint32_t ui = blockIdx.x * TREE_BLOCK + threadIdx.x; ui < nb; ui += stride) {
        const int64_t i = ui;
        Delta<K> delta{codes, (int64_t)nk, codes[i]};
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
        for (uint32_t div = 2, t = ((uint32_t)l + 1) / 2; div <= maxdiv;
             div <<= 1, t = ((uint32_t)l + div - 1) / div)
            if (delta(i + (s + (int64_t)t) * d) > dnode) s += t;
        const int64_t gamma = i + s * d + (d < 0 ? d : 0);
        left_child[i] = (int32_t)gamma;
        const bool leaf_l = (i < j ? i : j) == gamma;
        const bool leaf_r = (i > j ? i : j) == gamma + 1;
        hll[i] = leaf_l;
        hlr[i] = leaf_r;
        if (!leaf_l) parent[gamma] = (int32_t)i;
        if (!leaf_r) parent[gamma + 1] = (int32_t)i;
        if (i == 0) parent[0] = 0;
    }
}

__global__ __launch_bounds__(TREE_BLOCK) void k_edge_count(const uint8_t* __restrict__ prefix_n,
                                                           const int32_t* __restrict__ parent,
                                                           uint32_t* __restrict__ ec, uint32_t n_brt_cap,
                                                           const uint32_t* __restrict__ d_n_keys) {
    uint32_t nb = n_brt_cap;
    if (d_n_keys) {
        const uint32_t u = *d_n_keys;
        const uint32_t m = u > 0 ? u - 1 : 0;
        nb = m < n_brt_cap ? m : n_brt_cap;
    }
    const uint32_t stride = gridDim.x * TREE_BLOCK;
    for (uint32_t i = blockIdx.x * TREE_BLOCK + threadIdx.x; i < nb; i += stride)
    {
        uint32_t p = (uint32_t)parent[i];
        p = p < nb ? p : 0u;  // parent indices of a valid tree are < nb
        ec[i] = i == 0 ? 1u : (uint32_t)(prefix_n[i] / 3 - prefix_n[p] / 3);
    }
}

// Nearest radix ancestor-or-self with a non-zero edge count (at most three
// steps: along a run of zero counts the prefix length falls by one per step
// while its level stays the same; the root's count is 1).
TC_DEV uint32_t contributing(uint32_t i, uint32_t nb, const uint32_t* __restrict__ ec,
                             const int32_t* __restrict__ parent) {
    // Bounded walk: a valid tree needs at most 3 steps; malformed input (e.g.
    // duplicate codes) must not loop or index out of range.
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

template <typename K>
__global__ __launch_bounds__(TREE_BLOCK) void k_octree_emit(
    tc_oct_node* __restrict__ nodes, uint32_t node_cap, const uint32_t* __restrict__ offsets,
    const uint32_t* __restrict__ ec, const K* __restrict__ codes, const uint8_t* __restrict__ prefix_n,
    uint32_t n_brt_cap, const uint32_t* __restrict__ d_n_keys, float min_coord, float range) {
    constexpr int BITS = CodeTraits<K>::BITS;
    uint32_t nb = n_brt_cap;
    if (d_n_keys) {
        const uint32_t u = *d_n_keys;
        const uint32_t m = u > 0 ? u - 1 : 0;
        nb = m < n_brt_cap ? m : n_brt_cap;
    }
    if (nb == 0) return;
    const int root_level = prefix_n[0] / 3;
    const uint32_t stride = gridDim.x * TREE_BLOCK;
    for (uint32_t i = blockIdx.x * TREE_BLOCK + threadIdx.x; i < nb; i += stride) {
        const uint32_t cnt = ec[i];
        if (cnt == 0) continue;
        const uint32_t off = offsets[i];
        const K code = codes[i];
        const int bottom = prefix_n[i] / 3;
        for (uint32_t j = 0; j < cnt; ++j) {
            const uint32_t idx = off + j;
            if (idx >= node_cap) break;
            const int level = bottom - (int)j;
            const int sh = BITS - 3 * level;
            float c[4];
            CodeTraits<K>::decode((code >> sh) << sh, min_coord, range, c);
            const float cell = __fdiv_rn(range, (float)(1u << (level - root_level)));
            int32_t ch[8] = {0, 0, 0, 0, 0, 0, 0, 0};
            int32_t mask = 0;
            if (j > 0) {  // chain-internal child: the node one level deeper
                const uint32_t slot = child_slot<K>(code, level + 1);
                ch[slot] = (int32_t)(idx - 1);
                mask = 1 << slot;
            }
            int4* dst = reinterpret_cast<int4*>(nodes + idx);
            dst[0] = make_int4(ch[0], ch[1], ch[2], ch[3]);
            dst[1] = make_int4(ch[4], ch[5], ch[6], ch[7]);
            dst[2] = make_int4(__float_as_int(c[0]), __float_as_int(c[1]), __float_as_int(c[2]),
                               __float_as_int(c[3]));
            dst[3] = make_int4(__float_as_int(cell), mask, 0, 0);
        }
    }
}

template <typename K>
__global__ __launch_bounds__(TREE_BLOCK) void k_octree_link(
    tc_oct_node* __restrict__ nodes, uint32_t node_cap, const uint32_t* __restrict__ offsets,
    const uint32_t* __restrict__ ec, const K* __restrict__ codes, const uint8_t* __restrict__ prefix_n,
    const int32_t* __restrict__ parent, const int32_t* __restrict__ left_child,
    const uint8_t* __restrict__ hll, const uint8_t* __restrict__ hlr, uint32_t n_brt_cap,
    const uint32_t* __restrict__ d_n_keys, int link_leaves) {
    uint32_t nb = n_brt_cap;
    if (d_n_keys) {
        const uint32_t u = *d_n_keys;
        const uint32_t m = u > 0 ? u - 1 : 0;
        nb = m < n_brt_cap ? m : n_brt_cap;
    }
    const uint32_t stride = gridDim.x * TREE_BLOCK;
    for (uint32_t i = blockIdx.x * TREE_BLOCK + threadIdx.x; i < nb; i += stride) {
        const uint32_t cnt = ec[i];
        if (i > 0 && cnt > 0) {
            // Chain top -> bottom node of the nearest contributing ancestor.
            const int top_level = prefix_n[i] / 3 - (int)cnt + 1;
            const uint32_t a = contributing((uint32_t)parent[i] < nb ? (uint32_t)parent[i] : 0u, nb, ec, parent);
            const uint32_t slot = child_slot<K>(codes[i], top_level);
            const uint32_t pidx = offsets[a];
            const uint32_t top = offsets[i] + cnt - 1;
            if (pidx < node_cap) {
                nodes[pidx].children[slot] = (int32_t)top;
                atomicOr(&nodes[pidx].child_node_mask, 1 << slot);
            }
        }
        if (link_leaves) {
            const int leaf_level = prefix_n[i] / 3 + 1;
            const uint8_t has[2] = {hll[i], hlr[i]};
#pragma unroll
            for (int side = 0; side < 2; ++side) {
                if (!has[side]) continue;
                const int32_t leaf = left_child[i] + side;
                const uint32_t slot = child_slot<K>(codes[leaf], leaf_level);
                const uint32_t a = contributing(i, nb, ec, parent);
                const uint32_t pidx = offsets[a];
                if (pidx < node_cap) {
                    nodes[pidx].children[slot] = leaf;
                    // SetLeaf, shader/octree.comp:27-30: clears the leaf-mask bit.
                    atomicAnd(&nodes[pidx].child_leaf_mask, ~(1 << slot));
                }
            }
        }
    }
}

int tree_grid(tc_ctx* ctx, uint32_t n) { return tc_persistent_grid(ctx, n, TREE_BLOCK); }

template <typename K>
int brt_impl(tc_ctx* ctx, const K* codes, uint32_t n_keys, const uint32_t* d_n_keys, uint8_t* prefix_n,
             uint8_t* hll, uint8_t* hlr, int32_t* left, int32_t* parent, hipStream_t s) {
    if (!ctx) return TC_ERR_INVALID;
    if (n_keys > 1 && (!codes || !prefix_n || !hll || !hlr || !left || !parent)) return TC_ERR_INVALID;
    if (n_keys > TC_MAX_SORT_N) return TC_ERR_TOO_LARGE;
    if (n_keys < 2) return TC_OK;
    ProfScope prof(ctx, s, K_BRT);
    hipLaunchKernelGGL(k_build_radix_tree<K>, dim3(tree_grid(ctx, n_keys - 1)), dim3(TREE_BLOCK), 0, s,
                       codes, n_keys, d_n_keys, prefix_n, hll, hlr, left, parent);
    TC_LAUNCH_CHECK();
    return TC_OK;
}

template <typename K>
int octree_impl(tc_ctx* ctx, tc_oct_node* nodes, uint32_t cap, const uint32_t* offsets, const uint32_t* ec,
                const K* codes, const uint8_t* prefix_n, const int32_t* parent, const int32_t* left,
                const uint8_t* hll, const uint8_t* hlr, uint32_t n_brt, const uint32_t* d_n_keys,
                float min_coord, float range, int link_leaves, hipStream_t s) {
    if (!ctx) return TC_ERR_INVALID;
    if (n_brt > TC_MAX_SORT_N) return TC_ERR_TOO_LARGE;
    if (n_brt == 0) return TC_OK;
    if (!nodes || !offsets || !ec || !codes || !prefix_n || !parent) return TC_ERR_INVALID;
    if (link_leaves && (!left || !hll || !hlr)) return TC_ERR_INVALID;
    if (reinterpret_cast<uintptr_t>(nodes) & 15u) return TC_ERR_INVALID;
    const int grid = tree_grid(ctx, n_brt);
    {
        ProfScope prof(ctx, s, K_OCT_EMIT);
        hipLaunchKernelGGL(k_octree_emit<K>, dim3(grid), dim3(TREE_BLOCK), 0, s, nodes, cap, offsets, ec, codes,
                           prefix_n, n_brt, d_n_keys, min_coord, range);
        TC_LAUNCH_CHECK();
    }
    {
        ProfScope prof(ctx, s, K_OCT_LINK);
        hipLaunchKernelGGL(k_octree_link<K>, dim3(grid), dim3(TREE_BLOCK), 0, s, nodes, cap, offsets, ec, codes,
                           prefix_n, parent, left, hll, hlr, n_brt, d_n_keys, link_leaves);
        TC_LAUNCH_CHECK();
    }
    return TC_OK;
}

}  // namespace

extern "C" int tc_build_radix_tree_u32(tc_ctx* ctx, const uint32_t* codes, uint32_t n_keys,
                                       const uint32_t* d_n_keys, uint8_t* prefix_n, uint8_t* hll, uint8_t* hlr,
                                       int32_t* left, int32_t* parent, tc_stream stream) {
    return brt_impl<uint32_t>(ctx, codes, n_keys, d_n_keys, prefix_n, hll, hlr, left, parent, as_stream(stream));
}
extern "C" int tc_build_radix_tree_u64(tc_ctx* ctx, const uint64_t* codes, uint32_t n_keys,
                                       const uint32_t* d_n_keys, uint8_t* prefix_n, uint8_t* hll, uint8_t* hlr,
                                       int32_t* left, int32_t* parent, tc_stream stream) {
    return brt_impl<uint64_t>(ctx, codes, n_keys, d_n_keys, prefix_n, hll, hlr, left, parent, as_stream(stream));
}

extern "C" int tc_edge_count(tc_ctx* ctx, const uint8_t* prefix_n, const int32_t* parent, uint32_t* ec,
                             uint32_t n_brt, const uint32_t* d_n_keys, tc_stream stream) {
    if (!ctx || (n_brt && (!prefix_n || !parent || !ec))) return TC_ERR_INVALID;
    if (n_brt > TC_MAX_SORT_N) return TC_ERR_TOO_LARGE;
    if (n_brt == 0) return TC_OK;
    ProfScope prof(ctx, as_stream(stream), K_EDGE);
    hipLaunchKernelGGL(k_edge_count, dim3(tree_grid(ctx, n_brt)), dim3(TREE_BLOCK), 0, as_stream(stream),
                       prefix_n, parent, ec, n_brt, d_n_keys);
    TC_LAUNCH_CHECK();
    return TC_OK;
}

extern "C" int tc_build_octree_u32(tc_ctx* ctx, tc_oct_node* nodes, uint32_t cap, const uint32_t* offsets,
                                   const uint32_t* ec, const uint32_t* codes, const uint8_t* prefix_n,
                                   const int32_t* parent, const int32_t* left, const uint8_t* hll,
                                   const uint8_t* hlr, uint32_t n_brt, const uint32_t* d_n_keys,
                                   float min_coord, float range, int link_leaves, tc_stream stream) {
    return octree_impl<uint32_t>(ctx, nodes, cap, offsets, ec, codes, prefix_n, parent, left, hll, hlr, n_brt,
                                 d_n_keys, min_coord, range, link_leaves, as_stream(stream));
}
extern "C" int tc_build_octree_u64(tc_ctx* ctx, tc_oct_node* nodes, uint32_t cap, const uint32_t* offsets,
                                   const uint32_t* ec, const uint64_t* codes, const uint8_t* prefix_n,
                                   const int32_t* parent, const int32_t* left, const uint8_t* hll,
                                   const uint8_t* hlr, uint32_t n_brt, const uint32_t* d_n_keys,
                                   float min_coord, float range, int link_leaves, tc_stream stream) {
    return octree_impl<uint64_t>(ctx, nodes, cap, offsets, ec, codes, prefix_n, parent, left, hll, hlr, n_brt,
                                 d_n_keys, min_coord, range, link_leaves, as_stream(stream));
}
