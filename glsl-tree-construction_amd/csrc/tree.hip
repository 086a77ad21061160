// tree.hip -- Karras binary radix tree, edge count and octree emission
// (plain, one cloud per call).
//
// Reference: k_BuildRadixTree, shader/build_radix_tree.comp:71-153;
// countEdges, shader/edge_count.comp:30-41; k_MakeOctNodes and
// k_LinkLeafNodes, shader/octree.comp:126-236; root set-up src/main.rs:140-153.
// Published algorithm: Karras, "Maximizing Parallelism in the Construction of
// BVHs, Octrees, and k-d Trees" (HPG 2012).
//
// gfx950 design: one lane per radix node in grid-stride loops over a fixed
// grid, so the node count may be read on the device (no host round trip after
// unique).  The radix-tree kernel can also emit the edge counts (a parent knows
// its children's prefix lengths from their range end points).  The octree is
// one kernel: each contributing radix node writes its whole chain, each node as
// four 16 B stores, with the bottom node's children gathered by a bounded walk
// down the radix tree -- no clearing pass, no atomics, every node written once.
#include "tree_core.hpp"

namespace {

constexpr int TREE_BLOCK = 256;
constexpr int BRT_TILE = 1024;  // radix nodes per workgroup step (4 per lane)
constexpr int BRT_HALO = 128;   // codes staged in LDS on each side of the tile
// Ranges below this many keys take the LDS-only search (a probe lands at most
// BRT_LIMIT keys from its node, inside the halo).  128 against 64: uniform 2^26
// codes 0.711 -> 0.693 ms (half the long-range nodes), identical outputs; 256
// gained nothing more and 32 lost 18%.
constexpr uint32_t BRT_LIMIT = 128;
constexpr int OCT_TILE = 1024;  // radix nodes per workgroup step of the octree kernel
// Links staged on each side of the tile for the walks below its nodes.  The
// walks that leave the window read global memory instead; the kernel's time is
// the same with 128, 256 or 512 (2^26 builds: 70.6 / 69.4 / 70.9 us Gaussian,
// 1.368 / 1.372 / 1.371 ms uniform), and 128 stages the fewest bytes.
constexpr int OCT_HALO = 128;

TC_DEV uint32_t resolve_nodes(uint32_t n_brt_cap, const uint32_t* d_n_keys) {
    if (!d_n_keys) return n_brt_cap;
    const uint32_t u = *d_n_keys;
    const uint32_t m = u > 0 ? u - 1 : 0;
    return m < n_brt_cap ? m : n_brt_cap;
}

// Radix tree over tiles of BRT_TILE nodes.  The workgroup copies the codes of
// its tile plus BRT_HALO on each side into LDS; every node first runs the
// search bounded to ranges below BRT_LIMIT keys, whose probes all fall inside
// the window (LDS reads only, a lane's four nodes interleaved).  The few nodes with longer ranges
// (about 20 per tile) are collected and searched in full afterwards, each by a
// group of BRT_LONG_GROUP lanes: searched one per lane, their chains of
// dependent global probes took longer than the rest of the tile (2^26
// Gaussian-mixture build: median 12 us of a 22 us tile).
template <typename K>
__global__ __launch_bounds__(TREE_BLOCK) void k_build_radix_tree(
    const K* __restrict__ codes, uint32_t n_keys_cap, const uint32_t* __restrict__ d_n_keys,
    uint8_t* __restrict__ prefix_n, uint8_t* __restrict__ hll, uint8_t* __restrict__ hlr,
    int32_t* __restrict__ left_child, int32_t* __restrict__ parent, uint32_t* __restrict__ ec,
    uint32_t* __restrict__ epoch_seq) {
    static_assert(BRT_HALO >= BRT_LIMIT, "bounded searches must stay inside the window");
    __shared__ K s_win[BRT_TILE + 2 * BRT_HALO];
    __shared__ uint32_t s_long[BRT_TILE];
    __shared__ uint32_t s_nlong;
    // Advance the scan epoch tag for the offsets scan that follows in the build.
    if (blockIdx.x == 0 && threadIdx.x == 0) *epoch_seq = (*epoch_seq + 1u) & 0x3FFFFFFFu;
    const uint32_t nb = resolve_nodes(n_keys_cap > 0 ? n_keys_cap - 1 : 0, d_n_keys);
    const int64_t nk = (int64_t)nb + 1;
    const uint32_t tiles = ceil_div_u32(nb, BRT_TILE);
    const BrtArrays out{prefix_n, hll, hlr, left_child, parent, ec};
    for (uint32_t t = blockIdx.x; t < tiles; t += gridDim.x) {
        // Window [ws, ws + BRT_TILE + 2 BRT_HALO); slots outside [0, nk) hold
        // the all-ones sentinel ("no common prefix").
        const int64_t t0 = (int64_t)t * BRT_TILE;
        const int64_t ws = t0 - BRT_HALO;
        __syncthreads();  // the previous tile is done with the window
        if (threadIdx.x == 0) s_nlong = 0;
        stage_codes<TREE_BLOCK, BRT_TILE + 2 * BRT_HALO>(s_win, codes, ws, 0, nk, (K)~(K)0);
        __syncthreads();
        // The lane's nodes, searched together; slots past the last node
        // search the tile's first node (always a node) and are not stored.
        constexpr int NPL = BRT_TILE / TREE_BLOCK;
        uint32_t idx[NPL];
#pragma unroll
        for (int k = 0; k < NPL; ++k) {
            const int64_t i = t0 + k * TREE_BLOCK + threadIdx.x;
            idx[k] = (uint32_t)(i < nb ? i : t0);
        }
        BrtResult r[NPL];
        const uint32_t near =
            brt_nodes_near<K, BRT_LIMIT, NPL, sizeof(K) == 4>(as_lds<K>(s_win), (int32_t)ws, idx, ec != nullptr, r);
#pragma unroll
        for (int k = 0; k < NPL; ++k) {
            const int64_t i = t0 + k * TREE_BLOCK + threadIdx.x;
            if (i >= nb) continue;
            if ((near >> k) & 1u)
                brt_store(out, (uint32_t)i, r[k]);
            else
                s_long[atomicAdd(&s_nlong, 1u)] = (uint32_t)i;
        }
        __syncthreads();
        const int64_t w0 = ws > 0 ? ws : 0;
        const int64_t w1 = ws + BRT_TILE + 2 * BRT_HALO < nk ? ws + BRT_TILE + 2 * BRT_HALO : nk;
        const CodeView<K> view{codes, as_lds<K>(s_win + (w0 - ws)), (int32_t)w0, (uint32_t)(w1 - w0)};
        const uint32_t nlong = s_nlong;
        constexpr uint32_t G = BRT_LONG_GROUP;
        for (uint32_t q = threadIdx.x / G; q < nlong; q += TREE_BLOCK / G) {
            BrtResult r;
            // Every probe up to BRT_LIMIT held in the window search.
            brt_node_group<K, G>(view, (uint32_t)nk, s_long[q], ec != nullptr, &r, (uint32_t)__builtin_ctz(BRT_LIMIT) + 1u);
            if (threadIdx.x % G == 0) brt_store(out, s_long[q], r);
        }
    }
}

__global__ __launch_bounds__(TREE_BLOCK) void k_edge_count(const uint8_t* __restrict__ prefix_n,
                                                           const int32_t* __restrict__ parent,
                                                           uint32_t* __restrict__ ec, uint32_t n_brt_cap,
                                                           const uint32_t* __restrict__ d_n_keys) {
    const uint32_t nb = resolve_nodes(n_brt_cap, d_n_keys);
    const uint32_t stride = gridDim.x * TREE_BLOCK;
    for (uint32_t i = blockIdx.x * TREE_BLOCK + threadIdx.x; i < nb; i += stride) {
        uint32_t p = (uint32_t)parent[i];
        p = p < nb ? p : 0u;  // parent indices of a valid tree are < nb
        ec[i] = i == 0 ? 1u : (uint32_t)(prefix_n[i] / 3 - prefix_n[p] / 3);
    }
}

template <typename K>
// Eight workgroups per CU resident (61 VGPRs, nothing spilled): with fewer,
// a grid of 8 per CU left the remaining workgroups to run after the rest
// (measured at 6 per CU: 1.78 against 1.52 ms on 65 M radix nodes).
__global__ __launch_bounds__(TREE_BLOCK, 8) void k_octree_build(
    tc_oct_node* __restrict__ nodes, uint32_t node_cap, const uint32_t* __restrict__ offsets,
    const uint32_t* __restrict__ ec, const K* __restrict__ codes, const uint8_t* __restrict__ prefix_n,
    const int32_t* __restrict__ left_child, const uint8_t* __restrict__ hll, const uint8_t* __restrict__ hlr,
    uint32_t n_brt_cap, const uint32_t* __restrict__ d_n_keys, int link_leaves, float min_coord, float range) {
    constexpr int W = OCT_TILE + 2 * OCT_HALO;
    __shared__ int32_t s_left[W];
    __shared__ uint16_t s_meta[W];
    __shared__ K s_code[W];
    __shared__ uint32_t s_off[W];
    __shared__ uint16_t s_list[OCT_TILE];  // emitting nodes, tile-relative
    __shared__ uint32_t s_tmp[TREE_BLOCK / 64 + 2];
    const uint32_t nb = resolve_nodes(n_brt_cap, d_n_keys);
    if (nb == 0) return;
    const int root_level = prefix_n[0] / 3;
    const uint32_t tiles = ceil_div_u32(nb, OCT_TILE);
    for (uint32_t t = blockIdx.x; t < tiles; t += gridDim.x) {
        // Stage the tree arrays of the tile and a halo: the walks below a node
        // stay inside its key range, so nearly all of their reads hit LDS.
        const int64_t t0 = (int64_t)t * OCT_TILE;
        const int64_t w0 = t0 > OCT_HALO ? t0 - OCT_HALO : 0;
        const int64_t w1 = t0 + OCT_TILE + OCT_HALO < nb ? t0 + OCT_TILE + OCT_HALO : nb;
        __syncthreads();
        stage_tree<TREE_BLOCK, W, K>(s_left, s_meta, s_code, s_off, left_child, ec, hll, hlr, prefix_n, codes,
                                     offsets, w0, w1);
        __syncthreads();
        // Only nodes with ec > 0 emit anything (typically well under half):
        // compact them so that every lane of a wave has one.  Thread t looks
        // at nodes 4t .. 4t + 3 of the tile: one block scan.
        constexpr int NPT = OCT_TILE / TREE_BLOCK;
        uint32_t flags = 0;
#pragma unroll
        for (int k = 0; k < NPT; ++k) {
            const int64_t i = t0 + NPT * threadIdx.x + k;
            flags |= ((i < nb && (s_meta[i - w0] & 255u) != 0u) ? 1u : 0u) << k;
        }
        uint32_t count;
        uint32_t rank = block_exclusive_scan<TREE_BLOCK>(__popc(flags), s_tmp, &count);
#pragma unroll
        for (int k = 0; k < NPT; ++k)
            if ((flags >> k) & 1u) s_list[rank++] = (uint16_t)(NPT * threadIdx.x + k);
        __syncthreads();
        const TreeView<K> lv{left_child, hll, hlr, ec, prefix_n, codes, offsets, as_lds<int32_t>(s_left),
                             as_lds<uint16_t>(s_meta), as_lds<K>(s_code), as_lds<uint32_t>(s_off), (int32_t)w0,
                             (uint32_t)(w1 - w0)};
#pragma unroll 1
        for (uint32_t q = threadIdx.x; q < count; q += TREE_BLOCK)
            octree_build_node<K>(nodes, node_cap, 0u, lv, (uint32_t)t0 + s_list[q], nb, root_level, link_leaves,
                                 min_coord, range);
    }
}

int tree_grid(tc_ctx* ctx, uint32_t n) { return tc_persistent_grid(ctx, n, TREE_BLOCK); }

template <typename K>
int brt_impl(tc_ctx* ctx, const K* codes, uint32_t n_keys, const uint32_t* d_n_keys, uint8_t* prefix_n,
             uint8_t* hll, uint8_t* hlr, int32_t* left, int32_t* parent, uint32_t* ec, hipStream_t s) {
    if (!ctx) return TC_ERR_INVALID;
    if (n_keys > 1 && (!codes || !prefix_n || !hll || !hlr || !left || !parent)) return TC_ERR_INVALID;
    if (n_keys > TC_MAX_SORT_N) return TC_ERR_TOO_LARGE;
    if (n_keys < 2) return TC_OK;
    ProfScope prof(ctx, s, K_BRT);
    // Four times the resident workgroups (8 per CU): on 4,761 tiles a grid of
    // 8 per CU runs whole tiles in 3 rounds where 2.3 suffice (0.080 -> 0.061
    // ms at the 2^26 Gaussian-mixture build).
    const int grid = tc_persistent_grid(ctx, ceil_div_u32(n_keys - 1, BRT_TILE), 1, 32);
    hipLaunchKernelGGL(k_build_radix_tree<K>, dim3(grid), dim3(TREE_BLOCK), 0, s, codes,
                       n_keys, d_n_keys, prefix_n, hll, hlr, left, parent, ec, tc_scan_epoch(ctx));
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
    if (!left || !hll || !hlr) return TC_ERR_INVALID;
    if (reinterpret_cast<uintptr_t>(nodes) & 15u) return TC_ERR_INVALID;
    const int grid = tc_persistent_grid(ctx, ceil_div_u32(n_brt, OCT_TILE), 1, 32);  // 4 x resident
    ProfScope prof(ctx, s, K_OCT_EMIT);
    hipLaunchKernelGGL(k_octree_build<K>, dim3(grid), dim3(TREE_BLOCK), 0, s, nodes, cap, offsets, ec, codes,
                       prefix_n, left, hll, hlr, n_brt, d_n_keys, link_leaves, min_coord, range);
    TC_LAUNCH_CHECK();
    return TC_OK;
}

}  // namespace

extern "C" int tc_build_radix_tree_u32(tc_ctx* ctx, const uint32_t* codes, uint32_t n_keys,
                                       const uint32_t* d_n_keys, uint8_t* prefix_n, uint8_t* hll, uint8_t* hlr,
                                       int32_t* left, int32_t* parent, tc_stream stream) {
    return brt_impl<uint32_t>(ctx, codes, n_keys, d_n_keys, prefix_n, hll, hlr, left, parent, nullptr,
                              as_stream(stream));
}
extern "C" int tc_build_radix_tree_u64(tc_ctx* ctx, const uint64_t* codes, uint32_t n_keys,
                                       const uint32_t* d_n_keys, uint8_t* prefix_n, uint8_t* hll, uint8_t* hlr,
                                       int32_t* left, int32_t* parent, tc_stream stream) {
    return brt_impl<uint64_t>(ctx, codes, n_keys, d_n_keys, prefix_n, hll, hlr, left, parent, nullptr,
                              as_stream(stream));
}

extern "C" int tc_build_radix_tree_ec_u32(tc_ctx* ctx, const uint32_t* codes, uint32_t n_keys,
                                          const uint32_t* d_n_keys, uint8_t* prefix_n, uint8_t* hll,
                                          uint8_t* hlr, int32_t* left, int32_t* parent, uint32_t* edge_count,
                                          tc_stream stream) {
    if (n_keys > 1 && !edge_count) return TC_ERR_INVALID;
    return brt_impl<uint32_t>(ctx, codes, n_keys, d_n_keys, prefix_n, hll, hlr, left, parent, edge_count,
                              as_stream(stream));
}
extern "C" int tc_build_radix_tree_ec_u64(tc_ctx* ctx, const uint64_t* codes, uint32_t n_keys,
                                          const uint32_t* d_n_keys, uint8_t* prefix_n, uint8_t* hll,
                                          uint8_t* hlr, int32_t* left, int32_t* parent, uint32_t* edge_count,
                                          tc_stream stream) {
    if (n_keys > 1 && !edge_count) return TC_ERR_INVALID;
    return brt_impl<uint64_t>(ctx, codes, n_keys, d_n_keys, prefix_n, hll, hlr, left, parent, edge_count,
                              as_stream(stream));
}

extern "C" int tc_edge_count(tc_ctx* ctx, const uint8_t* prefix_n, const int32_t* parent, uint32_t* ec,
                             uint32_t n_brt, const uint32_t* d_n_keys, tc_stream stream) {
    if (!ctx || (n_brt && (!prefix_n || !parent || !ec))) return TC_ERR_INVALID;
    if (n_brt > TC_MAX_SORT_N) return TC_ERR_TOO_LARGE;
    if (n_brt == 0) return TC_OK;
    ProfScope prof(ctx, as_stream(stream), K_EDGE);
    hipLaunchKernelGGL(k_edge_count, dim3(tree_grid(ctx, n_brt)), dim3(TREE_BLOCK), 0, as_stream(stream), prefix_n,
                       parent, ec, n_brt, d_n_keys);
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
