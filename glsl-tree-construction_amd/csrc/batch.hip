// batch.hip -- many clouds of uneven sizes built in one persistent, segmented
// launch per stage, optionally captured as a HIP graph.
//
// No reference counterpart: the reference builds one cloud per call and creates
// a Vulkan device per stage (src/main.rs:60-170).  Each cloud's outputs here
// are bit-identical to a plain single-cloud build (tc_build_tree) of the same
// points.
//
// Layout: clouds are concatenated.  Cloud c owns points [pt_off[c],
// pt_off[c+1]); its sort tiles [first_tile[c], first_tile[c+1]) never cross
// into another cloud; its unique codes and radix-node slots are
// [ubase[c], ubase[c+1]) (U_c keys, U_c - 1 nodes, the last slot unused); its
// octree nodes are [node_base[c], node_base[c+1]).  Stored tree links are
// cloud-local indices.
//
// Every kernel reads the batch shape from device memory and uses a capacity
// grid (persistent loops), so one captured graph replays for any batch that
// fits the capacity: the host only uploads the new plan before the replay.
#include <cstring>
#include <new>
#include <vector>

#include "sort_core.hpp"
#include "tree_core.hpp"

struct tc_batch {
    tc_ctx* ctx = nullptr;
    uint32_t cap_clouds = 0, cap_points = 0, cap_tiles = 0;
    int code_bits = 30;
    int block = 256, kpt = 16;
    uint32_t tile = 4096;
    // Device plan block (one upload per batch): [desc 8][pt_off C+1][first_tile C+1][min C][range C]
    // [in_ptr C x u64][out C x tc_tree_ptrs].  desc: [0] clouds [1] tiles [2] points [3] flags.
    uint32_t* plan = nullptr;
    size_t plan_words = 0;
    uint32_t* h_plan = nullptr;  // pinned staging
    hipEvent_t plan_copied = nullptr;
    // Device work arrays.
    uint32_t* seg_hist = nullptr;  // [cap_clouds * digits * 256]
    uint32_t* lb = nullptr;        // 2 x cap_tiles x 256
    uint32_t* ctrs = nullptr;      // [16] per-pass tile counters
    uint32_t* ubase = nullptr;     // [cap_clouds + 1]
    uint32_t* node_base = nullptr; // [cap_clouds + 1]
    uint32_t* totals = nullptr;    // [4]: unique total, node total
    // Outputs (concatenated).
    void* codes = nullptr;
    void* alt = nullptr;
    void* uniq = nullptr;
    uint8_t* prefix_n = nullptr;
    uint8_t* hll = nullptr;
    uint8_t* hlr = nullptr;
    int32_t* left = nullptr;
    int32_t* parent = nullptr;
    uint32_t* ec = nullptr;
    uint32_t* offsets = nullptr;  // [cap_points + 1]
    tc_oct_node* nodes = nullptr;
    uint64_t node_cap = 0;
    // Host copy of the current plan shape.
    uint32_t n_clouds = 0, n_points = 0, n_tiles = 0;
    bool in_ptrs = false;  // the current plan reads clouds from per-cloud buffers
    // Graph.
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    const float* graph_pts = nullptr;
    int graph_link = -1;
};

namespace {

constexpr int SEG_BLOCK = 256;
// Workgroups per CU of the segmented tree kernels (SEG_BLOCK = one wave per
// SIMD, so also waves per SIMD): their launch bounds ask the compiler for this
// occupancy and their grids hold four times this many per CU.  Both fit 8
// without spills since the window staging keeps its per-lane load addresses
// out of the tile loops (the octree kernel needed 111 VGPRs before).
constexpr int SEG_BRT_PER_CU = 8;
constexpr int SEG_OCT_PER_CU = 8;
constexpr int SEG_ITEMS = 8;  // point loads in flight per lane in the Morton kernel
constexpr int SEG_BRT_TILE = 1024;  // slots per workgroup step of the tree kernels
constexpr int SEG_OCT_HALO = 128;   // LDS window halo of the octree walk
constexpr int SEG_SEARCH_HALO = 128;     // LDS window halo of the radix-tree searches
constexpr uint32_t SEG_BRT_LIMIT = 128;  // radix ranges searched in LDS only (within SEG_SEARCH_HALO; 64 before: 1,024-cloud replay 5.11 -> 5.05 ms)

// Plan views.
constexpr uint32_t PLAN_IN_PTRS = 1u;   // desc[3]: cloud c is read from in_ptr[c]
constexpr uint32_t PLAN_OUT_PTRS = 2u;  // desc[3]: outputs are written to out[c] (nodes only there)

struct Plan {
    const uint32_t* desc;  // [0] n_clouds [1] n_tiles [2] n_points [3] flags
    const uint32_t* pt_off;
    const uint32_t* first_tile;
    const float* min_c;
    const float* range_c;
    const uint64_t* in_ptr;
    const tc_tree_ptrs* out;
};

size_t plan_words(uint32_t C) { return 8 + 2 * ((size_t)C + 1) + 2 * (size_t)C + 2 * (size_t)C + 18 * (size_t)C; }

template <typename T>
Plan plan_at(T* base, uint32_t C) {
    Plan p;
    p.desc = base;
    p.pt_off = base + 8;
    p.first_tile = p.pt_off + (C + 1);
    p.min_c = reinterpret_cast<const float*>(p.first_tile + (C + 1));
    p.range_c = p.min_c + C;
    // 8 + 2(C+1) + 2C words: an even count, so the 64-bit arrays are aligned.
    p.in_ptr = reinterpret_cast<const uint64_t*>(p.range_c + C);
    p.out = reinterpret_cast<const tc_tree_ptrs*>(p.in_ptr + C);
    return p;
}

Plan plan_view(const tc_batch* b) { return plan_at(b->plan, b->cap_clouds); }

// Morton codes + per-cloud digit histograms, one sort tile at a time (the
// histogram is fused into the producer, so the codes are read only by the
// binning passes).  Blocks take contiguous runs of tiles and flush their LDS
// histogram to the cloud's counters when the cloud changes.
template <typename K>
__global__ __launch_bounds__(SEG_BLOCK) void k_seg_morton_hist(Plan p, const float4* __restrict__ pts,
                                                               K* __restrict__ codes, uint32_t* __restrict__ seg_hist,
                                                               uint32_t tile_size) {
    constexpr int DIG = (int)sizeof(K);
    constexpr bool TABLE = sizeof(K) == 4;  // 30-bit codes: bit spreading from an LDS table
    __shared__ uint32_t s_h[DIG * 256];
    __shared__ uint32_t s_split[TABLE ? 1024 : 1];
    __shared__ uint32_t s_cloud;
    const uint32_t C = p.desc[0], T = p.desc[1];
    const uint32_t chunk = ceil_div_u32(T, gridDim.x);
    const uint32_t t0 = blockIdx.x * chunk;
    const uint32_t t1 = t0 + chunk < T ? t0 + chunk : T;
    const uint32_t tid = threadIdx.x;
    for (uint32_t i = tid; i < DIG * 256; i += SEG_BLOCK) s_h[i] = 0;
    if constexpr (TABLE)
        for (uint32_t i = tid; i < 1024; i += SEG_BLOCK) s_split[i] = d_split3_10(i);
    uint32_t cur = 0xFFFFFFFFu;
    for (uint32_t t = t0; t < t1; ++t) {
        if (tid < 64) {
            // Tiles run in order: after the first, the cloud is usually the same one.
            const uint32_t c = cur == 0xFFFFFFFFu ? seg_find_wave(p.first_tile, C, t)
                                                  : seg_find_near(p.first_tile, C, t, cur);
            if (tid == 0) s_cloud = c;
        }
        __syncthreads();
        const uint32_t c = s_cloud;
        if (c != cur) {
            if (cur != 0xFFFFFFFFu) {
                for (uint32_t i = tid; i < DIG * 256; i += SEG_BLOCK) {
                    const uint32_t v = s_h[i];
                    if (v) atomicAdd(&seg_hist[(size_t)cur * DIG * 256 + i], v);
                    s_h[i] = 0;
                }
                __syncthreads();
            }
            cur = c;
        }
        const uint32_t start = p.pt_off[c] + (t - p.first_tile[c]) * tile_size;
        const uint32_t end_c = p.pt_off[c + 1];
        const uint32_t end = start + tile_size < end_c ? start + tile_size : end_c;
        const float mn = p.min_c[c], rg = p.range_c[c];
        const uint32_t c0 = p.pt_off[c];
        // Typed as global memory so the loads are global_load, not flat.
        using v4 = float __attribute__((ext_vector_type(4)));
        using gptr = const __attribute__((address_space(1))) v4*;
        const gptr src = (p.desc[3] & PLAN_IN_PTRS) ? (gptr)reinterpret_cast<const v4*>(p.in_ptr[c])
                                                    : (gptr)reinterpret_cast<const v4*>(pts + c0);
        // SEG_ITEMS point loads in flight per lane before any is encoded.
        for (uint32_t i0 = start + tid; i0 < end; i0 += SEG_BLOCK * SEG_ITEMS) {
            v4 qv[SEG_ITEMS];
#pragma unroll
            for (int u = 0; u < SEG_ITEMS; ++u) {
                const uint32_t i = i0 + u * SEG_BLOCK;
                if (i < end) qv[u] = ld_stream(&src[i - c0]);
            }
            K code[SEG_ITEMS];
#pragma unroll
            for (int u = 0; u < SEG_ITEMS; ++u) {
                const uint32_t i = i0 + u * SEG_BLOCK;
                code[u] = 0;
                if (i < end) {
                    if constexpr (TABLE) {
                        const MortonQuant q(mn, rg, 1024.0f);
                        code[u] = s_split[q(qv[u].x) & 1023u] | (s_split[q(qv[u].y) & 1023u] << 1) |
                                  (s_split[q(qv[u].z) & 1023u] << 2);
                    } else {
                        code[u] = d_encode63(qv[u].x, qv[u].y, qv[u].z, mn, rg);
                    }
                    codes[i] = code[u];
                }
            }
            // The wave-uniform shortcut of lds_count_digits only pays for
            // points in spatial order: taken for this round when the first
            // point's top digit is shared by the wave (item 0 is a point for
            // every lane still in the loop).
            const uint32_t top = (uint32_t)(code[0] >> (8 * (DIG - 1)));
            const bool uni = __ballot(top == __builtin_amdgcn_readfirstlane(top)) == __ballot(true);
#pragma unroll
            for (int u = 0; u < SEG_ITEMS; ++u) {
                if (i0 + u * SEG_BLOCK < end) {
                    if (uni) {
                        lds_count_digits<K, DIG>(s_h, code[u]);
                    } else {
#pragma unroll
                        for (int d = 0; d < DIG; ++d)
                            atomicAdd(&s_h[d * 256 + (uint32_t)((code[u] >> (8 * d)) & 255u)], 1u);
                    }
                }
            }
        }
        __syncthreads();
    }
    if (cur != 0xFFFFFFFFu) {
        for (uint32_t i = tid; i < DIG * 256; i += SEG_BLOCK) {
            const uint32_t v = s_h[i];
            if (v) atomicAdd(&seg_hist[(size_t)cur * DIG * 256 + i], v);
        }
    }
}

// Per cloud and digit: counts -> exclusive prefixes, in place.
template <int DIG>
__global__ __launch_bounds__(SEG_BLOCK) void k_seg_prefix(Plan p, uint32_t* __restrict__ seg_hist) {
    __shared__ uint32_t s_tmp[SEG_BLOCK / 64 + 2];
    const uint32_t C = p.desc[0];
    for (uint32_t c = blockIdx.x; c < C; c += gridDim.x) {
        for (int d = 0; d < DIG; ++d) {
            uint32_t* h = seg_hist + ((size_t)c * DIG + d) * 256;
            const uint32_t v = h[threadIdx.x];
            uint32_t tot;
            const uint32_t ex = block_exclusive_scan<SEG_BLOCK>(v, s_tmp, &tot);
            h[threadIdx.x] = ex;
        }
    }
}

// One OneSweep pass over every cloud: tiles claimed in order, look-back
// confined to the cloud's own tiles.
template <typename K, int BLOCK, int KPT, int RB>
__global__ __launch_bounds__(BLOCK) void k_seg_onesweep(Plan p, const K* __restrict__ sk, K* __restrict__ dk,
                                                        uint32_t shift, uint32_t pass, const uint32_t* __restrict__ seg_hist,
                                                        uint32_t* __restrict__ lb_cur, uint32_t* __restrict__ lb_next,
                                                        uint32_t spin_limit) {
    using Sh = OnesweepShared<K, BLOCK, KPT, RB>;
    constexpr int DIG = (int)sizeof(K);
    __shared__ Sh sm;
    const uint32_t C = p.desc[0], T = p.desc[1];
    // Grid = tile capacity, tile = workgroup id (see k_onesweep).
    const uint32_t tile = blockIdx.x;
    if (tile >= T) return;
    if (threadIdx.x < 64) {
        const uint32_t c = seg_find_wave(p.first_tile, C, tile);
        if (threadIdx.x == 0) sm.aux = c;
    }
    __syncthreads();
    const uint32_t c = sm.aux;
    const uint32_t first = p.first_tile[c];
    const uint32_t seg_base = p.pt_off[c];
    const uint32_t base = seg_base + (tile - first) * Sh::TILE;
    const uint32_t end = p.pt_off[c + 1];
    const uint32_t cnt = end - base < Sh::TILE ? end - base : Sh::TILE;
    onesweep_tile<K, false, BLOCK, KPT, RB>(sm, sk, dk, nullptr, nullptr, base, cnt, shift, tile, first, seg_base,
                                            seg_hist + ((size_t)c * DIG + pass) * 256, seg_base, lb_cur, lb_next,
                                            LB_ROW, spin_limit);
}

// Fix-ups after a segmented scan: empty segments get the base of the next
// non-empty one; entry C gets the total.
__global__ __launch_bounds__(SEG_BLOCK) void k_seg_finalize(const uint32_t* __restrict__ desc,
                                                            const uint32_t* __restrict__ seg_off,
                                                            uint32_t* __restrict__ base_out,
                                                            const uint32_t* __restrict__ total) {
    const uint32_t C = desc[0];
    const uint32_t tot = *total;
    for (uint32_t c = blockIdx.x * SEG_BLOCK + threadIdx.x; c <= C; c += gridDim.x * SEG_BLOCK) {
        if (c == C) {
            base_out[C] = tot;
            continue;
        }
        if (seg_off[c] != seg_off[c + 1]) continue;  // written by the scan
        uint32_t k = c + 1;
        while (k < C && seg_off[k] == seg_off[k + 1]) ++k;
        // k is the next non-empty segment, or C.  Its base is the exclusive
        // prefix at its first element; read it from the scan's output array
        // position instead of base_out (which the scan writes concurrently
        // only for non-empty segments, already finished here).
        base_out[c] = (k == C) ? tot : base_out[k];
    }
}

// Unique-key slots: radix tree (and edge counts) per cloud.  The last slot of
// each cloud has no radix node; its edge count is 0 so the node scan skips it.
template <typename K>
__global__ __launch_bounds__(SEG_BLOCK, SEG_BRT_PER_CU) void k_seg_brt(Plan p, const uint32_t* __restrict__ ubase,
                                                       const K* __restrict__ uniq, uint8_t* __restrict__ prefix_n,
                                                       uint8_t* __restrict__ hll, uint8_t* __restrict__ hlr,
                                                       int32_t* __restrict__ left, int32_t* __restrict__ parent,
                                                       uint32_t* __restrict__ ec) {
    // Tiles of SEG_BRT_TILE slots with the unique codes of the tile and a halo
    // staged in LDS (a cache of the concatenated array: each cloud still reads
    // only its own codes, by cloud-local index).  Bounded searches first, the
    // long ranges afterwards (see k_build_radix_tree).  With per-cloud output
    // buffers (tc_batch_set_io) the unique codes and radix arrays are also
    // stored there.
    // s_win: the concatenated codes of [ws, ws + W) (all-ones outside the
    // array); s_fast: the same with all-ones outside the tile's first cloud, so
    // that cloud's nodes search without bounds checks (as k_build_radix_tree).
    constexpr int W = SEG_BRT_TILE + 2 * SEG_SEARCH_HALO;
    static_assert(SEG_SEARCH_HALO >= (int)SEG_BRT_LIMIT, "bounded searches must stay inside the window");
    __shared__ K s_win[W];
    __shared__ K s_fast[W];
    __shared__ uint32_t s_long[SEG_BRT_TILE];
    __shared__ uint32_t s_nlong;
    __shared__ uint32_t s_c0, s_b0, s_e0;
    const uint32_t C = p.desc[0];
    const bool to_clouds = (p.desc[3] & PLAN_OUT_PTRS) != 0;
    const uint32_t total = ubase[C];
    const uint32_t tiles = ceil_div_u32(total, SEG_BRT_TILE);
    const BrtArrays all{prefix_n, hll, hlr, left, parent, ec};  // concatenated, addressed at b + index
    auto cloud_arrays = [&](uint32_t c) {
        const tc_tree_ptrs& o = p.out[c];
        return BrtArrays{o.prefix_n, o.has_leaf_left, o.has_leaf_right, o.left_child, o.parent, o.edge_count};
    };
    for (uint32_t t = blockIdx.x; t < tiles; t += gridDim.x) {
        const int64_t t0 = (int64_t)t * SEG_BRT_TILE;
        const int64_t ws = t0 - SEG_SEARCH_HALO;
        __syncthreads();
        if (threadIdx.x < 64) {
            // The cloud of the tile's first slot and its slot range [s_b0, s_e0).
            const uint32_t c = seg_find_wave(ubase, C, (uint32_t)t0);
            if (threadIdx.x == 0) {
                s_c0 = c;
                s_b0 = ubase[c];
                s_e0 = ubase[c + 1];
                s_nlong = 0;
            }
        }
        stage_codes<SEG_BLOCK, W>(s_win, uniq, ws, 0, total, (K)~(K)0);
        __syncthreads();
        const uint32_t c0 = s_c0, b0 = s_b0, e0 = s_e0;
        for (uint32_t q = threadIdx.x; q < (uint32_t)W; q += SEG_BLOCK) {
            const int64_t g = ws + q;
            s_fast[q] = (g >= (int64_t)b0 && g < (int64_t)e0) ? s_win[q] : (K)~(K)0;
        }
        __syncthreads();
        // Cloud, first slot and key count of slot g: from LDS while g is in the
        // tile's first cloud (no dependent global loads before the search).
        auto cloud_of = [&](uint32_t g, uint32_t& c, uint32_t& b, uint32_t& nk) {
            if (g < e0) {
                c = c0;
                b = b0;
                nk = e0 - b0;
            } else {
                c = seg_find_near(ubase, C, g, c0);
                b = ubase[c];
                nk = ubase[c + 1] - b;
            }
        };
        // The lane's slots in the tile's first cloud, searched together (as in
        // k_build_radix_tree); other slots search the tile's first slot (in
        // that cloud, every probe inside the window) and discard the result.
        constexpr int NPL = SEG_BRT_TILE / SEG_BLOCK;
        uint32_t idx[NPL];
#pragma unroll
        for (int k = 0; k < NPL; ++k) {
            const uint32_t g = (uint32_t)t0 + k * SEG_BLOCK + threadIdx.x;
            idx[k] = (g < e0 && g < total) ? g - b0 : (uint32_t)t0 - b0;
        }
        BrtResult rr[NPL];
        const uint32_t near =
            brt_nodes_near<K, SEG_BRT_LIMIT, NPL>(as_lds<K>(s_fast), (int32_t)(ws - (int64_t)b0), idx, true, rr);
#pragma unroll
        for (int k = 0; k < NPL; ++k) {
            const uint32_t g = (uint32_t)t0 + k * SEG_BLOCK + threadIdx.x;
            if (g >= total) continue;
            uint32_t c, b, nk;
            cloud_of(g, c, b, nk);
            const uint32_t i = g - b;
            if (to_clouds) static_cast<K*>(p.out[c].unique)[i] = s_win[g - ws];
            if (i + 1 < nk) {
                // Nodes of later clouds in the tile (tiles that straddle a
                // cloud boundary) take the full search below.
                if (c == c0 && ((near >> k) & 1u)) {
                    brt_store(all, i, rr[k], b);
                    if (to_clouds) brt_store(cloud_arrays(c), i, rr[k]);
                } else {
                    s_long[atomicAdd(&s_nlong, 1u)] = g;
                }
            } else {
                ec[g] = 0u;
            }
        }
        __syncthreads();
        const uint32_t nlong = s_nlong;
        constexpr uint32_t G = BRT_LONG_GROUP;  // lanes per long node
        for (uint32_t q = threadIdx.x / G; q < nlong; q += SEG_BLOCK / G) {
            const uint32_t g = s_long[q];
            uint32_t c, b, nk;
            cloud_of(g, c, b, nk);
            // The raw window holds real codes for every index of the cloud inside it.
            const CodeView<K> view{uniq + b, as_lds<K>(s_win), (int32_t)(ws - (int64_t)b), (uint32_t)W};
            BrtResult r;
            brt_node_group<K, G>(view, nk, g - b, true, &r);
            if (threadIdx.x % G == 0) {
                brt_store(all, g - b, r, b);
                if (to_clouds) brt_store(cloud_arrays(c), g - b, r);
            }
        }
    }
}

template <typename K>
__global__ __launch_bounds__(SEG_BLOCK, SEG_OCT_PER_CU) void k_seg_octree(const uint32_t* __restrict__ desc, const uint32_t* __restrict__ ubase,
                                                          const uint32_t* __restrict__ node_base, tc_oct_node* __restrict__ nodes,
                                                          uint32_t node_cap, const uint32_t* __restrict__ offsets,
                                                          const uint32_t* __restrict__ ec, const K* __restrict__ uniq,
                                                          const uint8_t* __restrict__ prefix_n, const int32_t* __restrict__ left,
                                                          const uint8_t* __restrict__ hll, const uint8_t* __restrict__ hlr,
                                                          int link_leaves, Plan p) {
    // Tiles of slots with the links of the tile and a halo staged in LDS (see
    // k_octree_build); the window is a cache of the concatenated arrays.
    constexpr int W = SEG_BRT_TILE + 2 * SEG_OCT_HALO;
    __shared__ int32_t s_left[W];
    __shared__ uint16_t s_meta[W];
    __shared__ K s_code[W];
    __shared__ uint32_t s_off[W];
    __shared__ uint16_t s_list[SEG_BRT_TILE];  // emitting slots, tile-relative
    __shared__ uint32_t s_tmp[SEG_BLOCK / 64 + 2];
    __shared__ uint32_t s_c0, s_b0, s_e0, s_nb0;
    __shared__ int s_rl0;
    __shared__ float s_min0, s_range0;
    const uint32_t C = desc[0];
    const uint32_t total = ubase[C];
    const uint32_t tiles = ceil_div_u32(total, SEG_BRT_TILE);
    for (uint32_t t = blockIdx.x; t < tiles; t += gridDim.x) {
        const int64_t t0 = (int64_t)t * SEG_BRT_TILE;
        const int64_t w0 = t0 > SEG_OCT_HALO ? t0 - SEG_OCT_HALO : 0;
        const int64_t w1 = t0 + SEG_BRT_TILE + SEG_OCT_HALO < total ? t0 + SEG_BRT_TILE + SEG_OCT_HALO : total;
        __syncthreads();
        if (threadIdx.x < 64) {
            // The tile's first cloud: slot range, node base, root level, quantiser.
            // (prefix_n[s_b0] is only used when the cloud has a radix node.)
            const uint32_t c = seg_find_wave(ubase, C, (uint32_t)t0);
            if (threadIdx.x == 0) {
                s_c0 = c;
                s_b0 = ubase[c];
                s_e0 = ubase[c + 1];
                s_nb0 = node_base[c];
                s_rl0 = prefix_n[s_b0] / 3;
                s_min0 = p.min_c[c];
                s_range0 = p.range_c[c];
            }
        }
        stage_tree<SEG_BLOCK, W, K>(s_left, s_meta, s_code, s_off, left, ec, hll, hlr, prefix_n, uniq, offsets, w0,
                                    w1);
        __syncthreads();
        const uint32_t c0 = s_c0;
        const bool to_clouds = (p.desc[3] & PLAN_OUT_PTRS) != 0;
        if (to_clouds) {  // cloud-local node offsets into the cloud's own buffer
#pragma unroll 1
            for (int k = 0; k < SEG_BRT_TILE / SEG_BLOCK; ++k) {
                const uint32_t g = (uint32_t)t0 + k * SEG_BLOCK + threadIdx.x;
                if (g >= total) break;
                const uint32_t c = seg_find_near(ubase, C, g, c0);
                p.out[c].offsets[g - ubase[c]] = offsets[g] - node_base[c];
            }
        }
        // Compact the slots with ec > 0 (the last slot of a cloud has ec 0);
        // thread t looks at slots 4t .. 4t + 3 of the tile.
        constexpr int NPT = SEG_BRT_TILE / SEG_BLOCK;
        uint32_t flags = 0;
#pragma unroll
        for (int k = 0; k < NPT; ++k) {
            const int64_t g = t0 + NPT * threadIdx.x + k;
            flags |= ((g < total && (s_meta[g - w0] & 255u) != 0u) ? 1u : 0u) << k;
        }
        uint32_t count;
        uint32_t rank = block_exclusive_scan<SEG_BLOCK>(__popc(flags), s_tmp, &count);
#pragma unroll
        for (int k = 0; k < NPT; ++k)
            if ((flags >> k) & 1u) s_list[rank++] = (uint16_t)(NPT * threadIdx.x + k);
        __syncthreads();
#pragma unroll 1
        for (uint32_t q = threadIdx.x; q < count; q += SEG_BLOCK) {
            const uint32_t g = (uint32_t)t0 + s_list[q];
            // Cloud data from LDS while g is in the tile's first cloud.
            uint32_t c = s_c0, b = s_b0, e = s_e0, nbase = s_nb0;
            int root_level = s_rl0;
            float mn = s_min0, rg = s_range0;
            if (g >= e) {
                c = seg_find_near(ubase, C, g, c0);
                b = ubase[c];
                e = ubase[c + 1];
                nbase = node_base[c];
                root_level = prefix_n[b] / 3;
                mn = p.min_c[c];
                rg = p.range_c[c];
            }
            const uint32_t nk = e - b;
            const uint32_t i = g - b;
            if (i + 1 >= nk) continue;
            const TreeView<K> lv{left, hll, hlr, ec, prefix_n, uniq, offsets, as_lds<int32_t>(s_left),
                                 as_lds<uint16_t>(s_meta), as_lds<K>(s_code), as_lds<uint32_t>(s_off),
                                 (int32_t)(w0 - (int64_t)b), (uint32_t)(w1 - w0), b};
            // Nodes go to the cloud's own buffer when there is one (then only there).
            tc_oct_node* dst = to_clouds ? p.out[c].nodes : nodes;
            const uint32_t cap = to_clouds ? d_octree_node_bound(nk, CodeTraits<K>::BITS) : node_cap;
            octree_build_node<K>(dst, cap, nbase, lv, i, nk - 1, root_level, link_leaves, mn, rg,
                                 to_clouds ? nbase : 0u);
        }
    }
}

template <typename K, int BLOCK, int KPT>
void launch_seg_pass(tc_batch* b, hipStream_t s, uint32_t grid, const K* src, K* dst, uint32_t pass) {
    const Plan p = plan_view(b);
    uint32_t* lb0 = b->lb + (size_t)(pass & 1) * b->cap_tiles * LB_ROW;
    uint32_t* lb1 = b->lb + (size_t)((pass + 1) & 1) * b->cap_tiles * LB_ROW;
    uint32_t* nxt = (pass + 1 < sizeof(K)) ? lb1 : nullptr;
    // The codes are Morton codes: the top digit has fewer live bits.
    if (pass + 1 == sizeof(K))
        hipLaunchKernelGGL((k_seg_onesweep<K, BLOCK, KPT, morton_top_digit_bits<K>()>), dim3(grid), dim3(BLOCK), 0, s,
                           p, src, dst, 8u * pass, pass, b->seg_hist, lb0, nxt, b->ctx->spin_limit);
    else
        hipLaunchKernelGGL((k_seg_onesweep<K, BLOCK, KPT, 8>), dim3(grid), dim3(BLOCK), 0, s, p, src, dst, 8u * pass,
                           pass, b->seg_hist, lb0, nxt, b->ctx->spin_limit);
}

template <typename K>
int seg_pass(tc_batch* b, hipStream_t s, const K* src, K* dst, uint32_t pass) {
    // One workgroup per tile of capacity (a replayed graph keeps this grid).
    const uint32_t grid = b->cap_tiles ? b->cap_tiles : 1;
#define TC_SEG_CASE(B, KP)                                            \
    if (b->block == B && b->kpt == KP) {                              \
        launch_seg_pass<K, B, KP>(b, s, grid, src, dst, pass);        \
        TC_LAUNCH_CHECK();                                            \
        return TC_OK;                                                 \
    }
    TC_SEG_CASE(256, 8)
    TC_SEG_CASE(512, 8)
    TC_SEG_CASE(512, 12)
    TC_SEG_CASE(512, 16)
    TC_SEG_CASE(512, 20)
    TC_SEG_CASE(512, 24)
    TC_SEG_CASE(512, 28)
    TC_SEG_CASE(512, 32)
    TC_SEG_CASE(512, 36)
    TC_SEG_CASE(1024, 16)
#undef TC_SEG_CASE
    return TC_ERR_INVALID;
}

template <typename K>
int batch_pipeline(tc_batch* b, const float* pts, int link_leaves, hipStream_t s) {
    tc_ctx* ctx = b->ctx;
    constexpr int DIG = (int)sizeof(K);
    const Plan p = plan_view(b);
    const uint32_t C = b->cap_clouds;
    K* codes = static_cast<K*>(b->codes);
    K* alt = static_cast<K*>(b->alt);
    K* uniq = static_cast<K*>(b->uniq);
    ProfScope prof(ctx, s, K_SEG);
    // Clear per-cloud histogram counters, tile counters and look-back region 0.
    TC_TRY(tc_fill_u32(ctx, b->seg_hist, 0u, (size_t)C * DIG * 256, s));
    TC_TRY(tc_fill_u32(ctx, b->ctrs, 0u, 16, s));
    TC_TRY(tc_fill_u32(ctx, b->lb, 0u, (size_t)b->cap_tiles * 256, s));
    // 16 workgroups per CU, each a contiguous run of tiles (4 per CU: 0.667 ms
    // per replay of the 1,024-cloud batch; 16: 0.646 ms).
    uint32_t grid = (uint32_t)ctx->num_cus * 16;
    if (grid > b->cap_tiles) grid = b->cap_tiles;
    hipLaunchKernelGGL(k_seg_morton_hist<K>, dim3(grid), dim3(SEG_BLOCK), 0, s, p,
                       reinterpret_cast<const float4*>(pts), codes, b->seg_hist, b->tile);
    TC_LAUNCH_CHECK();
    uint32_t cgrid = C < (uint32_t)ctx->num_cus * 4 ? C : (uint32_t)ctx->num_cus * 4;
    hipLaunchKernelGGL(k_seg_prefix<DIG>, dim3(cgrid ? cgrid : 1), dim3(SEG_BLOCK), 0, s, p, b->seg_hist);
    TC_LAUNCH_CHECK();
    K* src = codes;
    K* dst = alt;
    for (int pass = 0; pass < DIG; ++pass) {
        TC_TRY(seg_pass<K>(b, s, src, dst, (uint32_t)pass));
        K* t = src; src = dst; dst = t;
    }
    // Even pass count: sorted codes are back in `codes`.
    TC_TRY(tc_launch_scan(ctx, DIG == 4 ? MODE_SEG_UNIQUE32 : MODE_SEG_UNIQUE64, codes, nullptr, uniq, nullptr,
                          b->totals, b->cap_points, nullptr, p.pt_off, C, p.desc, b->ubase, s));
    const uint32_t fgrid = ceil_div_u32(C + 1, SEG_BLOCK);
    hipLaunchKernelGGL(k_seg_finalize, dim3(fgrid), dim3(SEG_BLOCK), 0, s, p.desc, p.pt_off, b->ubase, b->totals);
    TC_LAUNCH_CHECK();
    // The u64 radix-tree kernel's two LDS windows (24 KiB) fit 6 per CU.
    const int brt_per_cu = sizeof(K) == 4 ? SEG_BRT_PER_CU : 6;
    const uint32_t tgrid = (uint32_t)tc_persistent_grid(ctx, b->cap_points, SEG_BLOCK, 4 * brt_per_cu);
    hipLaunchKernelGGL(k_seg_brt<K>, dim3(tgrid), dim3(SEG_BLOCK), 0, s, p, b->ubase, uniq, b->prefix_n,
                       b->hll, b->hlr, b->left, b->parent, b->ec);
    TC_LAUNCH_CHECK();
    TC_TRY(tc_launch_scan(ctx, MODE_SEG_SCAN, b->ec, nullptr, b->offsets, nullptr, b->totals + 1, b->cap_points,
                          nullptr, b->ubase, C, p.desc, b->node_base, s));
    hipLaunchKernelGGL(k_seg_finalize, dim3(fgrid), dim3(SEG_BLOCK), 0, s, p.desc, b->ubase, b->node_base,
                       b->totals + 1);
    TC_LAUNCH_CHECK();
    const uint32_t ncap = (uint32_t)(b->node_cap < 0xFFFFFFFFull ? b->node_cap : 0xFFFFFFFFull);
    const uint32_t ogrid = (uint32_t)tc_persistent_grid(ctx, b->cap_points, SEG_BLOCK, 4 * SEG_OCT_PER_CU);
    hipLaunchKernelGGL(k_seg_octree<K>, dim3(ogrid), dim3(SEG_BLOCK), 0, s, p.desc, b->ubase, b->node_base,
                       b->nodes, ncap, b->offsets, b->ec, uniq, b->prefix_n, b->left, b->hll, b->hlr, link_leaves, p);
    TC_LAUNCH_CHECK();
    return TC_OK;
}

int run_pipeline(tc_batch* b, const float* pts, int link_leaves, hipStream_t s) {
    return b->code_bits == 30 ? batch_pipeline<uint32_t>(b, pts, link_leaves, s)
                              : batch_pipeline<uint64_t>(b, pts, link_leaves, s);
}

void free_batch(tc_batch* b) {
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    if (b->plan_copied) (void)hipEventDestroy(b->plan_copied);
    if (b->h_plan) (void)hipHostFree(b->h_plan);
    void* ptrs[] = {b->plan, b->seg_hist, b->lb, b->ctrs, b->ubase, b->node_base, b->totals, b->codes, b->alt,
                    b->uniq, b->prefix_n, b->hll, b->hlr, b->left, b->parent, b->ec, b->offsets, b->nodes};
    for (void* q : ptrs)
        if (q) (void)hipFree(q);
    delete b;
}

}  // namespace

extern "C" int tc_batch_create(tc_ctx* ctx, uint32_t cap_clouds, uint32_t cap_points, int code_bits,
                               tc_batch** out) {
    if (!ctx || !out || cap_clouds == 0 || (code_bits != 30 && code_bits != 63)) return TC_ERR_INVALID;
    if (cap_points > TC_MAX_SORT_N) return TC_ERR_TOO_LARGE;
    *out = nullptr;
    if (hipSetDevice(ctx->device) != hipSuccess) return TC_ERR_HIP;
    tc_batch* b = new (std::nothrow) tc_batch();
    if (!b) return TC_ERR_NOMEM;
    b->ctx = ctx;
    b->cap_clouds = cap_clouds;
    b->cap_points = cap_points;
    b->code_bits = code_bits;
    // Launch shape of the binning passes: the tuned one for the average cloud.
    // One segmented launch sorts every cloud's keys: the shape tuned for the
    // total key count (measured: 1,024 clouds of 140.9 M points in all take
    // 7.41 ms per replay with 512 x 20 or larger tiles, 7.86 ms with the
    // 512 x 8 tiles tuned for their average size).
    const SortShape sh = tc_sort_shape_8bit(ctx, code_bits == 30 ? 4 : 8, false, cap_points);
    b->block = sh.block;
    b->kpt = sh.kpt;
    b->tile = (uint32_t)(b->block * b->kpt);
    b->cap_tiles = ceil_div_u32(cap_points, b->tile) + cap_clouds;
    const size_t C = cap_clouds, N = cap_points ? cap_points : 1;
    const size_t ks = code_bits == 30 ? 4 : 8;
    const int dig = code_bits == 30 ? 4 : 8;
    b->plan_words = plan_words(cap_clouds);
    // Node bound: per cloud it is tc_octree_node_bound(n_c); the sum over clouds
    // is at most the bound for the total plus one root per cloud.
    b->node_cap = tc_octree_node_bound(N, code_bits) + C;
    bool ok = hipMalloc((void**)&b->plan, b->plan_words * 4) == hipSuccess &&
              hipHostMalloc((void**)&b->h_plan, b->plan_words * 4) == hipSuccess &&
              hipEventCreateWithFlags(&b->plan_copied, hipEventDisableTiming) == hipSuccess &&
              hipMalloc((void**)&b->seg_hist, C * dig * 256 * 4) == hipSuccess &&
              hipMalloc((void**)&b->lb, 2 * (size_t)b->cap_tiles * 256 * 4) == hipSuccess &&
              hipMalloc((void**)&b->ctrs, 64) == hipSuccess &&
              hipMalloc((void**)&b->ubase, (C + 1) * 4) == hipSuccess &&
              hipMalloc((void**)&b->node_base, (C + 1) * 4) == hipSuccess &&
              hipMalloc((void**)&b->totals, 16) == hipSuccess && hipMalloc(&b->codes, N * ks) == hipSuccess &&
              hipMalloc(&b->alt, N * ks) == hipSuccess && hipMalloc(&b->uniq, N * ks) == hipSuccess &&
              hipMalloc((void**)&b->prefix_n, N) == hipSuccess && hipMalloc((void**)&b->hll, N) == hipSuccess &&
              hipMalloc((void**)&b->hlr, N) == hipSuccess && hipMalloc((void**)&b->left, N * 4) == hipSuccess &&
              hipMalloc((void**)&b->parent, N * 4) == hipSuccess && hipMalloc((void**)&b->ec, N * 4) == hipSuccess &&
              hipMalloc((void**)&b->offsets, (N + 1) * 4) == hipSuccess &&
              hipMalloc((void**)&b->nodes, b->node_cap * sizeof(tc_oct_node)) == hipSuccess;
    if (!ok) {
        free_batch(b);
        return TC_ERR_NOMEM;
    }
    std::memset(b->h_plan, 0, b->plan_words * 4);
    if (hipMemset(b->plan, 0, b->plan_words * 4) != hipSuccess || hipDeviceSynchronize() != hipSuccess) {
        free_batch(b);
        return TC_ERR_HIP;
    }
    *out = b;
    return TC_OK;
}

extern "C" int tc_batch_destroy(tc_batch* b) {
    if (!b) return TC_ERR_INVALID;
    (void)hipSetDevice(b->ctx->device);
    (void)hipDeviceSynchronize();
    free_batch(b);
    return TC_OK;
}

extern "C" int tc_batch_set_io(tc_batch* b, uint32_t n_clouds, const uint32_t* counts, const float* mins,
                               const float* ranges, const float* const* cloud_pts, const tc_tree_ptrs* outs,
                               tc_stream stream) {
    if (!b || (n_clouds && !counts) || n_clouds > b->cap_clouds) return TC_ERR_INVALID;
    const uint32_t C = b->cap_clouds;
    uint64_t pts = 0, tiles = 0;
    for (uint32_t c = 0; c < n_clouds; ++c) {
        pts += counts[c];
        tiles += ceil_div_u32(counts[c], b->tile);
        const float r = ranges ? ranges[c] : 1024.0f;
        if (!(r != 0.0f)) return TC_ERR_INVALID;
        if (cloud_pts && counts[c] && (!cloud_pts[c] || (reinterpret_cast<uintptr_t>(cloud_pts[c]) & 15u)))
            return TC_ERR_INVALID;
    }
    if (pts > b->cap_points) return TC_ERR_CAPACITY;
    if (tiles > b->cap_tiles) return TC_ERR_CAPACITY;
    // The previous upload must have left the staging buffer.
    TC_HIP(hipEventSynchronize(b->plan_copied));
    uint32_t* h = b->h_plan;
    const Plan hp = plan_at(h, C);
    uint32_t* pt_off = const_cast<uint32_t*>(hp.pt_off);
    uint32_t* first = const_cast<uint32_t*>(hp.first_tile);
    float* mn = const_cast<float*>(hp.min_c);
    float* rg = const_cast<float*>(hp.range_c);
    uint64_t* in_ptr = const_cast<uint64_t*>(hp.in_ptr);
    tc_tree_ptrs* out = const_cast<tc_tree_ptrs*>(hp.out);
    pts = 0;
    tiles = 0;
    for (uint32_t c = 0; c < n_clouds; ++c) {
        pt_off[c] = (uint32_t)pts;
        first[c] = (uint32_t)tiles;
        pts += counts[c];
        tiles += ceil_div_u32(counts[c], b->tile);
        mn[c] = mins ? mins[c] : 0.0f;
        rg[c] = ranges ? ranges[c] : 1024.0f;
        in_ptr[c] = cloud_pts ? reinterpret_cast<uint64_t>(cloud_pts[c]) : 0ull;
        if (outs) out[c] = outs[c];
    }
    pt_off[n_clouds] = (uint32_t)pts;
    first[n_clouds] = (uint32_t)tiles;
    h[0] = n_clouds;
    h[1] = (uint32_t)tiles;
    h[2] = (uint32_t)pts;
    h[3] = (cloud_pts ? PLAN_IN_PTRS : 0u) | (outs ? PLAN_OUT_PTRS : 0u);
    b->n_clouds = n_clouds;
    b->n_points = (uint32_t)pts;
    b->n_tiles = (uint32_t)tiles;
    b->in_ptrs = cloud_pts != nullptr;
    hipStream_t s = as_stream(stream);
    TC_HIP(hipMemcpyAsync(b->plan, h, b->plan_words * 4, hipMemcpyHostToDevice, s));
    TC_HIP(hipEventRecord(b->plan_copied, s));
    return TC_OK;
}

extern "C" int tc_batch_set(tc_batch* b, uint32_t n_clouds, const uint32_t* counts, const float* mins,
                            const float* ranges, tc_stream stream) {
    return tc_batch_set_io(b, n_clouds, counts, mins, ranges, nullptr, nullptr, stream);
}

extern "C" int tc_batch_build(tc_batch* b, const float* pts, int link_leaves, tc_stream stream) {
    if (!b || (b->n_points && !pts && !b->in_ptrs)) return TC_ERR_INVALID;
    if (reinterpret_cast<uintptr_t>(pts) & 15u) return TC_ERR_INVALID;
    const MarkRange mark(b->ctx, "tc_batch_build");
    return run_pipeline(b, pts, link_leaves, as_stream(stream));
}

extern "C" int tc_batch_capture(tc_batch* b, const float* pts, int link_leaves, tc_stream stream) {
    if (!b || (!pts && !b->in_ptrs) || (reinterpret_cast<uintptr_t>(pts) & 15u)) return TC_ERR_INVALID;
    hipStream_t s = as_stream(stream);
    if (!s) return TC_ERR_INVALID;  // the legacy stream cannot be captured
    if (b->exec) {
        (void)hipGraphExecDestroy(b->exec);
        b->exec = nullptr;
    }
    if (b->graph) {
        (void)hipGraphDestroy(b->graph);
        b->graph = nullptr;
    }
    const bool prof = b->ctx->prof_on;
    b->ctx->prof_on = false;  // no event records inside the graph
    TC_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    const int st = run_pipeline(b, pts, link_leaves, s);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(s, &g);
    b->ctx->prof_on = prof;
    if (st != TC_OK) {
        if (g) (void)hipGraphDestroy(g);
        return st;
    }
    if (e != hipSuccess || !g) return TC_ERR_HIP;
    b->graph = g;
    if (hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0) != hipSuccess) return TC_ERR_HIP;
    b->graph_pts = pts;
    b->graph_link = link_leaves;
    return TC_OK;
}

extern "C" int tc_batch_replay(tc_batch* b, tc_stream stream) {
    if (!b || !b->exec) return TC_ERR_INVALID;
    if (!b->graph_pts && !b->in_ptrs && b->n_points) return TC_ERR_INVALID;  // captured without an input array
    const MarkRange mark(b->ctx, "tc_batch_replay");
    ProfScope prof(b->ctx, as_stream(stream), K_SEG);
    TC_HIP(hipGraphLaunch(b->exec, as_stream(stream)));
    return TC_OK;
}

extern "C" int tc_batch_read_info(tc_batch* b, tc_cloud_info* info, tc_stream stream) {
    if (!b || (b->n_clouds && !info)) return TC_ERR_INVALID;
    TC_HIP(hipStreamSynchronize(as_stream(stream)));
    const uint32_t C = b->n_clouds;
    std::vector<uint32_t> ub(C + 1), nbase(C + 1);
    TC_HIP(hipMemcpy(ub.data(), b->ubase, (C + 1) * 4, hipMemcpyDeviceToHost));
    TC_HIP(hipMemcpy(nbase.data(), b->node_base, (C + 1) * 4, hipMemcpyDeviceToHost));
    const uint32_t* pt_off = b->h_plan + 8;
    for (uint32_t c = 0; c < C; ++c) {
        info[c].point_offset = pt_off[c];
        info[c].n_points = pt_off[c + 1] - pt_off[c];
        info[c].unique_offset = ub[c];
        info[c].n_unique = ub[c + 1] - ub[c];
        info[c].node_offset = nbase[c];
        info[c].n_nodes = nbase[c + 1] - nbase[c];
    }
    if (nbase[C] > b->node_cap) return TC_ERR_CAPACITY;
    return TC_OK;
}

extern "C" int tc_batch_arrays(tc_batch* b, tc_batch_view* v) {
    if (!b || !v) return TC_ERR_INVALID;
    v->code_bits = b->code_bits;
    v->n_clouds = b->n_clouds;
    v->n_points = b->n_points;
    v->sorted = b->codes;
    v->unique = b->uniq;
    v->prefix_n = b->prefix_n;
    v->has_leaf_left = b->hll;
    v->has_leaf_right = b->hlr;
    v->left_child = b->left;
    v->parent = b->parent;
    v->edge_count = b->ec;
    v->offsets = b->offsets;
    v->nodes = b->nodes;
    v->node_capacity = b->node_cap;
    return TC_OK;
}

extern "C" uint64_t tc_batch_memory_bytes(uint32_t max_clouds, uint32_t max_points, int code_bits) {
    const uint64_t C = max_clouds, N = max_points ? max_points : 1;
    const uint64_t ks = code_bits == 30 ? 4 : 8;
    const uint64_t dig = code_bits == 30 ? 4 : 8;
    const uint64_t tiles = ceil_div_u32(N, 256 * 8) + C;  // smallest tile shape: an upper bound
    const uint64_t node_cap = tc_octree_node_bound(N, code_bits) + C;
    return plan_words(max_clouds) * 4 + C * dig * 256 * 4 + 2 * tiles * 256 * 4 + 64 + 2 * (C + 1) * 4 + 16 +
           3 * N * ks + 3 * N + 3 * N * 4 + (N + 1) * 4 + node_cap * sizeof(tc_oct_node);
}

extern "C" int tc_batch_read_info_async(tc_batch* b, uint32_t* h_unique_base, uint32_t* h_node_base,
                                        tc_stream stream) {
    if (!b || !h_unique_base || !h_node_base) return TC_ERR_INVALID;
    const size_t bytes = ((size_t)b->n_clouds + 1) * 4;
    TC_HIP(hipMemcpyAsync(h_unique_base, b->ubase, bytes, hipMemcpyDeviceToHost, as_stream(stream)));
    TC_HIP(hipMemcpyAsync(h_node_base, b->node_base, bytes, hipMemcpyDeviceToHost, as_stream(stream)));
    return TC_OK;
}
