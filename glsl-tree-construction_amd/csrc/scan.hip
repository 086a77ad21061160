// scan.hip -- single-pass tile scans with decoupled look-back: unique
// compaction, exclusive scan, and the fused edge-count + exclusive scan.
//
// Reference: k_CountUnique, shader/unique.comp:31-45 and k_PartialSum,
// shader/partial_sum.comp:23-39 are single-thread loops ([1,1,1] dispatches,
// src/unique.rs:245, src/partial_sum.rs:246); countEdges,
// shader/edge_count.comp:30-41.  Published algorithm: Merrill & Garland,
// "Single-pass Parallel Prefix Scan with Decoupled Look-back" (2016).
//
// gfx950 design: persistent workgroups claim 4096-element tiles from an atomic
// counter (so every waited-on tile is owned by a running workgroup); each
// thread owns 16 consecutive elements; a block scan gives tile-local offsets;
// wave 0 publishes the tile aggregate and looks back over 64 predecessors at a
// time.  Tile status words are u64 [value:32 | flag:2 | epoch:30]; the epoch is
// a device-side launch sequence number, so status words never need clearing
// (also under graph replay).  The last workgroup to exit resets the tile
// counter and advances the sequence number.
#include "tc_internal.hpp"

namespace {

constexpr int SCAN_BLOCK = 256;
constexpr int SCAN_IPT = 16;
constexpr uint32_t SCAN_TILE = SCAN_BLOCK * SCAN_IPT;

enum ScanMode { MODE_SCAN = 0, MODE_UNIQUE32 = 1, MODE_UNIQUE64 = 2, MODE_ECSCAN = 3 };

// Scan control words (in ctx->small): [0] tile counter, [1] exit counter, [2] launch sequence.
struct ScanCtl {
    uint32_t tile_ctr;
    uint32_t exit_ctr;
    uint32_t seq;
};

TC_DEV uint64_t st_word(uint32_t flag, uint32_t epoch, uint32_t value) {
    return ((uint64_t)((epoch << 2) | flag) << 32) | value;
}

TC_DEV uint32_t wave_sum(uint32_t v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Look-back by wave 0; returns the exclusive prefix of `tile`.
TC_DEV uint32_t wave_lookback(const uint64_t* status, uint32_t tile, uint32_t epoch) {
    const uint32_t lane = lane_id();
    uint32_t excl = 0;
    int64_t top = (int64_t)tile - 1;
    while (true) {
        const int64_t idx = top - (int64_t)lane;
        uint32_t f, v;
        uint64_t incl_mask, upto;
        while (true) {
            if (idx >= 0) {
                const uint64_t w = ld_agent64(&status[idx]);
                const uint32_t hi = (uint32_t)(w >> 32);
                f = ((hi >> 2) == epoch) ? (hi & 3u) : 0u;
                v = (uint32_t)w;
            } else {
                f = 2u;  // before tile 0: an inclusive zero
                v = 0u;
            }
            incl_mask = __ballot(f == 2u);
            const uint64_t nr = __ballot(f == 0u);
            upto = incl_mask ? ((incl_mask & (~incl_mask + 1)) << 1) - 1 : ~0ull;
            if (incl_mask == (1ull << 63)) upto = ~0ull;
            if ((nr & upto) == 0) break;
            __builtin_amdgcn_s_sleep(1);
        }
        const bool take = (upto >> lane) & 1ull;
        excl += wave_sum(take ? v : 0u);
        if (incl_mask) break;
        top -= 64;
    }
    return excl;
}

template <int MODE>
__global__ __launch_bounds__(SCAN_BLOCK) void k_scan_tiles(const void* __restrict__ in0,
                                                           const void* __restrict__ in1,
                                                           void* __restrict__ out0, void* __restrict__ out1,
                                                           uint32_t* __restrict__ count_out, uint32_t n_cap,
                                                           const uint32_t* __restrict__ d_n_keys,
                                                           uint64_t* __restrict__ status,
                                                           ScanCtl* __restrict__ ctl) {
    __shared__ uint32_t s_tmp[SCAN_BLOCK / 64 + 2];
    __shared__ uint32_t s_tile;
    __shared__ uint32_t s_excl;

    const uint32_t tid = threadIdx.x;
    const uint32_t wave = tid >> 6;
    const uint32_t epoch = ld_acquire_agent(&ctl->seq) & 0x3FFFFFFFu;

    uint32_t n = n_cap;
    if (d_n_keys) {
        const uint32_t u = *d_n_keys;
        uint32_t m = (MODE == MODE_SCAN || MODE == MODE_ECSCAN) ? (u > 0 ? u - 1 : 0) : u;
        n = m < n_cap ? m : n_cap;
    }
    const uint32_t num_tiles = n ? ceil_div_u32(n, SCAN_TILE) : 1u;

    while (true) {
        if (tid == 0) s_tile = atomicAdd(&ctl->tile_ctr, 1u);
        __syncthreads();
        const uint32_t tile = s_tile;
        __syncthreads();
        if (tile >= num_tiles) break;

        const uint32_t base = tile * SCAN_TILE;
        const uint32_t i0 = base + tid * SCAN_IPT;
        uint32_t val[SCAN_IPT];
        uint64_t key64[MODE == MODE_UNIQUE64 ? SCAN_IPT : 1];
        uint32_t key32[MODE == MODE_UNIQUE32 ? SCAN_IPT : 1];

        if (MODE == MODE_SCAN) {
            const uint32_t* in = static_cast<const uint32_t*>(in0);
#pragma unroll
            for (int j = 0; j < SCAN_IPT; ++j) val[j] = (i0 + j < n) ? in[i0 + j] : 0u;
        } else if (MODE == MODE_UNIQUE32) {
            const uint32_t* in = static_cast<const uint32_t*>(in0);
            uint32_t prev = (i0 > 0 && i0 - 1 < n) ? in[i0 - 1] : 0u;
#pragma unroll
            for (int j = 0; j < SCAN_IPT; ++j) {
                const uint32_t i = i0 + j;
                const uint32_t k = i < n ? in[i] : 0u;
                key32[j] = k;
                val[j] = (i < n && (i == 0 || k != prev)) ? 1u : 0u;
                prev = k;
            }
        } else if (MODE == MODE_UNIQUE64) {
            const uint64_t* in = static_cast<const uint64_t*>(in0);
            uint64_t prev = (i0 > 0 && i0 - 1 < n) ? in[i0 - 1] : 0ull;
#pragma unroll
            for (int j = 0; j < SCAN_IPT; ++j) {
                const uint32_t i = i0 + j;
                const uint64_t k = i < n ? in[i] : 0ull;
                key64[j] = k;
                val[j] = (i < n && (i == 0 || k != prev)) ? 1u : 0u;
                prev = k;
            }
        } else {  // MODE_ECSCAN: edge count of radix node i, root entry 1
            const uint8_t* pn = static_cast<const uint8_t*>(in0);
            const int32_t* par = static_cast<const int32_t*>(in1);
#pragma unroll
            for (int j = 0; j < SCAN_IPT; ++j) {
                const uint32_t i = i0 + j;
                uint32_t e = 0;
                if (i < n) {
                    uint32_t p = (uint32_t)par[i];
                    p = p < n ? p : 0u;  // parent indices of a valid tree are < n
                    e = (i == 0) ? 1u : (uint32_t)(pn[i] / 3 - pn[p] / 3);
                }
                val[j] = e;
            }
        }

        uint32_t tsum = 0;
#pragma unroll
        for (int j = 0; j < SCAN_IPT; ++j) tsum += val[j];
        uint32_t block_total;
        const uint32_t texcl = block_exclusive_scan<SCAN_BLOCK>(tsum, s_tmp, &block_total);

        if (wave == 0) {
            const uint32_t lane = lane_id();
            if (tile == 0) {
                if (lane == 0) {
                    st_agent64(&status[0], st_word(2u, epoch, block_total));
                    s_excl = 0;
                }
            } else {
                if (lane == 0) st_agent64(&status[tile], st_word(1u, epoch, block_total));
                const uint32_t ex = wave_lookback(status, tile, epoch);
                if (lane == 0) {
                    st_agent64(&status[tile], st_word(2u, epoch, ex + block_total));
                    s_excl = ex;
                }
            }
        }
        __syncthreads();
        const uint32_t tile_excl = s_excl;
        uint32_t run = tile_excl + texcl;

        if (MODE == MODE_SCAN) {
            uint32_t* out = static_cast<uint32_t*>(out0);
#pragma unroll
            for (int j = 0; j < SCAN_IPT; ++j) {
                if (i0 + j < n) out[i0 + j] = run;
                run += val[j];
            }
        } else if (MODE == MODE_UNIQUE32) {
            uint32_t* out = static_cast<uint32_t*>(out0);
#pragma unroll
            for (int j = 0; j < SCAN_IPT; ++j) {
                if (val[j]) out[run] = key32[j];
                run += val[j];
            }
        } else if (MODE == MODE_UNIQUE64) {
            uint64_t* out = static_cast<uint64_t*>(out0);
#pragma unroll
            for (int j = 0; j < SCAN_IPT; ++j) {
                if (val[j]) out[run] = key64[j];
                run += val[j];
            }
        } else {
            uint32_t* ec = static_cast<uint32_t*>(out0);
            uint32_t* off = static_cast<uint32_t*>(out1);
#pragma unroll
            for (int j = 0; j < SCAN_IPT; ++j) {
                if (i0 + j < n) {
                    ec[i0 + j] = val[j];
                    off[i0 + j] = run;
                }
                run += val[j];
            }
        }
        if (tile == num_tiles - 1 && tid == SCAN_BLOCK - 1) {
            const uint32_t total = tile_excl + block_total;
            if (MODE == MODE_SCAN || MODE == MODE_ECSCAN) static_cast<uint32_t*>(MODE == MODE_SCAN ? out0 : out1)[n] = total;
            if (count_out) *count_out = total;
        }
        __syncthreads();
    }

    // Exit protocol: the last workgroup out resets the counter for the next launch.
    if (tid == 0) {
        __threadfence();
        const uint32_t t = atomicAdd(&ctl->exit_ctr, 1u);
        if (t == gridDim.x - 1) {
            st_agent(&ctl->tile_ctr, 0u);
            st_agent(&ctl->exit_ctr, 0u);
            st_agent(&ctl->seq, (epoch + 1u) & 0x3FFFFFFFu);
        }
    }
}

}  // namespace

size_t tc_scan_scratch_bytes(uint32_t n) { return ((size_t)ceil_div_u32(n ? n : 1, SCAN_TILE) + 1) * 8; }

int tc_persistent_grid(tc_ctx* ctx, uint32_t work_items, int block) {
    uint32_t blocks = ceil_div_u32(work_items ? work_items : 1, (uint32_t)block);
    const uint32_t cap = (uint32_t)ctx->num_cus * 8;
    return (int)(blocks < cap ? blocks : cap);
}

int tc_launch_scan_tiles(tc_ctx* ctx, int mode, const void* in0, const void* in1, void* out0, void* out1,
                         uint32_t* count_out, uint32_t n, const uint32_t* d_n_keys, hipStream_t s) {
    const uint32_t tiles = ceil_div_u32(n ? n : 1, SCAN_TILE);
    uint64_t* status = ctx->scan_status;
    if (!status) return TC_ERR_NOMEM;
    // Status words for up to 2^30 / 4096 tiles are preallocated with the context.
    if (tiles > (1u << 18)) return TC_ERR_TOO_LARGE;
    ScanCtl* ctl = reinterpret_cast<ScanCtl*>(ctx->small + SMALL_SCAN_CTR);
    const uint32_t cap = (uint32_t)ctx->num_cus * 8;
    const uint32_t grid = tiles < cap ? tiles : cap;
    ProfScope prof(ctx, s, mode == MODE_SCAN ? K_SCAN : mode == MODE_ECSCAN ? K_ECSCAN : K_UNIQUE);
    switch (mode) {
        case MODE_SCAN:
            hipLaunchKernelGGL(k_scan_tiles<MODE_SCAN>, dim3(grid), dim3(SCAN_BLOCK), 0, s, in0, in1, out0, out1,
                               count_out, n, d_n_keys, status, ctl);
            break;
        case MODE_UNIQUE32:
            hipLaunchKernelGGL(k_scan_tiles<MODE_UNIQUE32>, dim3(grid), dim3(SCAN_BLOCK), 0, s, in0, in1, out0,
                               out1, count_out, n, d_n_keys, status, ctl);
            break;
        case MODE_UNIQUE64:
            hipLaunchKernelGGL(k_scan_tiles<MODE_UNIQUE64>, dim3(grid), dim3(SCAN_BLOCK), 0, s, in0, in1, out0,
                               out1, count_out, n, d_n_keys, status, ctl);
            break;
        case MODE_ECSCAN:
            hipLaunchKernelGGL(k_scan_tiles<MODE_ECSCAN>, dim3(grid), dim3(SCAN_BLOCK), 0, s, in0, in1, out0,
                               out1, count_out, n, d_n_keys, status, ctl);
            break;
        default:
            return TC_ERR_INVALID;
    }
    TC_LAUNCH_CHECK();
    return TC_OK;
}

extern "C" int tc_unique_u32(tc_ctx* ctx, const uint32_t* sorted, uint32_t* out, uint32_t n,
                             uint32_t* d_n_unique, tc_stream stream) {
    if (!ctx || !d_n_unique || (n && (!sorted || !out))) return TC_ERR_INVALID;
    if (n > TC_MAX_SORT_N) return TC_ERR_TOO_LARGE;
    return tc_launch_scan_tiles(ctx, MODE_UNIQUE32, sorted, nullptr, out, nullptr, d_n_unique, n, nullptr,
                                as_stream(stream));
}

extern "C" int tc_unique_u64(tc_ctx* ctx, const uint64_t* sorted, uint64_t* out, uint32_t n,
                             uint32_t* d_n_unique, tc_stream stream) {
    if (!ctx || !d_n_unique || (n && (!sorted || !out))) return TC_ERR_INVALID;
    if (n > TC_MAX_SORT_N) return TC_ERR_TOO_LARGE;
    return tc_launch_scan_tiles(ctx, MODE_UNIQUE64, sorted, nullptr, out, nullptr, d_n_unique, n, nullptr,
                                as_stream(stream));
}

extern "C" int tc_exclusive_scan_u32(tc_ctx* ctx, const uint32_t* in, uint32_t* out, uint32_t n,
                                     const uint32_t* d_n_keys, tc_stream stream) {
    if (!ctx || !out || (n && !in)) return TC_ERR_INVALID;
    if (n > TC_MAX_SORT_N) return TC_ERR_TOO_LARGE;
    return tc_launch_scan_tiles(ctx, MODE_SCAN, in, nullptr, out, nullptr, nullptr, n, d_n_keys,
                                as_stream(stream));
}

extern "C" int tc_edge_count_scan(tc_ctx* ctx, const uint8_t* prefix_n, const int32_t* parent,
                                  uint32_t* edge_count, uint32_t* offsets, uint32_t n_brt,
                                  const uint32_t* d_n_keys, tc_stream stream) {
    if (!ctx || !offsets || (n_brt && (!prefix_n || !parent || !edge_count))) return TC_ERR_INVALID;
    if (n_brt > TC_MAX_SORT_N) return TC_ERR_TOO_LARGE;
    return tc_launch_scan_tiles(ctx, MODE_ECSCAN, prefix_n, parent, edge_count, offsets, nullptr, n_brt,
                                d_n_keys, as_stream(stream));
}
