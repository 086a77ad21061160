// sort.hip -- one-pass global histogram + OneSweep LSD digit-binning passes.
//
// Reference: k_GlobalHistogram, shader/histogram.comp:103-195 and
// k_DigitBinning, shader/new_radix_sort.comp:116-314, driven by
// test_radix_sort, src/main.rs:185-800.  Published algorithm: Adinets &
// Merrill, "Onesweep: A Faster Least Significant Digit Radix Sort for GPUs"
// (2022), with Merrill & Garland's decoupled look-back.
//
// gfx950 design:
//  * digit plan: a sort of KB-bit keys (32 / 64, or 30 / 63 for Morton codes)
//    runs either 8-bit passes (the last one narrower when KB is not a multiple
//    of 8) or "wide" passes of at most 11 bits whose widths differ by at most
//    one (30-bit codes: 10, 10, 10; 32-bit keys: 11, 11, 10; 63-bit codes:
//    11, 11, 11, 10, 10, 10).  The launch shape and the plan are chosen
//    together by the autotuner.
//  * histogram: one read of the keys (16 B vector loads) builds every pass's
//    digit histogram in LDS, merged with device atomics into 32 accumulator
//    copies (same-address device atomics serialise); a one-workgroup-per-pass
//    kernel sums the copies into exclusive prefixes and clears them for the
//    next call.  The histogram kernel also clears the pass-0 look-back words
//    and the tile counters.  In the pipeline the histogram is fused into the
//    Morton kernel instead.
//  * binning pass: a tile of BLOCK*KPT keys per workgroup (tile = workgroup
//    id).  Each wave ranks its 64*KPT keys with 64-lane ballots (one ballot per
//    digit bit gives the mask of lanes holding the same digit), so the rank is
//    stable.  Per-wave digit counters live in LDS; a block scan turns them into
//    tile-local offsets; keys are staged in LDS in digit order and written out
//    in runs (coalesced stores).  Look-back words are u32 (2-bit flag, 30-bit
//    count) published and read with agent-scope atomics, since the 8 XCDs' L2
//    caches are not coherent with each other; one thread per digit (or pair of
//    digits) walks back, four predecessors per round trip, with a bounded wait
//    and a self-computed fall-back.  Each pass clears its row of the other
//    look-back region for the next pass.
#include <cstring>

#include "sort_core.hpp"

namespace {

constexpr int HIST_BLOCK = 256;
// LDS sub-histogram copies per block of the 8-bit plans.  One copy shared by the
// block's four waves measured faster than one per wave (2^26 keys, four 8-bit
// digits: 0.0895 against 0.1006 ms on Morton codes, 0.0754 against 0.0810 on
// random keys): less LDS per block, more blocks resident.
constexpr int HIST_REP = 1;
constexpr int HG_ITEMS = 4;  // 16-byte key loads in flight per thread in the histogram kernel
constexpr uint32_t HSTRIDE = 2048;  // words per pass in histogram tables and accumulators (digits <= 11 bits)
constexpr uint32_t ACC_STRIDE = 8 * HSTRIDE;  // words per accumulator copy (8 passes)
static_assert(HIST_BLOCK == 256, "bins per thread in k_hist_finish");

template <typename K> struct KeyVec;
template <> struct KeyVec<uint32_t> { using V = uint4; static constexpr int N = 4; };
template <> struct KeyVec<uint64_t> { using V = ulonglong2; static constexpr int N = 2; };

template <typename K>
TC_DEV void vec_get(const typename KeyVec<K>::V& v, K* out);
template <>
TC_DEV void vec_get<uint32_t>(const uint4& v, uint32_t* out) {
    out[0] = v.x; out[1] = v.y; out[2] = v.z; out[3] = v.w;
}
template <>
TC_DEV void vec_get<uint64_t>(const ulonglong2& v, uint64_t* out) {
    out[0] = v.x; out[1] = v.y;
}

// Which key bits each pass sorts (passes <= 8, widths <= 11).
struct DigitPlan {
    int passes;
    uint32_t shift[8];
    uint32_t width[8];
};

// 8-bit plans: 8-bit digits, the last one key_bits - 8*(passes-1) wide.  Wide
// plans (max_bits 11): ceil(key_bits / 11) passes of balanced widths.
DigitPlan make_plan(int key_bits, int max_bits) {
    DigitPlan pl{};
    pl.passes = (key_bits + max_bits - 1) / max_bits;
    uint32_t sh = 0;
    for (int p = 0; p < pl.passes; ++p) {
        uint32_t w;
        if (max_bits == 8) {
            w = p + 1 < pl.passes ? 8u : (uint32_t)(key_bits - 8 * (pl.passes - 1));
        } else {
            const int base = key_bits / pl.passes, rem = key_bits % pl.passes;
            w = (uint32_t)(base + (p < rem ? 1 : 0));
        }
        pl.shift[p] = sh;
        pl.width[p] = w;
        sh += w;
    }
    return pl;
}

// Passes of the wide plans (known at compile time for the LDS layout).
template <typename K>
constexpr int wide_passes() {
    return sizeof(K) == 4 ? 3 : 6;
}

// LDS histograms of one block.  8-bit plans: HIST_REP copies of DIG x 256 bins.  Wide plans: one copy, pass p's 2^width[p] bins at
// offset sum_{q<p} 2^width[q] (at most 11, 11, 10 for u32 keys and
// 11, 11, 11, 11, 10, 10 for u64 keys).
template <typename K, bool WIDE>
struct HistLds {
    static constexpr int NP = WIDE ? wide_passes<K>() : (int)sizeof(K);
    static constexpr uint32_t WORDS = WIDE ? (sizeof(K) == 4 ? 5120u : 10240u) : (uint32_t)(HIST_REP * sizeof(K) * 256);
};

// Adds the digits of one key to the wide plan's LDS histograms (runtime shifts
// and widths); the wave-uniform shortcut of lds_count_digits applies.
template <typename K, int NP>
TC_DEV void lds_count_plan(uint32_t* h, K key, const DigitPlan& pl, const uint32_t* loff) {
    const uint64_t active = __ballot(true);
    const bool leader = mbcnt64(active) == 0u;
    const uint32_t n_active = (uint32_t)__popcll(active);
#pragma unroll
    for (int p = 0; p < NP; ++p) {
        const uint32_t v = (uint32_t)(key >> pl.shift[p]) & ((1u << pl.width[p]) - 1u);
        const uint32_t v0 = __builtin_amdgcn_readfirstlane(v);
        if (__ballot(v == v0) == active) {
            if (leader) atomicAdd(&h[loff[p] + v0], n_active);
        } else {
            atomicAdd(&h[loff[p] + v], 1u);
        }
    }
}

template <typename K, bool WIDE>
struct HistCounter {
    uint32_t* h;
    uint32_t loff[HistLds<K, WIDE>::NP];
    TC_DEV HistCounter(uint32_t* s_h, const DigitPlan& pl) {
        if constexpr (WIDE) {
            h = s_h;
            uint32_t o = 0;
#pragma unroll
            for (int p = 0; p < HistLds<K, WIDE>::NP; ++p) {
                loff[p] = o;
                o += 1u << pl.width[p];
            }
        } else {
            h = s_h + ((threadIdx.x / 64) % HIST_REP) * (sizeof(K) * 256);
        }
    }
    TC_DEV void add(K key, const DigitPlan& pl) {
        if constexpr (WIDE) lds_count_plan<K, HistLds<K, WIDE>::NP>(h, key, pl, loff);
        else lds_count_digits<K, (int)sizeof(K)>(h, key);
    }
    // Without the wave-uniform test: one LDS atomic per digit.
    TC_DEV void add_plain(K key, const DigitPlan& pl) {
        if constexpr (WIDE) {
#pragma unroll
            for (int p = 0; p < HistLds<K, WIDE>::NP; ++p)
                atomicAdd(&h[loff[p] + ((uint32_t)(key >> pl.shift[p]) & ((1u << pl.width[p]) - 1u))], 1u);
        } else {
#pragma unroll
            for (int d = 0; d < (int)sizeof(K); ++d) atomicAdd(&h[d * 256 + (uint32_t)((key >> (8 * d)) & 255u)], 1u);
        }
    }
    // The most significant digit of the plan (the last pass's).
    TC_DEV uint32_t top_digit(K key, const DigitPlan& pl) const {
        if constexpr (WIDE) {
            constexpr int L = HistLds<K, WIDE>::NP - 1;
            return (uint32_t)(key >> pl.shift[L]) & ((1u << pl.width[L]) - 1u);
        } else {
            return (uint32_t)(key >> (8 * (sizeof(K) - 1)));
        }
    }
};

// Merge a block's LDS histograms into accumulator copy blockIdx % HIST_COPIES
// (acc_stride words per copy, pass p's bins at p * HSTRIDE).  Device atomics to
// one address serialise (one per block and bin would make a chain of gridDim.x
// adds per bin at the end of the kernel); spread over the copies the chains
// are gridDim.x / HIST_COPIES long.  k_hist_finish sums the copies.
template <typename K, bool WIDE>
TC_DEV void hist_merge(const uint32_t* s_h, uint32_t* __restrict__ acc, uint32_t acc_stride, const DigitPlan& pl) {
    uint32_t* a = acc + (blockIdx.x % HIST_COPIES) * acc_stride;
    if constexpr (WIDE) {
        uint32_t o = 0;
        for (int p = 0; p < HistLds<K, WIDE>::NP; ++p) {
            const uint32_t bins = 1u << pl.width[p];
            for (uint32_t b = threadIdx.x; b < bins; b += HIST_BLOCK) {
                const uint32_t v = s_h[o + b];
                if (v) atomicAdd(&a[p * HSTRIDE + b], v);
            }
            o += bins;
        }
    } else {
        constexpr uint32_t DIG = sizeof(K);
        for (uint32_t b = threadIdx.x; b < DIG * 256; b += HIST_BLOCK) {
            uint32_t v = 0;
#pragma unroll
            for (int r = 0; r < HIST_REP; ++r) v += s_h[r * DIG * 256 + b];
            if (v) atomicAdd(&a[(b >> 8) * HSTRIDE + (b & 255u)], v);
        }
    }
}

// One workgroup per pass: the pass's counts summed over the accumulator
// copies, written as an exclusive prefix (digit-major, the layout of
// shader/histogram.comp:103-195, pass p at p * out_stride); the copies are
// cleared for the next use.  Thread t owns bins [t*BPT, t*BPT + BPT).
// It also advances the scan epoch tag (as a scan's own one-lane epoch kernel
// would), so a pipeline whose next scan follows this sort can skip that launch.
__global__ __launch_bounds__(HIST_BLOCK) void k_hist_finish(uint32_t* __restrict__ acc, uint32_t acc_stride,
                                                            uint32_t* __restrict__ hist_out, uint32_t out_stride,
                                                            DigitPlan pl, uint32_t* __restrict__ epoch_seq) {
    __shared__ uint32_t s_tmp[HIST_BLOCK / 64 + 2];
    const uint32_t p = blockIdx.x, tid = threadIdx.x;
    if (p == 0 && tid == 0) *epoch_seq = (*epoch_seq + 1u) & 0x3FFFFFFFu;
    const uint32_t bins = 1u << pl.width[p];
    const uint32_t bpt = bins > HIST_BLOCK ? bins / HIST_BLOCK : 1u;
    const uint32_t b0 = tid * bpt;
    uint32_t c[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    uint32_t sum = 0;
    if (b0 < bins) {
        for (uint32_t j = 0; j < bpt; ++j) {
            uint32_t v = 0;
            // All the copies' loads issued together (one round trip; eight at a
            // time took four: the build 0.9888 -> 0.9853 ms).
            uint32_t part[HIST_COPIES];
#pragma unroll
            for (int r = 0; r < HIST_COPIES; ++r) part[r] = acc[r * acc_stride + p * HSTRIDE + b0 + j];
#pragma unroll
            for (int r = 0; r < HIST_COPIES; ++r) {
                v += part[r];
                acc[r * acc_stride + p * HSTRIDE + b0 + j] = 0u;
            }
            c[j] = v;
            sum += v;
        }
    }
    if (tid >= bins) {
        // A narrow pass of an 8-bit plan: its accumulator row has 256 bins.
        for (int r = 0; r < HIST_COPIES; ++r) acc[r * acc_stride + p * HSTRIDE + tid] = 0u;
    }
    uint32_t tot;
    uint32_t run = block_exclusive_scan<HIST_BLOCK>(sum, s_tmp, &tot);
    if (b0 < bins) {
        for (uint32_t j = 0; j < bpt; ++j) {
            hist_out[p * out_stride + b0 + j] = run;
            run += c[j];
        }
    }
}

template <typename K, bool WIDE>
__global__ __launch_bounds__(HIST_BLOCK) void k_global_histogram(
    const K* __restrict__ keys, uint32_t n, DigitPlan pl, uint32_t* __restrict__ acc, uint32_t acc_stride,
    uint32_t* __restrict__ zero_ptr, uint32_t zero_words) {
    constexpr int VN = KeyVec<K>::N;
    constexpr uint32_t LW = HistLds<K, WIDE>::WORDS;
    __shared__ uint32_t s_h[LW];

    const uint32_t tid = threadIdx.x;
    for (uint32_t i = tid; i < LW; i += HIST_BLOCK) s_h[i] = 0;
    __syncthreads();

    HistCounter<K, WIDE> hc(s_h, pl);
    const uint32_t gtid = blockIdx.x * HIST_BLOCK + tid;
    const uint32_t gstride = gridDim.x * HIST_BLOCK;
    const uint32_t nvec = n / VN;
    const typename KeyVec<K>::V* kv = reinterpret_cast<const typename KeyVec<K>::V*>(keys);
    // HG_ITEMS 16-byte loads in flight per thread, then the vectors' keys are counted.
    constexpr uint32_t CHUNK = HIST_BLOCK * HG_ITEMS;
    const uint32_t full = nvec / CHUNK;
    for (uint32_t c = blockIdx.x; c < full; c += gridDim.x) {
        typename KeyVec<K>::V v[HG_ITEMS];
#pragma unroll
        for (int u = 0; u < HG_ITEMS; ++u) v[u] = ld_stream(&kv[c * CHUNK + u * HIST_BLOCK + tid]);
        K k[HG_ITEMS][VN];
#pragma unroll
        for (int u = 0; u < HG_ITEMS; ++u) vec_get<K>(v[u], k[u]);
        // The wave-uniform test only when the chunk's first top digit is
        // shared by the wave (as in k_morton_hist): on keys in no particular
        // order it never pays (2^26 random u32 keys: 85.5 -> 81.5 us with the
        // grid below).
        const uint32_t top = hc.top_digit(k[0][0], pl);
        if (__ballot(top == __builtin_amdgcn_readfirstlane(top)) == __ballot(true)) {
#pragma unroll
            for (int u = 0; u < HG_ITEMS; ++u)
#pragma unroll
                for (int j = 0; j < VN; ++j) hc.add(k[u][j], pl);
        } else {
#pragma unroll
            for (int u = 0; u < HG_ITEMS; ++u)
#pragma unroll
                for (int j = 0; j < VN; ++j) hc.add_plain(k[u][j], pl);
        }
    }
    for (uint32_t v = full * CHUNK + gtid; v < nvec; v += gstride) {
        K k[VN];
        vec_get<K>(kv[v], k);
#pragma unroll
        for (int j = 0; j < VN; ++j) hc.add(k[j], pl);
    }
    for (uint32_t i = nvec * VN + gtid; i < n; i += gstride) hc.add(keys[i], pl);
    // Clear the pass-0 look-back words and the tile counters for the passes.
    for (uint32_t w = gtid; w < zero_words; w += gstride) zero_ptr[w] = 0u;
    __syncthreads();
    hist_merge<K, WIDE>(s_h, acc, acc_stride, pl);
}

template <typename K>
TC_DEV K encode_point(const float4& p, float min_coord, float range) {
    if constexpr (sizeof(K) == 4) return d_encode30(p.x, p.y, p.z, min_coord, range);
    else return d_encode63(p.x, p.y, p.z, min_coord, range);
}

// 30-bit codes with the bit spreading (d_split3_10) read from an LDS table of
// its 1024 values: one LDS read per axis instead of eight VALU operations.
TC_DEV uint32_t encode30_table(const float4& p, const MortonQuant& q, const uint32_t* s_split) {
    return s_split[q(p.x) & 1023u] | (s_split[q(p.y) & 1023u] << 1) | (s_split[q(p.z) & 1023u] << 2);
}

// Morton codes with the sort's global histogram fused in: the codes are counted
// while they are in registers, so the sort does not read them an extra time
// (k_ComputeMortonCode, shader/morton.comp:60-67 + k_GlobalHistogram,
// shader/histogram.comp:103-195).  Each thread keeps MH_ITEMS point loads in
// flight to cover HBM latency.
constexpr int MH_ITEMS = 8;

template <typename K, bool WIDE>
__global__ __launch_bounds__(HIST_BLOCK) void k_morton_hist(
    const float4* __restrict__ pts, K* __restrict__ codes, uint32_t* __restrict__ index, uint32_t n,
    float min_coord, float range, DigitPlan pl, uint32_t* __restrict__ acc, uint32_t acc_stride,
    uint32_t* __restrict__ zero_ptr, uint32_t zero_words) {
    constexpr uint32_t CHUNK = HIST_BLOCK * MH_ITEMS;
    constexpr uint32_t LW = HistLds<K, WIDE>::WORDS;
    constexpr bool TABLE = sizeof(K) == 4;
    __shared__ uint32_t s_h[LW];
    __shared__ uint32_t s_split[TABLE ? 1024 : 1];
    const uint32_t tid = threadIdx.x;
    for (uint32_t i = tid; i < LW; i += HIST_BLOCK) s_h[i] = 0;
    if constexpr (TABLE)
        for (uint32_t i = tid; i < 1024; i += HIST_BLOCK) s_split[i] = d_split3_10(i);
    __syncthreads();
    HistCounter<K, WIDE> hc(s_h, pl);
    const MortonQuant q(min_coord, range, 1024.0f);
    auto encode = [&](const float4& p) -> K {
        if constexpr (TABLE) return encode30_table(p, q, s_split);
        else return encode_point<K>(p, min_coord, range);
    };
    const uint32_t full = n / CHUNK;
    for (uint32_t c = blockIdx.x; c < full; c += gridDim.x) {
        const uint32_t b0 = c * CHUNK + tid;
        float4 p[MH_ITEMS];
#pragma unroll
        for (int u = 0; u < MH_ITEMS; ++u) p[u] = ld_stream(&pts[b0 + u * HIST_BLOCK]);
        K code[MH_ITEMS];
#pragma unroll
        for (int u = 0; u < MH_ITEMS; ++u) {
            code[u] = encode(p[u]);
            codes[b0 + u * HIST_BLOCK] = code[u];
            if (index) index[b0 + u * HIST_BLOCK] = b0 + u * HIST_BLOCK;  // sort payload: the point's index
        }
        // The wave-uniform test (one LDS atomic per wave for a digit shared by
        // all lanes) only pays for points in spatial order; it is taken for the
        // chunk when the first point's top digit is shared by the wave, else
        // every digit is counted directly (the counts are the same either way).
        const uint32_t top = hc.top_digit(code[0], pl);
        if (__ballot(top == __builtin_amdgcn_readfirstlane(top)) == __ballot(true)) {
#pragma unroll
            for (int u = 0; u < MH_ITEMS; ++u) hc.add(code[u], pl);
        } else {
#pragma unroll
            for (int u = 0; u < MH_ITEMS; ++u) hc.add_plain(code[u], pl);
        }
    }
    const uint32_t gtid = blockIdx.x * HIST_BLOCK + tid;
    const uint32_t gstride = gridDim.x * HIST_BLOCK;
    for (uint32_t i = full * CHUNK + gtid; i < n; i += gstride) {
        const K code = encode(pts[i]);
        codes[i] = code;
        if (index) index[i] = i;
        hc.add(code, pl);
    }
    for (uint32_t w = gtid; w < zero_words; w += gstride) zero_ptr[w] = 0u;
    __syncthreads();
    hist_merge<K, WIDE>(s_h, acc, acc_stride, pl);
}

template <typename K, bool PAIRS, int BLOCK, int KPT, int RB>
__global__ __launch_bounds__(BLOCK) void k_onesweep(
    const K* __restrict__ sk, K* __restrict__ dk, const uint32_t* __restrict__ sv,
    uint32_t* __restrict__ dv, uint32_t n, uint32_t shift, const uint32_t* __restrict__ digit_prefix,
    uint32_t* __restrict__ lb_cur, uint32_t* __restrict__ lb_next, uint32_t next_row, uint32_t spin_limit) {
    using Sh = OnesweepShared<K, BLOCK, KPT, RB>;
    __shared__ Sh sm;
    const uint32_t num_tiles = ceil_div_u32(n, Sh::TILE);
    // One tile per workgroup, tile = workgroup id: no claim counter (claims on
    // one counter serialise at about 11 ns each on MI355X, a floor of 0.19 ms
    // for 16384 tiles).  The look-back's bounded wait with its self-computed
    // fall-back keeps this correct and live whatever order workgroups start in.
    const uint32_t tile = blockIdx.x;
    if (tile >= num_tiles) return;
    const uint32_t base = tile * Sh::TILE;
    const uint32_t cnt = n - base < Sh::TILE ? n - base : Sh::TILE;
    onesweep_tile<K, PAIRS, BLOCK, KPT, RB, !PAIRS>(sm, sk, dk, sv, dv, base, cnt, shift, tile, 0u, 0u, digit_prefix,
                                                    0u, lb_cur, lb_next, next_row, spin_limit);
}

// Copy of a sort result back into the caller's buffer (plans with an odd number
// of passes end in the ping-pong buffer): 16 B vectors when both ends are
// 16-byte aligned, words otherwise.
__global__ __launch_bounds__(256) void k_copy_vec(const uint4* __restrict__ src, uint4* __restrict__ dst, size_t nvec,
                                                  const uint32_t* __restrict__ tail_src, uint32_t* __restrict__ tail_dst,
                                                  uint32_t tail) {
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < nvec; i += (size_t)gridDim.x * 256)
        dst[i] = ld_stream(&src[i]);
    if (blockIdx.x == 0 && threadIdx.x < tail) tail_dst[threadIdx.x] = tail_src[threadIdx.x];
}

__global__ __launch_bounds__(256) void k_copy_words(const uint32_t* __restrict__ src, uint32_t* __restrict__ dst,
                                                    size_t n) {
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) dst[i] = src[i];
}

int copy_words(tc_ctx* ctx, const void* src, void* dst, size_t words, hipStream_t s) {
    if (words == 0) return TC_OK;
    ProfScope prof(ctx, s, K_FILL);
    const size_t cap = (size_t)ctx->num_cus * 8;
    const uint32_t* s32 = static_cast<const uint32_t*>(src);
    uint32_t* d32 = static_cast<uint32_t*>(dst);
    if (((reinterpret_cast<uintptr_t>(src) | reinterpret_cast<uintptr_t>(dst)) & 15u) == 0) {
        const size_t nvec = words / 4;
        const uint32_t tail = (uint32_t)(words % 4);
        size_t blocks = (nvec + 255) / 256;
        if (blocks > cap) blocks = cap;
        if (blocks == 0) blocks = 1;
        hipLaunchKernelGGL(k_copy_vec, dim3((uint32_t)blocks), dim3(256), 0, s, reinterpret_cast<const uint4*>(src),
                           reinterpret_cast<uint4*>(dst), nvec, s32 + nvec * 4, d32 + nvec * 4, tail);
    } else {
        size_t blocks = (words + 255) / 256;
        if (blocks > cap) blocks = cap;
        hipLaunchKernelGGL(k_copy_words, dim3((uint32_t)blocks), dim3(256), 0, s, s32, d32, words);
    }
    TC_LAUNCH_CHECK();
    return TC_OK;
}

// Look-back row stride (words) of a pass of the given width.
uint32_t row_words(uint32_t width) { return width < 8 ? 256u : (1u << width); }

template <typename K, bool PAIRS, int BLOCK, int KPT>
int launch_pass(uint32_t tiles, hipStream_t s, const K* sk, K* dk, const uint32_t* sv, uint32_t* dv, uint32_t n,
                uint32_t shift, uint32_t width, const uint32_t* prefix, uint32_t* lb_cur, uint32_t* lb_next,
                uint32_t next_row, uint32_t spin_limit) {
#define TC_RB_CASE(RB)                                                                                     \
    if (width == RB) {                                                                                     \
        hipLaunchKernelGGL((k_onesweep<K, PAIRS, BLOCK, KPT, RB>), dim3(tiles), dim3(BLOCK), 0, s, sk, dk, sv, dv, \
                           n, shift, prefix, lb_cur, lb_next, next_row, spin_limit);                      \
        return TC_OK;                                                                                      \
    }
    if constexpr (BLOCK >= 1024) {
        TC_RB_CASE(11)
        TC_RB_CASE(10)
    }
    TC_RB_CASE(8)
    if constexpr (sizeof(K) == 4) {
        TC_RB_CASE(6)  // top digit of 30-bit Morton codes
    } else {
        TC_RB_CASE(7)  // top digit of 63-bit Morton codes
    }
#undef TC_RB_CASE
    return TC_ERR_INVALID;
}

template <typename K, bool PAIRS>
int dispatch_pass(tc_ctx* ctx, SortShape sh, uint32_t tiles, hipStream_t s, const K* sk, K* dk,
                  const uint32_t* sv, uint32_t* dv, uint32_t n, uint32_t shift, uint32_t width,
                  const uint32_t* prefix, uint32_t* lb_cur, uint32_t* lb_next, uint32_t next_row) {
    ProfScope prof(ctx, s, K_BINNING);
#define TC_PASS_CASE(B, KP)                                                                                    \
    if (sh.block == B && sh.kpt == KP) {                                                                       \
        TC_TRY((launch_pass<K, PAIRS, B, KP>(tiles, s, sk, dk, sv, dv, n, shift, width, prefix, lb_cur, lb_next, \
                                             next_row, ctx->spin_limit)));                                    \
        TC_LAUNCH_CHECK();                                                                                     \
        return TC_OK;                                                                                          \
    }
    if (sh.bits == 8) {
        TC_PASS_CASE(256, 8)
        TC_PASS_CASE(512, 8)
        TC_PASS_CASE(512, 12)
        TC_PASS_CASE(512, 16)
        TC_PASS_CASE(512, 20)
        TC_PASS_CASE(512, 24)
        TC_PASS_CASE(512, 28)
        TC_PASS_CASE(512, 32)
        TC_PASS_CASE(512, 36)
        TC_PASS_CASE(1024, 16)
    } else {
        TC_PASS_CASE(1024, 8)
        TC_PASS_CASE(1024, 16)
    }
#undef TC_PASS_CASE
    return TC_ERR_INVALID;
}

// Histogram grids: 8 workgroups per CU for u32 keys (2^26 keys: 88.5 -> 86.7 us
// on the headline's Morton codes, 85.5 -> 81.5 us on random keys, with the
// per-chunk uniform test; 32 per CU before, when every key took the test) and
// 20 for u64 keys (unchanged).  The merge's atomic chains stay gridDim /
// HIST_COPIES long.
template <typename K>
uint32_t hist_grid(tc_ctx* ctx, uint32_t n) {
    uint32_t blocks = ceil_div_u32(n, HIST_BLOCK * 16);
    uint32_t cap = (uint32_t)ctx->num_cus * (sizeof(K) == 4 ? 8u : 20u);
    if (blocks < 1) blocks = 1;
    return blocks < cap ? blocks : cap;
}

// The Morton + histogram kernel over 30-bit codes: 8 workgroups per CU (2^26
// Gaussian-mixture points: 229.6 us against 242.9 us at 32 per CU, which also
// merges four times as many LDS histograms at the end).
template <typename K>
uint32_t morton_hist_grid(tc_ctx* ctx, uint32_t n) {
    if constexpr (sizeof(K) != 4) return hist_grid<K>(ctx, n);
    uint32_t blocks = ceil_div_u32(n, HIST_BLOCK * 16);
    const uint32_t cap = (uint32_t)ctx->num_cus * 8u;
    if (blocks < 1) blocks = 1;
    return blocks < cap ? blocks : cap;
}

// Scratch of one sort: [64 tile counters][region 0][region 1][hist], plus the
// histogram accumulators (zero between uses, acc_stride words per copy) and the
// scan epoch word that the histogram-finish kernel advances.
struct SortScratch {
    uint32_t* ctr;
    uint32_t* region[2];
    uint32_t* hist;  // passes x HSTRIDE
    uint32_t* acc;
    uint32_t acc_stride;
    uint32_t* epoch;
};

uint32_t plan_row_max(const DigitPlan& pl) {
    uint32_t r = 256;
    for (int p = 0; p < pl.passes; ++p) r = row_words(pl.width[p]) > r ? row_words(pl.width[p]) : r;
    return r;
}

size_t scratch_words(uint32_t tiles, uint32_t row_max) { return 64 + 2 * (size_t)tiles * row_max + 8 * HSTRIDE; }

int sort_scratch(tc_ctx* ctx, uint32_t tiles, uint32_t row_max, SortScratch* ss) {
    const size_t region_words = (size_t)tiles * row_max;
    TC_TRY(tc_ensure_scratch(ctx, scratch_words(tiles, row_max) * 4));
    uint32_t* base = static_cast<uint32_t*>(ctx->scratch);
    ss->ctr = base;
    ss->region[0] = base + 64;
    ss->region[1] = base + 64 + region_words;
    ss->hist = base + 64 + 2 * region_words;
    ss->acc = ctx->hist_acc;
    ss->acc_stride = ACC_STRIDE;
    ss->epoch = tc_scan_epoch(ctx);
    return TC_OK;
}

template <typename K>
int histogram_launch(tc_ctx* ctx, const K* keys, uint32_t n, const DigitPlan& pl, bool wide, uint32_t* acc,
                     uint32_t acc_stride, uint32_t* zero_ptr, uint32_t zero_words, hipStream_t s) {
    ProfScope prof(ctx, s, K_HIST);
    if (wide)
        hipLaunchKernelGGL((k_global_histogram<K, true>), dim3(hist_grid<K>(ctx, n)), dim3(HIST_BLOCK), 0, s, keys, n,
                           pl, acc, acc_stride, zero_ptr, zero_words);
    else
        hipLaunchKernelGGL((k_global_histogram<K, false>), dim3(hist_grid<K>(ctx, n)), dim3(HIST_BLOCK), 0, s, keys,
                           n, pl, acc, acc_stride, zero_ptr, zero_words);
    TC_LAUNCH_CHECK();
    return TC_OK;
}

int hist_finish(tc_ctx* ctx, const DigitPlan& pl, uint32_t* acc, uint32_t acc_stride, uint32_t* out,
                uint32_t out_stride, uint32_t* epoch, hipStream_t s) {
    hipLaunchKernelGGL(k_hist_finish, dim3(pl.passes), dim3(HIST_BLOCK), 0, s, acc, acc_stride, out, out_stride, pl,
                       epoch);
    TC_LAUNCH_CHECK();
    return TC_OK;
}

// The launches of one sort of key_bits-bit keys with the plan of shape sh
// (histogram, then one binning pass per digit) on the scratch `ss` (tiles
// look-back rows of up to plan_row_max words per region).  With pts != NULL the
// keys are first computed from the points by the fused Morton + histogram
// kernel, which also sets vals[i] = i when vals is given (the sorted values are
// then the permutation from sorted position to point index).  The result ends
// in keys (and vals): a plan with an odd number of passes has the Morton
// kernel write into the ping-pong buffers, or copies plain keys back at the end.
template <typename K>
int sort_run(tc_ctx* ctx, SortShape sh, const SortScratch& ss, uint32_t tiles, K* keys, uint32_t* vals, K* alt,
             uint32_t* valt, uint32_t n, int key_bits, hipStream_t s, const float* pts, float min_coord,
             float range) {
    const DigitPlan pl = make_plan(key_bits, sh.bits);
    const bool wide = sh.bits != 8;
    if (pl.passes > 8 || (wide && pl.passes != wide_passes<K>())) return TC_ERR_INVALID;
    const bool odd = (pl.passes & 1) != 0;
    // The histogram kernel also clears the tile counters and look-back region 0,
    // as one range: the region must follow the 64 counters.
    if (ss.region[0] != ss.ctr + 64) return TC_ERR_INVALID;
    const uint32_t zero_words = 64 + tiles * row_words(pl.width[0]);
    // Every n > 0 sort launches k_hist_finish (both branches below), which
    // advances the scan epoch: tc_build_tree's unique scan relies on that to
    // skip its own epoch kernel.
    K* src = keys;
    K* dst = alt;
    uint32_t* vsrc = vals;
    uint32_t* vdst = valt;
    if (pts) {
        if (odd) {  // codes (and indices) go to the ping-pong buffers, so the last pass ends in keys
            src = alt;
            dst = keys;
            vsrc = valt;
            vdst = vals;
        }
        ProfScope prof(ctx, s, K_MORTON);
        if (wide)
            hipLaunchKernelGGL((k_morton_hist<K, true>), dim3(morton_hist_grid<K>(ctx, n)), dim3(HIST_BLOCK), 0, s,
                               reinterpret_cast<const float4*>(pts), src, vsrc, n, min_coord, range, pl, ss.acc,
                               ss.acc_stride, ss.ctr, zero_words);
        else
            hipLaunchKernelGGL((k_morton_hist<K, false>), dim3(morton_hist_grid<K>(ctx, n)), dim3(HIST_BLOCK), 0, s,
                               reinterpret_cast<const float4*>(pts), src, vsrc, n, min_coord, range, pl, ss.acc,
                               ss.acc_stride, ss.ctr, zero_words);
        TC_LAUNCH_CHECK();
    } else {
        TC_TRY(histogram_launch<K>(ctx, keys, n, pl, wide, ss.acc, ss.acc_stride, ss.ctr, zero_words, s));
    }
    TC_TRY(hist_finish(ctx, pl, ss.acc, ss.acc_stride, ss.hist, HSTRIDE, ss.epoch, s));
    for (int p = 0; p < pl.passes; ++p) {
        uint32_t* cur = ss.region[p & 1];
        uint32_t* nxt = (p + 1 < pl.passes) ? ss.region[(p + 1) & 1] : nullptr;
        const uint32_t next_row = (p + 1 < pl.passes) ? row_words(pl.width[p + 1]) : 0u;
        const uint32_t* prefix = ss.hist + (size_t)p * HSTRIDE;
        if (vals) {
            TC_TRY((dispatch_pass<K, true>(ctx, sh, tiles, s, src, dst, vsrc, vdst, n, pl.shift[p], pl.width[p],
                                           prefix, cur, nxt, next_row)));
        } else {
            TC_TRY((dispatch_pass<K, false>(ctx, sh, tiles, s, src, dst, nullptr, nullptr, n, pl.shift[p],
                                            pl.width[p], prefix, cur, nxt, next_row)));
        }
        K* t = src; src = dst; dst = t;
        uint32_t* tv = vsrc; vsrc = vdst; vdst = tv;
    }
    if (src != keys) {  // odd plan on keys given in place: copy the result back
        TC_TRY(copy_words(ctx, src, keys, (size_t)n * sizeof(K) / 4, s));
        if (vals) TC_TRY(copy_words(ctx, vsrc, vals, n, s));
    }
    return TC_OK;
}

int plain_key_bits(int key_bytes) { return key_bytes == 4 ? 32 : 64; }
int morton_key_bits(int key_bytes) { return key_bytes == 4 ? 30 : 63; }

// Sort keys (and values) in place on the context's scratch.
template <typename K>
int sort_impl(tc_ctx* ctx, K* keys, uint32_t* vals, K* alt, uint32_t* valt, uint32_t n, hipStream_t s,
              const float* pts = nullptr, float min_coord = 0.f, float range = 1.f) {
    if (!ctx || (n && (!keys || !alt))) return TC_ERR_INVALID;
    if (vals && !valt) return TC_ERR_INVALID;
    if (n > TC_MAX_SORT_N) return TC_ERR_TOO_LARGE;
    if (tc_overlaps(keys, (size_t)n * sizeof(K), alt, (size_t)n * sizeof(K))) return TC_ERR_INVALID;
    if (vals && tc_overlaps(vals, (size_t)n * 4, valt, (size_t)n * 4)) return TC_ERR_INVALID;
    if (n == 0) return TC_OK;
    if (reinterpret_cast<uintptr_t>(keys) & 15u) return TC_ERR_INVALID;
    const int kb = pts ? morton_key_bits((int)sizeof(K)) : plain_key_bits((int)sizeof(K));
    const SortShape sh = tc_sort_shape(ctx, (int)sizeof(K), vals != nullptr, n);
    const uint32_t tiles = ceil_div_u32(n, (uint32_t)(sh.block * sh.kpt));
    SortScratch ss;
    TC_TRY(sort_scratch(ctx, tiles, plan_row_max(make_plan(kb, sh.bits)), &ss));
    return sort_run<K>(ctx, sh, ss, tiles, keys, vals, alt, valt, n, kb, s, pts, min_coord, range);
}

// The 8-bit shape used where only 8-bit passes are supported (the single-pass
// API, the segmented batch sort, the no-alloc sort): the shape itself if it is
// one, else the 8-bit shape of the same tile size, else 512 x 16.
SortShape eight_bit_shape(SortShape sh) {
    if (sh.bits == 8) return sh;
    for (const SortShape& c : SORT_SHAPES)
        if (c.bits == 8 && c.block * c.kpt == sh.block * sh.kpt) return c;
    return SortShape{512, 16, 8};
}

template <typename K>
int binning_pass_impl(tc_ctx* ctx, const K* sk, K* dk, const uint32_t* sv, uint32_t* dv, uint32_t n,
                      const uint32_t* hist, uint32_t shift, hipStream_t s) {
    if (!ctx || (n && (!sk || !dk || !hist))) return TC_ERR_INVALID;
    if ((sv == nullptr) != (dv == nullptr)) return TC_ERR_INVALID;
    if (shift % 8 != 0 || shift >= 8 * sizeof(K)) return TC_ERR_INVALID;
    if (n > TC_MAX_SORT_N) return TC_ERR_TOO_LARGE;
    if (tc_overlaps(sk, (size_t)n * sizeof(K), dk, (size_t)n * sizeof(K))) return TC_ERR_INVALID;
    if (sv && tc_overlaps(sv, (size_t)n * 4, dv, (size_t)n * 4)) return TC_ERR_INVALID;
    if (n == 0) return TC_OK;
    const SortShape sh = eight_bit_shape(tc_sort_shape(ctx, (int)sizeof(K), sv != nullptr, n));
    const uint32_t tiles = ceil_div_u32(n, (uint32_t)(sh.block * sh.kpt));
    SortScratch ss;
    TC_TRY(sort_scratch(ctx, tiles, 256, &ss));
    TC_TRY(tc_fill_u32(ctx, ss.ctr, 0u, 64 + (size_t)tiles * 256, s));
    const uint32_t* prefix = hist + 256 * (shift / 8);
    if (sv) return dispatch_pass<K, true>(ctx, sh, tiles, s, sk, dk, sv, dv, n, shift, 8, prefix, ss.region[0],
                                          nullptr, 0);
    return dispatch_pass<K, false>(ctx, sh, tiles, s, sk, dk, nullptr, nullptr, n, shift, 8, prefix, ss.region[0],
                                   nullptr, 0);
}

// The reference's histogram layout: 8-bit digits, hist[256 * d + b].
template <typename K>
int global_histogram_api(tc_ctx* ctx, const K* keys, uint32_t n, uint32_t* hist, hipStream_t s) {
    if (!ctx || !hist || (n && !keys)) return TC_ERR_INVALID;
    if (reinterpret_cast<uintptr_t>(keys) & 15u) return TC_ERR_INVALID;
    const DigitPlan pl = make_plan(plain_key_bits((int)sizeof(K)), 8);
    TC_TRY(histogram_launch<K>(ctx, keys, n, pl, false, ctx->hist_acc, ACC_STRIDE, nullptr, 0, s));
    return hist_finish(ctx, pl, ctx->hist_acc, ACC_STRIDE, hist, 256, tc_scan_epoch(ctx), s);
}

}  // namespace

size_t tc_sort_scratch_max_bytes(uint32_t n) {
    size_t words = 0;
    for (const SortShape& sh : SORT_SHAPES) {
        const uint32_t tiles = ceil_div_u32(n ? n : 1, (uint32_t)(sh.block * sh.kpt));
        const uint32_t row = sh.bits == 8 ? 256u : 2048u;
        const size_t w = scratch_words(tiles, row);
        words = w > words ? w : words;
    }
    return words * 4;
}

size_t tc_hist_acc_bytes() { return (size_t)HIST_COPIES * ACC_STRIDE * 4; }

extern "C" int tc_global_histogram_u32(tc_ctx* ctx, const uint32_t* keys, uint32_t n, uint32_t* hist,
                                       tc_stream stream) {
    return global_histogram_api<uint32_t>(ctx, keys, n, hist, as_stream(stream));
}
extern "C" int tc_global_histogram_u64(tc_ctx* ctx, const uint64_t* keys, uint32_t n, uint32_t* hist,
                                       tc_stream stream) {
    return global_histogram_api<uint64_t>(ctx, keys, n, hist, as_stream(stream));
}
extern "C" int tc_binning_pass_u32(tc_ctx* ctx, const uint32_t* sk, uint32_t* dk, const uint32_t* sv,
                                   uint32_t* dv, uint32_t n, const uint32_t* hist, uint32_t shift,
                                   tc_stream stream) {
    return binning_pass_impl<uint32_t>(ctx, sk, dk, sv, dv, n, hist, shift, as_stream(stream));
}
extern "C" int tc_binning_pass_u64(tc_ctx* ctx, const uint64_t* sk, uint64_t* dk, const uint32_t* sv,
                                   uint32_t* dv, uint32_t n, const uint32_t* hist, uint32_t shift,
                                   tc_stream stream) {
    return binning_pass_impl<uint64_t>(ctx, sk, dk, sv, dv, n, hist, shift, as_stream(stream));
}
extern "C" int tc_sort_keys_u32(tc_ctx* ctx, uint32_t* keys, uint32_t* alt, uint32_t n, tc_stream stream) {
    return sort_impl<uint32_t>(ctx, keys, nullptr, alt, nullptr, n, as_stream(stream));
}
extern "C" int tc_sort_pairs_u32(tc_ctx* ctx, uint32_t* keys, uint32_t* vals, uint32_t* alt, uint32_t* valt,
                                 uint32_t n, tc_stream stream) {
    if (!vals) return TC_ERR_INVALID;
    return sort_impl<uint32_t>(ctx, keys, vals, alt, valt, n, as_stream(stream));
}
extern "C" int tc_morton_sort_u32(tc_ctx* ctx, const float* pts, uint32_t* keys, uint32_t* alt, uint32_t n,
                                  float min_coord, float range, tc_stream stream) {
    if (n && (!pts || (reinterpret_cast<uintptr_t>(pts) & 15u))) return TC_ERR_INVALID;
    return sort_impl<uint32_t>(ctx, keys, nullptr, alt, nullptr, n, as_stream(stream), pts, min_coord, range);
}
extern "C" int tc_morton_sort_pairs_u32(tc_ctx* ctx, const float* pts, uint32_t* keys, uint32_t* index,
                                        uint32_t* keys_alt, uint32_t* index_alt, uint32_t n, float min_coord,
                                        float range, tc_stream stream) {
    if (n && (!pts || !index || (reinterpret_cast<uintptr_t>(pts) & 15u))) return TC_ERR_INVALID;
    return sort_impl<uint32_t>(ctx, keys, index, keys_alt, index_alt, n, as_stream(stream), pts, min_coord, range);
}
extern "C" int tc_morton_sort_pairs_u64(tc_ctx* ctx, const float* pts, uint64_t* keys, uint32_t* index,
                                        uint64_t* keys_alt, uint32_t* index_alt, uint32_t n, float min_coord,
                                        float range, tc_stream stream) {
    if (n && (!pts || !index || (reinterpret_cast<uintptr_t>(pts) & 15u))) return TC_ERR_INVALID;
    return sort_impl<uint64_t>(ctx, keys, index, keys_alt, index_alt, n, as_stream(stream), pts, min_coord, range);
}
extern "C" int tc_morton_sort_u64(tc_ctx* ctx, const float* pts, uint64_t* keys, uint64_t* alt, uint32_t n,
                                  float min_coord, float range, tc_stream stream) {
    if (n && (!pts || (reinterpret_cast<uintptr_t>(pts) & 15u))) return TC_ERR_INVALID;
    return sort_impl<uint64_t>(ctx, keys, nullptr, alt, nullptr, n, as_stream(stream), pts, min_coord, range);
}
extern "C" int tc_sort_keys_u64(tc_ctx* ctx, uint64_t* keys, uint64_t* alt, uint32_t n, tc_stream stream) {
    return sort_impl<uint64_t>(ctx, keys, nullptr, alt, nullptr, n, as_stream(stream));
}
extern "C" int tc_sort_pairs_u64(tc_ctx* ctx, uint64_t* keys, uint32_t* vals, uint64_t* alt, uint32_t* valt,
                                 uint32_t n, tc_stream stream) {
    if (!vals) return TC_ERR_INVALID;
    return sort_impl<uint64_t>(ctx, keys, vals, alt, valt, n, as_stream(stream));
}

// ------------------------------------------------- sort that allocates nothing
// tc_sort_keys_u32_noalloc keeps its scratch in one file-scope device array
// (each GPU holds its own copy, at its own address, until the process ends):
// tile counters, two look-back regions sized for 2^24 keys in tiles of at
// least 8192 keys with 8-bit digits, the histogram, the scan epoch word and
// the histogram accumulators.  Host code reaches it only through the address
// hipGetSymbolAddress returns on each device (tc_ctx::noalloc_scratch, set by
// tc_create with the context's device current); the array's name itself is a
// host-side stand-in that must never be passed to a kernel.
namespace {

constexpr uint32_t NA_MAX_N = 1u << 24;
constexpr uint32_t NA_MIN_TILE = 8192;
constexpr size_t NA_REGION_WORDS = (size_t)(NA_MAX_N / NA_MIN_TILE) * 256;
constexpr uint32_t NA_ACC_STRIDE = 4 * HSTRIDE;  // u32 keys: 4 passes
// The sort's layout: the tile counters directly followed by look-back region 0
// (the histogram kernel clears both as one range).
constexpr size_t NA_CTR = 0;
constexpr size_t NA_R0 = 64;
constexpr size_t NA_R1 = NA_R0 + NA_REGION_WORDS;
constexpr size_t NA_HIST = NA_R1 + NA_REGION_WORDS;
constexpr size_t NA_EPOCH = NA_HIST + 4 * HSTRIDE;
constexpr size_t NA_ACC = NA_EPOCH + 64;
constexpr size_t NA_WORDS = NA_ACC + (size_t)HIST_COPIES * NA_ACC_STRIDE;  // 1,319,040 words = 5.03 MiB

static __device__ __align__(256) uint32_t scratch[NA_WORDS];

}  // namespace

int tc_noalloc_init(tc_ctx* ctx) {
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) {
        (void)hipGetLastError();
        ctx->noalloc_scratch = nullptr;
        return TC_ERR_HIP;
    }
    ctx->noalloc_scratch = static_cast<uint32_t*>(ptr);
    return TC_OK;
}

extern "C" int tc_sort_keys_u32_noalloc(tc_ctx* ctx, uint32_t* keys, uint32_t* alt, uint32_t n, tc_stream stream) {
    if (!ctx || (n && (!keys || !alt))) return TC_ERR_INVALID;
    if (n > NA_MAX_N) return TC_ERR_TOO_LARGE;
    if (tc_overlaps(keys, (size_t)n * 4, alt, (size_t)n * 4)) return TC_ERR_INVALID;
    if (n == 0) return TC_OK;
    if (reinterpret_cast<uintptr_t>(keys) & 15u) return TC_ERR_INVALID;
    uint32_t* base = ctx->noalloc_scratch;
    if (!base) return TC_ERR_HIP;
    hipStream_t s = as_stream(stream);
    // The tuned shape (as an 8-bit plan) unless its tiles are too small for
    // the static regions; then the first built 8-bit shape whose tiles are
    // (every n <= 2^24 fits those).
    SortShape sh = eight_bit_shape(tc_sort_shape(ctx, 4, false, n));
    if ((size_t)ceil_div_u32(n, (uint32_t)(sh.block * sh.kpt)) * 256 > NA_REGION_WORDS) {
        for (const SortShape& c : SORT_SHAPES) {
            if (c.bits == 8 && (uint32_t)(c.block * c.kpt) >= NA_MIN_TILE) {
                sh = c;
                break;
            }
        }
    }
    const uint32_t tiles = ceil_div_u32(n, (uint32_t)(sh.block * sh.kpt));
    SortScratch ss{base + NA_CTR, {base + NA_R0, base + NA_R1}, base + NA_HIST, base + NA_ACC, NA_ACC_STRIDE,
                   base + NA_EPOCH};
    // The accumulators must start at zero: every sort leaves them so, but a
    // call that failed part-way may not have.
    TC_TRY(tc_fill_u32(ctx, ss.acc, 0u, (size_t)HIST_COPIES * NA_ACC_STRIDE, s));
    return sort_run<uint32_t>(ctx, sh, ss, tiles, keys, nullptr, alt, nullptr, n, 32, s, nullptr, 0.f, 1.f);
}

// ------------------------------------------------------------------ autotuner
namespace {

__global__ void k_random_keys(uint32_t* __restrict__ out, size_t words, uint32_t seed) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < words; i += (size_t)gridDim.x * blockDim.x) {
        uint32_t x = (uint32_t)i * 0x9E3779B9u ^ seed;
        x ^= x >> 16;
        x *= 0x7FEB352Du;
        x ^= x >> 15;
        x *= 0x846CA68Bu;
        x ^= x >> 16;
        out[i] = x;
    }
}

int log2_bucket(uint32_t n) {
    int b = 0;
    while (b < 31 && (1u << b) < n) ++b;
    return b;
}

int shape_index(int block, int kpt, int bits) {
    for (int i = 0; i < N_SORT_SHAPES; ++i)
        if (SORT_SHAPES[i].block == block && SORT_SHAPES[i].kpt == kpt && SORT_SHAPES[i].bits == bits) return i;
    return -1;
}

}  // namespace

SortShape tc_sort_shape(const tc_ctx* ctx, int key_bytes, bool pairs, uint32_t n) {
    if (ctx->tune_on) {
        const int kb = key_bytes == 8 ? 1 : 0, pv = pairs ? 1 : 0;
        const int b = log2_bucket(n);
        for (int d = 0; d < 32; ++d) {
            for (int x : {b - d, b + d}) {
                if (x < 0 || x > 31) continue;
                const int s = ctx->tune[kb][pv][x];
                if (s >= 0 && s < N_SORT_SHAPES) return SORT_SHAPES[s];
            }
        }
    }
    return SortShape{ctx->sort_block, ctx->sort_kpt, ctx->sort_bits};
}

SortShape tc_sort_shape_8bit(const tc_ctx* ctx, int key_bytes, bool pairs, uint32_t n) {
    return eight_bit_shape(tc_sort_shape(ctx, key_bytes, pairs, n));
}

// Time every launch shape and digit plan on random keys at sizes 2^12, 2^14,
// ... up to max_n, for u32 and u64 keys, with and without a u32 payload, and
// keep the fastest per size.  Blocking; uses its own stream and buffers.
extern "C" int tc_autotune_sort(tc_ctx* ctx, uint32_t max_n) {
    if (!ctx || max_n < (1u << 12)) return TC_ERR_INVALID;
    if (max_n > TC_MAX_SORT_N) return TC_ERR_TOO_LARGE;
    TC_HIP(hipSetDevice(ctx->device));
    hipStream_t s = nullptr;
    TC_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    const size_t n = max_n;
    void *k0 = nullptr, *k = nullptr, *ka = nullptr, *v = nullptr, *va = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int st = TC_OK;
    if (hipMalloc(&k0, n * 8) != hipSuccess || hipMalloc(&k, n * 8) != hipSuccess ||
        hipMalloc(&ka, n * 8) != hipSuccess || hipMalloc(&v, n * 4) != hipSuccess ||
        hipMalloc(&va, n * 4) != hipSuccess || hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)
        st = TC_ERR_NOMEM;
    if (st == TC_OK) {
        hipLaunchKernelGGL(k_random_keys, dim3(ctx->num_cus * 4), dim3(256), 0, s, static_cast<uint32_t*>(k0), n * 2,
                           0x5EEDu);
        if (hipGetLastError() != hipSuccess) st = TC_ERR_HIP;
    }
    const bool prof = ctx->prof_on;
    ctx->prof_on = false;
    const int saved_block = ctx->sort_block, saved_kpt = ctx->sort_kpt, saved_bits = ctx->sort_bits;
    const bool saved_on = ctx->tune_on;
    ctx->tune_on = false;
    int8_t table[2][2][32];
    std::memset(table, -1, sizeof(table));
    for (int kb = 0; kb < 2 && st == TC_OK; ++kb) {
        for (int pv = 0; pv < 2 && st == TC_OK; ++pv) {
            for (int lg = 12; lg <= 30 && (1u << lg) <= max_n && st == TC_OK; lg += 2) {
                const uint32_t m = 1u << lg;
                float best = 1e30f;
                int best_s = -1;
                for (int sh = 0; sh < N_SORT_SHAPES && st == TC_OK; ++sh) {
                    ctx->sort_block = SORT_SHAPES[sh].block;
                    ctx->sort_kpt = SORT_SHAPES[sh].kpt;
                    ctx->sort_bits = SORT_SHAPES[sh].bits;
                    float t_min = 1e30f;
                    for (int rep = 0; rep < 4 && st == TC_OK; ++rep) {
                        const size_t kbytes = (size_t)m * (kb ? 8 : 4);
                        if (hipMemcpyAsync(k, k0, kbytes, hipMemcpyDeviceToDevice, s) != hipSuccess) st = TC_ERR_HIP;
                        if (st != TC_OK) break;
                        (void)hipEventRecord(e0, s);
                        if (kb == 0)
                            st = pv ? tc_sort_pairs_u32(ctx, (uint32_t*)k, (uint32_t*)v, (uint32_t*)ka, (uint32_t*)va, m, s)
                                    : tc_sort_keys_u32(ctx, (uint32_t*)k, (uint32_t*)ka, m, s);
                        else
                            st = pv ? tc_sort_pairs_u64(ctx, (uint64_t*)k, (uint32_t*)v, (uint64_t*)ka, (uint32_t*)va, m, s)
                                    : tc_sort_keys_u64(ctx, (uint64_t*)k, (uint64_t*)ka, m, s);
                        (void)hipEventRecord(e1, s);
                        if (st == TC_OK && hipEventSynchronize(e1) != hipSuccess) st = TC_ERR_HIP;
                        float ms = 0.f;
                        if (st == TC_OK && hipEventElapsedTime(&ms, e0, e1) != hipSuccess) st = TC_ERR_HIP;
                        if (rep > 0 && ms < t_min) t_min = ms;  // rep 0 warms up
                    }
                    if (t_min < best) {
                        best = t_min;
                        best_s = sh;
                    }
                }
                table[kb][pv][lg] = (int8_t)best_s;
            }
        }
    }
    ctx->prof_on = prof;
    ctx->sort_block = saved_block;
    ctx->sort_kpt = saved_kpt;
    ctx->sort_bits = saved_bits;
    if (st == TC_OK) {
        std::memcpy(ctx->tune, table, sizeof(table));
        ctx->tune_on = true;
    } else {
        ctx->tune_on = saved_on;
    }
    (void)hipStreamSynchronize(s);
    for (void* p : {k0, k, ka, v, va})
        if (p) (void)hipFree(p);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipStreamDestroy(s);
    return st;
}

extern "C" int tc_tuned_sort_shape(tc_ctx* ctx, int key_bits, int pairs, uint32_t n, int* block, int* kpt) {
    int bits = 0;
    return tc_tuned_sort_plan(ctx, key_bits, pairs, n, block, kpt, &bits);
}

extern "C" int tc_tuned_sort_plan(tc_ctx* ctx, int key_bits, int pairs, uint32_t n, int* block, int* kpt,
                                  int* digit_bits) {
    if (!ctx || !block || !kpt || !digit_bits || (key_bits != 32 && key_bits != 64)) return TC_ERR_INVALID;
    const SortShape sh = tc_sort_shape(ctx, key_bits / 8, pairs != 0, n);
    *block = sh.block;
    *kpt = sh.kpt;
    *digit_bits = sh.bits;
    return TC_OK;
}

extern "C" int tc_set_tuned_sort_shape(tc_ctx* ctx, int key_bits, int pairs, int log2_n, int block, int kpt) {
    return tc_set_tuned_sort_plan(ctx, key_bits, pairs, log2_n, block, kpt, 8);
}

extern "C" int tc_set_tuned_sort_plan(tc_ctx* ctx, int key_bits, int pairs, int log2_n, int block, int kpt,
                                      int digit_bits) {
    if (!ctx || (key_bits != 32 && key_bits != 64) || log2_n < 0 || log2_n > 31) return TC_ERR_INVALID;
    const int idx = shape_index(block, kpt, digit_bits);
    if (idx < 0) return TC_ERR_INVALID;
    ctx->tune[key_bits == 64 ? 1 : 0][pairs ? 1 : 0][log2_n] = (int8_t)idx;
    ctx->tune_on = true;
    return TC_OK;
}

extern "C" int tc_set_sort_plan(tc_ctx* ctx, int block, int kpt, int digit_bits) {
    if (!ctx) return TC_ERR_INVALID;
    if (block == 0) block = ctx->sort_block;
    if (kpt == 0) kpt = ctx->sort_kpt;
    if (digit_bits == 0) digit_bits = ctx->sort_bits;
    if (shape_index(block, kpt, digit_bits) < 0) return TC_ERR_INVALID;
    ctx->sort_block = block;
    ctx->sort_kpt = kpt;
    ctx->sort_bits = digit_bits;
    ctx->tune_on = false;  // an explicit shape replaces tuned ones
    return TC_OK;
}

extern "C" int tc_sort_shape_list(int* blocks, int* kpts, int max) {
    return tc_sort_plan_list(blocks, kpts, nullptr, max);
}

extern "C" int tc_sort_plan_list(int* blocks, int* kpts, int* digit_bits, int max) {
    for (int i = 0; i < N_SORT_SHAPES && i < max; ++i) {
        if (blocks) blocks[i] = SORT_SHAPES[i].block;
        if (kpts) kpts[i] = SORT_SHAPES[i].kpt;
        if (digit_bits) digit_bits[i] = SORT_SHAPES[i].bits;
    }
    return N_SORT_SHAPES;
}

extern "C" int tc_clear_sort_tuning(tc_ctx* ctx) {
    if (!ctx) return TC_ERR_INVALID;
    std::memset(ctx->tune, -1, sizeof(ctx->tune));
    ctx->tune_on = false;
    return TC_OK;
}
