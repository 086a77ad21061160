// sort.hip -- one-pass global histogram + OneSweep LSD digit-binning passes.
//
// Reference: k_GlobalHistogram, shader/histogram.comp:103-195 and
// k_DigitBinning, shader/new_radix_sort.comp:116-314, driven by
// test_radix_sort, src/main.rs:185-800.  Published algorithm: Adinets &
// Merrill, "Onesweep: A Faster Least Significant Digit Radix Sort for GPUs"
// (2022), with Merrill & Garland's decoupled look-back.
//
// gfx950 design:
//  * histogram: one read of the keys (16 B vector loads) builds all 4 (u32) or
//    8 (u64) digit histograms in LDS sub-histograms, merged with device atomics
//    into 32 accumulator copies (same-address device atomics serialise); a
//    one-workgroup-per-digit kernel sums the copies into exclusive prefixes and
//    clears them for the next call.  The histogram kernel also clears the
//    pass-0 look-back words and the per-pass tile counters.
//    In the pipeline the histogram is fused into the Morton kernel instead.
//  * binning pass: a tile of BLOCK*KPT keys per workgroup (tile = workgroup
//    id).  Each wave ranks its 64*KPT keys with 64-lane ballots (8 ballots give
//    the mask of lanes holding the same digit), so the rank is stable.  Per-wave
//    digit counters live in LDS; a block scan turns them into tile-local
//    offsets; keys are staged in LDS in digit order and written out in runs
//    (coalesced stores).  Look-back words are u32 (2-bit flag, 30-bit count)
//    published and read with agent-scope atomics, since the 8 XCDs' L2 caches
//    are not coherent with each other; one thread per digit walks back, four
//    predecessors per round trip, with a bounded wait and a self-computed
//    fall-back.  Each pass clears the other look-back region for the next pass.
#include <cstring>

#include "sort_core.hpp"

namespace {

constexpr int HIST_BLOCK = 256;
constexpr int HIST_REP = 4;  // LDS sub-histogram copies per block (one per wave)
constexpr int HG_ITEMS = 4;  // 16-byte key loads in flight per thread in the histogram kernel
static_assert(HIST_BLOCK == 256, "one thread per bin in k_hist_finish");

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

// Merge a block's LDS sub-histograms into accumulator copy blockIdx % HIST_COPIES
// (ctx->hist_acc, HIST_COPIES x 8 x 256 words).  Device atomics to one address
// serialise (one per block and bin would make a chain of gridDim.x adds per
// bin at the end of the kernel); spread over the copies the chains are
// gridDim.x / HIST_COPIES long.  k_hist_finish sums the copies.
template <int DIG>
TC_DEV void hist_merge(const uint32_t* s_h, uint32_t* __restrict__ acc) {
    uint32_t* a = acc + (blockIdx.x % HIST_COPIES) * (8 * 256);
    for (uint32_t b = threadIdx.x; b < DIG * 256; b += HIST_BLOCK) {
        uint32_t v = 0;
#pragma unroll
        for (int r = 0; r < HIST_REP; ++r) v += s_h[r * DIG * 256 + b];
        if (v) atomicAdd(&a[b], v);
    }
}

// One workgroup per digit: the digit's counts summed over the accumulator
// copies, written as an exclusive prefix (digit-major, the layout of
// shader/histogram.comp:103-195); the copies are cleared for the next use.
// It also advances the scan epoch tag (as a scan's own one-lane epoch kernel
// would), so a pipeline whose next scan follows this sort can skip that launch.
__global__ __launch_bounds__(HIST_BLOCK) void k_hist_finish(uint32_t* __restrict__ acc,
                                                            uint32_t* __restrict__ hist_out,
                                                            uint32_t* __restrict__ epoch_seq) {
    __shared__ uint32_t s_tmp[HIST_BLOCK / 64 + 2];
    const uint32_t d = blockIdx.x, tid = threadIdx.x;
    if (d == 0 && tid == 0) *epoch_seq = (*epoch_seq + 1u) & 0x3FFFFFFFu;
    uint32_t c = 0;
#pragma unroll 8
    for (int r = 0; r < HIST_COPIES; ++r) {
        uint32_t* p = &acc[r * (8 * 256) + d * 256 + tid];
        c += *p;
        *p = 0u;
    }
    uint32_t tot;
    hist_out[d * 256 + tid] = block_exclusive_scan<HIST_BLOCK>(c, s_tmp, &tot);
}

template <typename K>
__global__ __launch_bounds__(HIST_BLOCK) void k_global_histogram(
    const K* __restrict__ keys, uint32_t n, uint32_t* __restrict__ acc, uint32_t* __restrict__ zero_ptr,
    uint32_t zero_words) {
    constexpr int DIG = (int)sizeof(K);
    constexpr int VN = KeyVec<K>::N;
    __shared__ uint32_t s_h[HIST_REP][DIG * 256];

    const uint32_t tid = threadIdx.x;
    for (uint32_t i = tid; i < HIST_REP * DIG * 256; i += HIST_BLOCK) (&s_h[0][0])[i] = 0;
    __syncthreads();

    uint32_t* h = s_h[(tid / 64) % HIST_REP];
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
#pragma unroll
        for (int u = 0; u < HG_ITEMS; ++u) {
            K k[VN];
            vec_get<K>(v[u], k);
#pragma unroll
            for (int j = 0; j < VN; ++j)
                lds_count_digits<K, DIG>(h, k[j]);
        }
    }
    for (uint32_t v = full * CHUNK + gtid; v < nvec; v += gstride) {
        K k[VN];
        vec_get<K>(kv[v], k);
#pragma unroll
        for (int j = 0; j < VN; ++j)
            lds_count_digits<K, DIG>(h, k[j]);
    }
    for (uint32_t i = nvec * VN + gtid; i < n; i += gstride) {
        const K k = keys[i];
        lds_count_digits<K, DIG>(h, k);
    }
    // Clear the pass-0 look-back words and the tile counters for the passes.
    for (uint32_t w = gtid; w < zero_words; w += gstride) zero_ptr[w] = 0u;
    __syncthreads();
    hist_merge<DIG>(&s_h[0][0], acc);
}

template <typename K>
TC_DEV K encode_point(const float4& p, float min_coord, float range) {
    if constexpr (sizeof(K) == 4) return d_encode30(p.x, p.y, p.z, min_coord, range);
    else return d_encode63(p.x, p.y, p.z, min_coord, range);
}

// Morton codes with the sort's global histogram fused in: the codes are counted
// while they are in registers, so the sort does not read them an extra time
// (k_ComputeMortonCode, shader/morton.comp:60-67 + k_GlobalHistogram,
// shader/histogram.comp:103-195).  Each thread keeps MH_ITEMS point loads in
// flight to cover HBM latency.
constexpr int MH_ITEMS = 8;

template <typename K>
__global__ __launch_bounds__(HIST_BLOCK) void k_morton_hist(
    const float4* __restrict__ pts, K* __restrict__ codes, uint32_t* __restrict__ index, uint32_t n,
    float min_coord, float range, uint32_t* __restrict__ acc, uint32_t* __restrict__ zero_ptr,
    uint32_t zero_words) {
    constexpr int DIG = (int)sizeof(K);
    constexpr uint32_t CHUNK = HIST_BLOCK * MH_ITEMS;
    __shared__ uint32_t s_h[HIST_REP][DIG * 256];
    const uint32_t tid = threadIdx.x;
    for (uint32_t i = tid; i < HIST_REP * DIG * 256; i += HIST_BLOCK) (&s_h[0][0])[i] = 0;
    __syncthreads();
    uint32_t* h = s_h[(tid / 64) % HIST_REP];
    const uint32_t full = n / CHUNK;
    for (uint32_t c = blockIdx.x; c < full; c += gridDim.x) {
        const uint32_t b0 = c * CHUNK + tid;
        float4 p[MH_ITEMS];
#pragma unroll
        for (int u = 0; u < MH_ITEMS; ++u) p[u] = ld_stream(&pts[b0 + u * HIST_BLOCK]);
#pragma unroll
        for (int u = 0; u < MH_ITEMS; ++u) {
            const K code = encode_point<K>(p[u], min_coord, range);
            codes[b0 + u * HIST_BLOCK] = code;
            if (index) index[b0 + u * HIST_BLOCK] = b0 + u * HIST_BLOCK;  // sort payload: the point's index
            lds_count_digits<K, DIG>(h, code);
        }
    }
    const uint32_t gtid = blockIdx.x * HIST_BLOCK + tid;
    const uint32_t gstride = gridDim.x * HIST_BLOCK;
    for (uint32_t i = full * CHUNK + gtid; i < n; i += gstride) {
        const K code = encode_point<K>(pts[i], min_coord, range);
        codes[i] = code;
        if (index) index[i] = i;
        lds_count_digits<K, DIG>(h, code);
    }
    for (uint32_t w = gtid; w < zero_words; w += gstride) zero_ptr[w] = 0u;
    __syncthreads();
    hist_merge<DIG>(&s_h[0][0], acc);
}

template <typename K, bool PAIRS, int BLOCK, int KPT, int RB>
__global__ __launch_bounds__(BLOCK) void k_onesweep(
    const K* __restrict__ sk, K* __restrict__ dk, const uint32_t* __restrict__ sv,
    uint32_t* __restrict__ dv, uint32_t n, uint32_t shift, const uint32_t* __restrict__ digit_prefix,
    uint32_t* __restrict__ lb_cur, uint32_t* __restrict__ lb_next, uint32_t spin_limit) {
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
    onesweep_tile<K, PAIRS, BLOCK, KPT, RB, !PAIRS>(sm, sk, dk, sv, dv, base, cnt, shift, tile, 0u, 0u, digit_prefix, 0u,
                                            lb_cur, lb_next, spin_limit);
}

// narrow: the pass ranks only the top digit's live Morton-code bits.
template <typename K, bool PAIRS, int BLOCK, int KPT>
void launch_pass(uint32_t tiles, hipStream_t s, const K* sk, K* dk, const uint32_t* sv, uint32_t* dv,
                 uint32_t n, uint32_t shift, const uint32_t* prefix, uint32_t* lb_cur, uint32_t* lb_next,
                 uint32_t spin_limit, bool narrow) {
    if (narrow)
        hipLaunchKernelGGL((k_onesweep<K, PAIRS, BLOCK, KPT, morton_top_digit_bits<K>()>), dim3(tiles), dim3(BLOCK),
                           0, s, sk, dk, sv, dv, n, shift, prefix, lb_cur, lb_next, spin_limit);
    else
        hipLaunchKernelGGL((k_onesweep<K, PAIRS, BLOCK, KPT, 8>), dim3(tiles), dim3(BLOCK), 0, s, sk, dk, sv, dv,
                           n, shift, prefix, lb_cur, lb_next, spin_limit);
}

template <typename K, bool PAIRS>
int dispatch_pass(tc_ctx* ctx, SortShape sh, uint32_t tiles, hipStream_t s, const K* sk, K* dk,
                  const uint32_t* sv, uint32_t* dv, uint32_t n, uint32_t shift, const uint32_t* prefix,
                  uint32_t* lb_cur, uint32_t* lb_next, bool narrow = false) {
    ProfScope prof(ctx, s, K_BINNING);
#define TC_PASS_CASE(B, KP)                                                                         \
    if (sh.block == B && sh.kpt == KP) {                                                            \
        launch_pass<K, PAIRS, B, KP>(tiles, s, sk, dk, sv, dv, n, shift, prefix, lb_cur, lb_next,   \
                                     ctx->spin_limit, narrow);                                      \
        TC_LAUNCH_CHECK();                                                                           \
        return TC_OK;                                                                                \
    }
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
#undef TC_PASS_CASE
    return TC_ERR_INVALID;
}

// Four times as many workgroups as are resident at once (eight per CU for
// u32 keys, five for u64 keys, whose 32 KiB of LDS sub-histograms fit five per
// CU): each streams a few chunks and the hardware hands the next workgroup to
// whichever CU frees up first.  Measured on the fused Morton + histogram
// kernel at 2^26 points: 0.248 ms with 8 workgroups per CU, 0.240 with 16,
// 0.233 with 32.  The merge's atomic chains stay gridDim / HIST_COPIES long.
template <typename K>
uint32_t hist_grid(tc_ctx* ctx, uint32_t n) {
    uint32_t blocks = ceil_div_u32(n, HIST_BLOCK * 16);
    uint32_t cap = (uint32_t)ctx->num_cus * (sizeof(K) == 4 ? 32u : 20u);
    if (blocks < 1) blocks = 1;
    return blocks < cap ? blocks : cap;
}

// Scratch of one sort: [64 tile counters][region 0][region 1][hist], plus the
// histogram accumulators (zero between uses) and the scan epoch word that the
// histogram-finish kernel advances.
struct SortScratch {
    uint32_t* ctr;
    uint32_t* region[2];
    uint32_t* hist;
    uint32_t* acc;
    uint32_t* epoch;
};

int sort_scratch(tc_ctx* ctx, uint32_t n, int digits, SortShape sh, SortScratch* ss, uint32_t* tiles_out) {
    const uint32_t tile = (uint32_t)(sh.block * sh.kpt);
    const uint32_t tiles = ceil_div_u32(n, tile);
    const size_t region_words = (size_t)tiles * 256;
    const size_t words = 64 + 2 * region_words + (size_t)digits * 256;
    TC_TRY(tc_ensure_scratch(ctx, words * 4));
    uint32_t* base = static_cast<uint32_t*>(ctx->scratch);
    ss->ctr = base;
    ss->region[0] = base + 64;
    ss->region[1] = base + 64 + region_words;
    ss->hist = base + 64 + 2 * region_words;
    ss->acc = ctx->hist_acc;
    ss->epoch = tc_scan_epoch(ctx);
    *tiles_out = tiles;
    return TC_OK;
}

template <typename K>
int histogram_impl(tc_ctx* ctx, const K* keys, uint32_t n, uint32_t* hist_out, uint32_t* acc, uint32_t* epoch,
                   uint32_t* zero_ptr, uint32_t zero_words, hipStream_t s) {
    ProfScope prof(ctx, s, K_HIST);
    hipLaunchKernelGGL((k_global_histogram<K>), dim3(hist_grid<K>(ctx, n)), dim3(HIST_BLOCK), 0, s, keys, n,
                       acc, zero_ptr, zero_words);
    TC_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_hist_finish, dim3(sizeof(K)), dim3(HIST_BLOCK), 0, s, acc, hist_out, epoch);
    TC_LAUNCH_CHECK();
    return TC_OK;
}

// The launches of one sort (histogram, then one binning pass per digit) on
// the scratch `ss` (tiles look-back rows per region).  With pts != NULL the
// keys are first computed from the points by the fused Morton + histogram
// kernel, which also sets vals[i] = i when vals is given (the sorted values are
// then the permutation from sorted position to point index).
template <typename K>
int sort_run(tc_ctx* ctx, SortShape sh, const SortScratch& ss, uint32_t tiles, K* keys, uint32_t* vals, K* alt,
             uint32_t* valt, uint32_t n, hipStream_t s, const float* pts, float min_coord, float range) {
    constexpr int DIG = (int)sizeof(K);
    const uint32_t region_words = tiles * 256;
    // The histogram kernel also clears the tile counters and look-back region 0,
    // as one range: the region must follow the 64 counters.
    if (ss.region[0] != ss.ctr + 64) return TC_ERR_INVALID;
    // Every n > 0 sort launches k_hist_finish (both branches below), which
    // advances the scan epoch: tc_build_tree's unique scan relies on that to
    // skip its own epoch kernel.
    if (pts) {
        ProfScope prof(ctx, s, K_MORTON);
        hipLaunchKernelGGL((k_morton_hist<K>), dim3(hist_grid<K>(ctx, n)), dim3(HIST_BLOCK), 0, s,
                           reinterpret_cast<const float4*>(pts), keys, vals, n, min_coord, range, ss.acc,
                           ss.ctr, 64 + region_words);
        TC_LAUNCH_CHECK();
        hipLaunchKernelGGL(k_hist_finish, dim3(DIG), dim3(HIST_BLOCK), 0, s, ss.acc, ss.hist, ss.epoch);
        TC_LAUNCH_CHECK();
    } else {
        TC_TRY(histogram_impl<K>(ctx, keys, n, ss.hist, ss.acc, ss.epoch, ss.ctr, 64 + region_words, s));
    }
    K* src = keys;
    K* dst = alt;
    uint32_t* vsrc = vals;
    uint32_t* vdst = valt;
    for (int p = 0; p < DIG; ++p) {
        uint32_t* cur = ss.region[p & 1];
        uint32_t* nxt = (p + 1 < DIG) ? ss.region[(p + 1) & 1] : nullptr;
        // Morton codes (computed here from pts) have a narrow top digit.
        const bool narrow = pts != nullptr && p == DIG - 1;
        if (vals) {
            TC_TRY((dispatch_pass<K, true>(ctx, sh, tiles, s, src, dst, vsrc, vdst, n, 8u * p, ss.hist + 256 * p,
                                           cur, nxt, narrow)));
        } else {
            TC_TRY((dispatch_pass<K, false>(ctx, sh, tiles, s, src, dst, nullptr, nullptr, n, 8u * p,
                                            ss.hist + 256 * p, cur, nxt, narrow)));
        }
        K* t = src; src = dst; dst = t;
        uint32_t* tv = vsrc; vsrc = vdst; vdst = tv;
    }
    // An even number of passes leaves the result in keys/vals.
    return TC_OK;
}

// Sort keys (and values) in place on the context's scratch.
template <typename K>
int sort_impl(tc_ctx* ctx, K* keys, uint32_t* vals, K* alt, uint32_t* valt, uint32_t n, hipStream_t s,
              const float* pts = nullptr, float min_coord = 0.f, float range = 1.f) {
    if (!ctx || (n && (!keys || !alt))) return TC_ERR_INVALID;
    if (vals && !valt) return TC_ERR_INVALID;
    if (n > TC_MAX_SORT_N) return TC_ERR_TOO_LARGE;
    if (n == 0) return TC_OK;
    if (reinterpret_cast<uintptr_t>(keys) & 15u) return TC_ERR_INVALID;
    const SortShape sh = tc_sort_shape(ctx, (int)sizeof(K), vals != nullptr, n);
    SortScratch ss;
    uint32_t tiles;
    TC_TRY(sort_scratch(ctx, n, (int)sizeof(K), sh, &ss, &tiles));
    return sort_run<K>(ctx, sh, ss, tiles, keys, vals, alt, valt, n, s, pts, min_coord, range);
}

template <typename K>
int binning_pass_impl(tc_ctx* ctx, const K* sk, K* dk, const uint32_t* sv, uint32_t* dv, uint32_t n,
                      const uint32_t* hist, uint32_t shift, hipStream_t s) {
    if (!ctx || (n && (!sk || !dk || !hist))) return TC_ERR_INVALID;
    if ((sv == nullptr) != (dv == nullptr)) return TC_ERR_INVALID;
    if (shift % 8 != 0 || shift >= 8 * sizeof(K)) return TC_ERR_INVALID;
    if (n > TC_MAX_SORT_N) return TC_ERR_TOO_LARGE;
    if (n == 0) return TC_OK;
    const SortShape sh = tc_sort_shape(ctx, (int)sizeof(K), sv != nullptr, n);
    SortScratch ss;
    uint32_t tiles;
    TC_TRY(sort_scratch(ctx, n, (int)sizeof(K), sh, &ss, &tiles));
    TC_TRY(tc_fill_u32(ctx, ss.ctr, 0u, 64 + (size_t)tiles * 256, s));
    const uint32_t* prefix = hist + 256 * (shift / 8);
    if (sv) return dispatch_pass<K, true>(ctx, sh, tiles, s, sk, dk, sv, dv, n, shift, prefix, ss.region[0],
                                          nullptr);
    return dispatch_pass<K, false>(ctx, sh, tiles, s, sk, dk, nullptr, nullptr, n, shift, prefix, ss.region[0],
                                   nullptr);
}

template <typename K>
int global_histogram_api(tc_ctx* ctx, const K* keys, uint32_t n, uint32_t* hist, hipStream_t s) {
    if (!ctx || !hist || (n && !keys)) return TC_ERR_INVALID;
    if (reinterpret_cast<uintptr_t>(keys) & 15u) return TC_ERR_INVALID;
    return histogram_impl<K>(ctx, keys, n, hist, ctx->hist_acc, tc_scan_epoch(ctx), nullptr, 0, s);
}

}  // namespace

size_t tc_sort_scratch_bytes(uint32_t n, int passes, int tile) {
    const size_t tiles = ceil_div_u32(n, (uint32_t)tile);
    return (64 + 2 * tiles * 256 + (size_t)passes * 256) * 4;
}

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
// tile counters, the histogram, the scan epoch word, the histogram
// accumulators and two look-back regions sized for 2^24 keys in tiles of at
// least 8192 keys.  Host code reaches it only through the address
// hipGetSymbolAddress returns on each device (tc_ctx::noalloc_scratch, set by
// tc_create with the context's device current); the array's name itself is a
// host-side stand-in that must never be passed to a kernel.
namespace {

constexpr uint32_t NA_MAX_N = 1u << 24;
constexpr uint32_t NA_MIN_TILE = 8192;
constexpr size_t NA_REGION_WORDS = (size_t)(NA_MAX_N / NA_MIN_TILE) * 256;
// The sort's layout: the tile counters directly followed by look-back region 0
// (the histogram kernel clears both as one range).
constexpr size_t NA_CTR = 0;
constexpr size_t NA_R0 = 64;
constexpr size_t NA_R1 = NA_R0 + NA_REGION_WORDS;
constexpr size_t NA_HIST = NA_R1 + NA_REGION_WORDS;
constexpr size_t NA_EPOCH = NA_HIST + 8 * 256;
constexpr size_t NA_ACC = NA_EPOCH + 64;
constexpr size_t NA_WORDS = NA_ACC + (size_t)HIST_COPIES * 8 * 256;  // 1,116,288 words = 4.26 MiB

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
    if (n == 0) return TC_OK;
    if (reinterpret_cast<uintptr_t>(keys) & 15u) return TC_ERR_INVALID;
    uint32_t* base = ctx->noalloc_scratch;
    if (!base) return TC_ERR_HIP;
    hipStream_t s = as_stream(stream);
    // The tuned shape unless its tiles are too small for the static regions;
    // then the first built shape whose tiles are (every n <= 2^24 fits those).
    SortShape sh = tc_sort_shape(ctx, 4, false, n);
    if ((size_t)ceil_div_u32(n, (uint32_t)(sh.block * sh.kpt)) * 256 > NA_REGION_WORDS) {
        for (const SortShape& c : SORT_SHAPES) {
            if ((uint32_t)(c.block * c.kpt) >= NA_MIN_TILE) {
                sh = c;
                break;
            }
        }
    }
    const uint32_t tiles = ceil_div_u32(n, (uint32_t)(sh.block * sh.kpt));
    SortScratch ss{base + NA_CTR, {base + NA_R0, base + NA_R1}, base + NA_HIST, base + NA_ACC, base + NA_EPOCH};
    // The accumulators must start at zero: every sort leaves them so, but a
    // call that failed part-way may not have.
    TC_TRY(tc_fill_u32(ctx, ss.acc, 0u, (size_t)HIST_COPIES * 8 * 256, s));
    return sort_run<uint32_t>(ctx, sh, ss, tiles, keys, nullptr, alt, nullptr, n, s, nullptr, 0.f, 1.f);
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
    return SortShape{ctx->sort_block, ctx->sort_kpt};
}

// Time every launch shape on random keys at sizes 2^12, 2^14, ... up to max_n,
// for u32 and u64 keys, with and without a u32 payload, and keep the fastest
// per size.  Blocking; uses its own stream and buffers.
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
    const int saved_block = ctx->sort_block, saved_kpt = ctx->sort_kpt;
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
    if (!ctx || !block || !kpt || (key_bits != 32 && key_bits != 64)) return TC_ERR_INVALID;
    const SortShape sh = tc_sort_shape(ctx, key_bits / 8, pairs != 0, n);
    *block = sh.block;
    *kpt = sh.kpt;
    return TC_OK;
}

extern "C" int tc_set_tuned_sort_shape(tc_ctx* ctx, int key_bits, int pairs, int log2_n, int block, int kpt) {
    if (!ctx || (key_bits != 32 && key_bits != 64) || log2_n < 0 || log2_n > 31) return TC_ERR_INVALID;
    int idx = -1;
    for (int i = 0; i < N_SORT_SHAPES; ++i)
        if (SORT_SHAPES[i].block == block && SORT_SHAPES[i].kpt == kpt) idx = i;
    if (idx < 0) return TC_ERR_INVALID;
    ctx->tune[key_bits == 64 ? 1 : 0][pairs ? 1 : 0][log2_n] = (int8_t)idx;
    ctx->tune_on = true;
    return TC_OK;
}

extern "C" int tc_sort_shape_list(int* blocks, int* kpts, int max) {
    for (int i = 0; i < N_SORT_SHAPES && i < max; ++i) {
        if (blocks) blocks[i] = SORT_SHAPES[i].block;
        if (kpts) kpts[i] = SORT_SHAPES[i].kpt;
    }
    return N_SORT_SHAPES;
}

extern "C" int tc_clear_sort_tuning(tc_ctx* ctx) {
    if (!ctx) return TC_ERR_INVALID;
    std::memset(ctx->tune, -1, sizeof(ctx->tune));
    ctx->tune_on = false;
    return TC_OK;
}
