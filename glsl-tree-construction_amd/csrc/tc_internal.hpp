// tc_internal.hpp -- shared definitions of libtc: context, scratch layout,
// error plumbing and wave64 device helpers for gfx950 (CDNA4).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <mutex>
#include <type_traits>
#include <vector>

#include "../../include/tc.h"

// True if the byte ranges [a, a + an) and [b, b + bn) share an address.  The
// sort and scan kernels read their input again when a look-back computes a
// stalled predecessor tile (and the binning passes scatter into the other
// buffer), so entry points refuse outputs that overlap an input.
inline bool tc_overlaps(const void* a, size_t an, const void* b, size_t bn) {
    const uintptr_t x = reinterpret_cast<uintptr_t>(a), y = reinterpret_cast<uintptr_t>(b);
    return an && bn && x < y + bn && y < x + an;
}

#define TC_HIP(call)                                   \
    do {                                               \
        hipError_t e_ = (call);                        \
        if (e_ != hipSuccess) return TC_ERR_HIP;       \
    } while (0)

#define TC_TRY(call)                                   \
    do {                                               \
        int s_ = (call);                               \
        if (s_ != TC_OK) return s_;                    \
    } while (0)

#define TC_LAUNCH_CHECK() TC_HIP(hipGetLastError())

// Lookback flag words of the sort: bits 0-1 flag, bits 2-31 count.
constexpr uint32_t LB_NOT_READY = 0u;
constexpr uint32_t LB_AGGREGATE = 1u;
constexpr uint32_t LB_INCLUSIVE = 2u;
constexpr uint32_t LB_FLAG_MASK = 3u;

// Scan tile status words: bits 0-31 value, bits 32-33 flag.
constexpr uint64_t ST_AGGREGATE = 1ull << 32;
constexpr uint64_t ST_INCLUSIVE = 2ull << 32;
constexpr uint64_t ST_FLAG_MASK = 3ull << 32;

// Kernel ids for the per-launch event profiler (tc_profile_*).
enum TcKernelId {
    K_MORTON = 0,
    K_HIST,
    K_BINNING,
    K_UNIQUE,
    K_SCAN,
    K_BRT,
    K_EDGE,
    K_ECSCAN,
    K_OCT_EMIT,
    K_OCT_LINK,
    K_FILL,
    K_DECODE,
    K_SEG,  // segmented (batched) kernels
    K_COUNT
};

// roctx marker entry points, resolved from the ROCm marker library when a
// context first enables markers (tc_set_markers); null until then.
struct TcRoctx {
    int (*push)(const char*) = nullptr;
    int (*pop)() = nullptr;
};
extern TcRoctx g_tc_roctx;
const char* tc_kernel_name_internal(int id);

struct ProfRec {
    int id;
    hipEvent_t a, b;
};

struct tc_ctx {
    int device = 0;
    // Event profiler: when on, every launch is bracketed by two events on its
    // stream; tc_profile_collect folds the elapsed times into per-kernel sums.
    bool prof_on = false;
    bool markers_on = false;  // roctx ranges around calls and launches
    std::vector<ProfRec> prof_pending;
    std::vector<hipEvent_t> prof_pool;
    double prof_ms[K_COUNT] = {};
    uint64_t prof_n[K_COUNT] = {};
    int num_cus = 256;
    // Sort launch parameters (used when no tuned shape applies): workgroup
    // size, keys per lane and digit plan (8: 8-bit digits; 11: digits of at
    // most 11 bits, fewer passes).
    int sort_block = 512;
    int sort_kpt = 20;
    int sort_bits = 8;
    // Tuned launch shapes (tc_autotune_sort): index into SORT_SHAPES per
    // [key bytes 4 / 8][keys only / pairs][ceil(log2 n)], -1 where not measured.
    bool tune_on = false;
    int8_t tune[2][2][32];
    // Look-back wait bound: polls without progress (about 0.8 us each) before
    // a workgroup computes the stalled predecessor's counts itself, one tile
    // read.  Short, since a predecessor that has not started may wait for
    // another stream's kernel to leave its XCD; tests set 0 (every
    // predecessor computed) and 1 (most of them) to exercise that path.
    uint32_t spin_limit = TC_LOOKBACK_SPIN_DEFAULT;
    // Grow-only device scratch.
    void* scratch = nullptr;
    size_t scratch_bytes = 0;
    // Permanent small device block: histogram accumulators (self-cleaning),
    // completion counters, pipeline counters.
    uint32_t* small = nullptr;
    // Sort histogram accumulators: HIST_COPIES x 8 x 256 counters, zero
    // between uses (the finishing kernel clears what it reads).
    uint32_t* hist_acc = nullptr;
    // Scan tile status words (epoch-tagged, never cleared): one per 8192-element
    // tile up to 2^30 elements.
    uint64_t* scan_status = nullptr;
    // This device's address of the static scratch array of the sort that
    // allocates nothing (tc_sort_keys_u32_noalloc); NULL if it was not found.
    uint32_t* noalloc_scratch = nullptr;
    // Host-side staging for the host-array wrappers.
    std::mutex mu;
};

// Copies of the sort histogram accumulator (ctx->hist_acc).
constexpr int HIST_COPIES = 32;

// Layout of ctx->small (u32 words; words 0-2055 unused).
constexpr size_t SMALL_TILE_CTR = 2056;     // 8 per-pass tile counters (+ 8 spare)
constexpr size_t SMALL_SCAN_CTR = 2072;     // scan tile counters (8)
constexpr size_t SMALL_WORDS = 4096;

// Scratch sizing helpers (bytes).
int tc_ensure_scratch(tc_ctx* ctx, size_t bytes);
// Sort scratch for n keys with any built shape and plan.
size_t tc_sort_scratch_max_bytes(uint32_t n);
// The sort's histogram accumulators (ctx->hist_acc).
size_t tc_hist_acc_bytes();

// Binning-pass launch shapes (workgroup size, keys per lane, digit plan: 8 =
// 8-bit digits, 11 = digits of at most 11 bits) the library is built with; the
// tuner picks among them.
struct SortShape {
    int block, kpt, bits;
};
constexpr int N_SORT_SHAPES = 12;
constexpr SortShape SORT_SHAPES[N_SORT_SHAPES] = {
    {256, 8, 8},  {512, 8, 8},  {512, 12, 8}, {512, 16, 8}, {512, 20, 8},   {512, 24, 8},
    {512, 28, 8}, {512, 32, 8}, {512, 36, 8}, {1024, 16, 8}, {1024, 8, 11}, {1024, 16, 11}};
// The shape a sort of n keys uses: the tuned one for the nearest measured size
// when the context has been tuned, else the context's fixed shape.
SortShape tc_sort_shape(const tc_ctx* ctx, int key_bytes, bool pairs, uint32_t n);
// The same restricted to 8-bit plans (the segmented batch sort's passes).
SortShape tc_sort_shape_8bit(const tc_ctx* ctx, int key_bytes, bool pairs, uint32_t n);
// Scan status words: one 64-byte line per tile, for up to TC_SCAN_MAX_TILES tiles.
constexpr uint32_t TC_SCAN_MAX_TILES = 1u << 17;
size_t tc_scan_status_bytes();
// Looks up this device's address of the no-alloc sort's static scratch
// (current device = ctx->device); TC_ERR_HIP if the lookup fails.
int tc_noalloc_init(tc_ctx* ctx);

static inline hipStream_t as_stream(tc_stream s) { return reinterpret_cast<hipStream_t>(s); }

// Brackets one kernel launch with profiler events (no-op when profiling is off).
hipEvent_t tc_prof_event(tc_ctx* ctx);
struct ProfScope {
    tc_ctx* ctx;
    hipStream_t s;
    int id;
    hipEvent_t a = nullptr;
    bool marked = false;
    ProfScope(tc_ctx* c, hipStream_t st, int kid) : ctx(c), s(st), id(kid) {
        if (ctx->markers_on && g_tc_roctx.push) marked = g_tc_roctx.push(tc_kernel_name_internal(id)) >= 0;
        if (ctx->prof_on) {
            a = tc_prof_event(ctx);
            if (a) (void)hipEventRecord(a, s);
        }
    }
    ~ProfScope() {
        if (marked) g_tc_roctx.pop();
        if (a) {
            hipEvent_t b = tc_prof_event(ctx);
            if (b) {
                (void)hipEventRecord(b, s);
                ctx->prof_pending.push_back({id, a, b});
            }
        }
    }
};

// A roctx range around a whole call (host-side: the time the call spends
// enqueuing its launches); no-op unless the context has markers on.
struct MarkRange {
    bool on;
    MarkRange(const tc_ctx* ctx, const char* name)
        : on(ctx && ctx->markers_on && g_tc_roctx.push && g_tc_roctx.push(name) >= 0) {}
    ~MarkRange() {
        if (on) g_tc_roctx.pop();
    }
};

__host__ __device__ static inline uint32_t ceil_div_u32(uint64_t a, uint64_t b) {
    return (uint32_t)((a + b - 1) / b);
}

// Modes of the single-pass tile scan (scan.hip).
enum ScanMode {
    MODE_SCAN = 0,
    MODE_UNIQUE32 = 1,
    MODE_UNIQUE64 = 2,
    MODE_ECSCAN = 3,
    MODE_SEG_UNIQUE32 = 4,
    MODE_SEG_UNIQUE64 = 5,
    MODE_SEG_SCAN = 6
};

// Entry points shared between translation units.
// tc_launch_scan: count_out (may be NULL) receives the scan's total, except in
// MODE_SCAN, where count_out[0] = element count and count_out[1] = total.
int tc_fill_u32(tc_ctx* ctx, uint32_t* ptr, uint32_t value, size_t count, hipStream_t s);
int tc_launch_scan(tc_ctx* ctx, int mode, const void* in0, const void* in1, void* out0, void* out1,
                   uint32_t* count_out, uint32_t n, const uint32_t* d_n_keys, const uint32_t* seg_off,
                   uint32_t n_seg_cap, const uint32_t* d_n_seg, uint32_t* seg_out, hipStream_t s,
                   bool epoch_advanced = false);
// The scans' epoch tag (device word).  A scan launched with epoch_advanced
// skips its one-lane epoch kernel: the kernel just before it on the stream has
// advanced the tag (the sort's k_hist_finish, the radix-tree kernel).
uint32_t* tc_scan_epoch(tc_ctx* ctx);
// Grid of a kernel that loops over its work (grid-stride): one workgroup per
// `block` items, at most per_cu per CU.  per_cu is the kernel's resident count
// or a multiple of it: a grid slightly larger than what fits leaves the extra
// workgroups to run after the others have finished all their work (a tail as
// long as the kernel), while four times the resident count makes each
// workgroup's share small enough for the hardware to balance the load.
int tc_persistent_grid(tc_ctx* ctx, uint32_t work_items, int block, int per_cu = 8);

// ---------------------------------------------------------------- device side
#define TC_DEV __device__ __forceinline__

TC_DEV uint32_t lane_id() { return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u)); }

// Loads of data that is read exactly once (the points, a plain key sort's
// source keys, unique's sorted input): non-temporal (measured at 2^26:
// Morton + histogram 0.289 -> 0.257 ms, binning pass 0.154 -> 0.147 ms; the
// HBM byte counts of the pass did not change).
template <typename T>
TC_DEV T ld_stream(const T* p) {
    return __builtin_nontemporal_load(p);
}
TC_DEV float4 ld_stream(const float4* p) {
    typedef float v4 __attribute__((ext_vector_type(4)));
    const v4 v = __builtin_nontemporal_load(reinterpret_cast<const v4*>(p));
    return make_float4(v.x, v.y, v.z, v.w);
}
TC_DEV uint4 ld_stream(const uint4* p) {
    typedef unsigned int v4 __attribute__((ext_vector_type(4)));
    const v4 v = __builtin_nontemporal_load(reinterpret_cast<const v4*>(p));
    return make_uint4(v.x, v.y, v.z, v.w);
}
TC_DEV ulonglong2 ld_stream(const ulonglong2* p) {
    typedef unsigned long long v2 __attribute__((ext_vector_type(2)));
    const v2 v = __builtin_nontemporal_load(reinterpret_cast<const v2*>(p));
    return make_ulonglong2(v.x, v.y);
}

// Segment lookups over a nondecreasing offset array off[0..n] (off[0] = 0):
// the segment of index x is the largest c in [0, n-1] with off[c] <= x, which
// skips empty segments.
//
// Binary search for it in [lo, n-1], given off[lo] <= x.
TC_DEV uint32_t seg_search(const uint32_t* off, uint32_t lo, uint32_t n, uint32_t x) {
    uint32_t hi = n;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (off[mid] <= x) lo = mid;
        else hi = mid;
    }
    return lo;
}

// The segment of x, found by a whole wave (all 64 lanes active, same x): each
// round tests 64 evenly spaced candidates at once, so 1,024 segments take 2
// dependent loads instead of 10.  off[c] <= x holds for a prefix of c (off is nondecreasing),
// so the ballot is a run of low lanes; the last of them brackets the answer.
TC_DEV uint32_t seg_find_wave(const uint32_t* off, uint32_t n, uint32_t x) {
    uint32_t lo = 0, hi = n;  // answer in [lo, hi), off[lo] <= x
    const uint32_t lane = __lane_id();
    while (hi - lo > 1) {
        const uint32_t step = (hi - lo + 63) / 64;
        const uint32_t k = lo + lane * step;
        const bool le = k < hi && off[k] <= x;
        const uint64_t m = __ballot(le);
        const uint32_t last = 63u - (uint32_t)__builtin_clzll(m);  // lane 0 always holds: off[lo] <= x
        lo = lo + last * step;
        hi = lo + step < hi ? lo + step : hi;
    }
    return lo;
}

// The segment of x, starting from c0, the segment of an index <= x (a tile's
// first slot): a few steps forward cover the usual case of a tile inside one or
// two clouds, a binary search the rest.
TC_DEV uint32_t seg_find_near(const uint32_t* off, uint32_t n, uint32_t x, uint32_t c0) {
    uint32_t c = c0;
    for (int k = 0; k < 4; ++k) {
        if (c + 1 < n && off[c + 1] <= x) ++c;
        else return c;
    }
    return seg_search(off, c, n, x);
}

// tc_octree_node_bound on the device: 1 + sum over levels of min(8^L, n/2).
TC_DEV uint32_t d_octree_node_bound(uint32_t n_keys, int code_bits) {
    if (n_keys < 2) return 0;
    const uint64_t half = n_keys / 2;
    uint64_t total = 1, p = 1;
    for (int L = 1; L <= (code_bits - 1) / 3; ++L) {
        p = p < (1ull << 60) ? p * 8 : p;
        total += p < half ? p : half;
    }
    return total < 0xFFFFFFFFull ? (uint32_t)total : 0xFFFFFFFFu;
}

// popcount(mask & lanes-below-me), the wave64 exclusive bit count.
TC_DEV uint32_t mbcnt64(uint64_t mask) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

// Adds the DIG 8-bit digits of `key` to the LDS histograms h[d * 256 + digit].
// When every active lane of the wave holds the same digit (the high digits of
// sorted or spatially coherent input), one lane adds the wave's count: 64
// same-address LDS atomics would otherwise serialise.  The branch on that test
// is wave-uniform; the leader and the count are computed once per key.
template <typename K, int DIG>
TC_DEV void lds_count_digits(uint32_t* h, K key) {
    const uint64_t active = __ballot(true);
    const bool leader = mbcnt64(active) == 0u;
    const uint32_t n_active = (uint32_t)__popcll(active);
#pragma unroll
    for (int d = 0; d < DIG; ++d) {
        const uint32_t v = (uint32_t)((key >> (8 * d)) & 255u);
        const uint32_t v0 = __builtin_amdgcn_readfirstlane(v);
        if (__ballot(v == v0) == active) {
            if (leader) atomicAdd(&h[d * 256 + v0], n_active);
        } else {
            atomicAdd(&h[d * 256 + v], 1u);
        }
    }
}

// Relaxed agent-scope loads and stores of look-back and status words: the
// flag and the count share one word, so a reader that sees the flag also sees
// its count and no acquire ordering is needed.
TC_DEV uint32_t ld_agent(const uint32_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
TC_DEV void st_agent(uint32_t* p, uint32_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
TC_DEV uint64_t ld_agent64(const uint64_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
TC_DEV void st_agent64(uint64_t* p, uint64_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Inclusive scan across the 64 lanes of a wave.
TC_DEV uint32_t wave_inclusive_scan(uint32_t v) {
    const uint32_t lane = lane_id();
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        uint32_t t = __shfl_up(v, off, 64);
        if (lane >= (uint32_t)off) v += t;
    }
    return v;
}

// Block-wide exclusive scan of one value per thread; every thread of the block
// must call it.  s_tmp holds >= BLOCK/64 words.  Returns the exclusive prefix;
// *total receives the block sum.
template <int BLOCK>
TC_DEV uint32_t block_exclusive_scan(uint32_t v, uint32_t* s_tmp, uint32_t* total) {
    constexpr int WAVES = BLOCK / 64;
    const uint32_t lane = lane_id();
    const uint32_t wave = threadIdx.x / 64;
    uint32_t inc = wave_inclusive_scan(v);
    if (lane == 63) s_tmp[wave] = inc;
    __syncthreads();
    uint32_t wave_off = 0, sum = 0;
#pragma unroll
    for (int w = 0; w < WAVES; ++w) {
        uint32_t t = s_tmp[w];
        wave_off += ((uint32_t)w < wave) ? t : 0u;
        sum += t;
    }
    __syncthreads();
    *total = sum;
    return wave_off + inc - v;
}

// ---------------------------------------------------------- Morton helpers
// morton3D_SplitBy3bits, shader/morton.comp:36-44.
TC_DEV uint32_t d_split3_10(uint32_t b) {
    uint32_t x = b & 0x000003ffu;
    x = (x | x << 16) & 0x030000ffu;
    x = (x | x << 8) & 0x0300f00fu;
    x = (x | x << 4) & 0x030c30c3u;
    x = (x | x << 2) & 0x09249249u;
    return x;
}
TC_DEV uint64_t d_split3_21(uint64_t b) {
    uint64_t x = b & 0x1fffffull;
    x = (x | x << 32) & 0x001f00000000ffffull;
    x = (x | x << 16) & 0x001f0000ff0000ffull;
    x = (x | x << 8) & 0x100f00f00f00f00full;
    x = (x | x << 4) & 0x10c30c30c30c30c3ull;
    x = (x | x << 2) & 0x1249249249249249ull;
    return x;
}
// morton3_d_get_third_bits, src/morton.rs:61-68.
TC_DEV uint32_t d_compact3_10(uint32_t m) {
    uint32_t x = m & 0x09249249u;
    x = (x ^ (x >> 2)) & 0x030c30c3u;
    x = (x ^ (x >> 4)) & 0x0300f00fu;
    x = (x ^ (x >> 8)) & 0x030000ffu;
    x = (x ^ (x >> 16)) & 0x000003ffu;
    return x;
}
TC_DEV uint64_t d_compact3_21(uint64_t m) {
    uint64_t x = m & 0x1249249249249249ull;
    x = (x ^ (x >> 2)) & 0x10c30c30c30c30c3ull;
    x = (x ^ (x >> 4)) & 0x100f00f00f00f00full;
    x = (x ^ (x >> 8)) & 0x001f0000ff0000ffull;
    x = (x ^ (x >> 16)) & 0x001f00000000ffffull;
    x = (x ^ (x >> 32)) & 0x00000000001fffffull;
    return x;
}

// GLSL uint(float) with out-of-range input made well defined: NaN and
// negatives give 0, values >= 2^32 give 0xFFFFFFFF, otherwise truncation.
// Branch-free: v_max_f32 returns the non-NaN operand, so NaN becomes 0.
TC_DEV uint32_t d_f2u_sat(float t) {
    const float c = fmaxf(t, 0.0f);
    return c >= 4294967296.0f ? 0xFFFFFFFFu : (uint32_t)c;
}

// scale / range when it is an exact power of two that is a normal fp32 value,
// else 0.  For such a ratio, (t / range) * scale and t * (scale / range) give
// the same quantised integer for every t: dividing by a power of two and
// multiplying by one are exact unless the quotient is subnormal, and a
// subnormal quotient times scale (<= 2^21) is below 1, which both forms
// truncate to 0 (an overflow gives +inf in both).  `scale` is a power of two.
TC_DEV float d_pow2_ratio(float range, float scale) {
    const uint32_t rb = __float_as_uint(range);
    const int re = (int)((rb >> 23) & 255u);
    if ((rb & 0x807FFFFFu) != 0u || re == 0 || re == 255) return 0.0f;  // not a positive normal power of two
    const int se = (int)((__float_as_uint(scale) >> 23) & 255u);
    const int e = se - re + 127;  // biased exponent of scale / range
    return (e >= 1 && e <= 254) ? __uint_as_float((uint32_t)e << 23) : 0.0f;
}

// The Morton quantiser of one cloud: per-axis integer grid coordinate
// ((v - min) / range) * scale with separately rounded fp32 steps
// (single_point_to_code_v2, shader/morton.comp:50-56), with one multiply in
// place of the division and the scaling when scale / range is an exact power
// of two (every range 2^k).
struct MortonQuant {
    float min_coord, range, scale, ratio;
    TC_DEV MortonQuant(float mn, float rg, float sc) : min_coord(mn), range(rg), scale(sc), ratio(d_pow2_ratio(rg, sc)) {}
    TC_DEV uint32_t operator()(float v) const {
        const float t = __fsub_rn(v, min_coord);
        if (ratio != 0.0f) return d_f2u_sat(__fmul_rn(t, ratio));
        return d_f2u_sat(__fmul_rn(__fdiv_rn(t, range), scale));
    }
};

// 30-bit code of a point (k_ComputeMortonCode, shader/morton.comp:60-67).
TC_DEV uint32_t d_encode30(float x, float y, float z, float min_coord, float range) {
    const MortonQuant q(min_coord, range, 1024.0f);
    return d_split3_10(q(x)) | (d_split3_10(q(y)) << 1) | (d_split3_10(q(z)) << 2);
}
// 63-bit code: the same quantiser with a 2^21 grid.
TC_DEV uint64_t d_encode63(float x, float y, float z, float min_coord, float range) {
    const MortonQuant q(min_coord, range, 2097152.0f);
    return d_split3_21(q(x)) | (d_split3_21(q(y)) << 1) | (d_split3_21(q(z)) << 2);
}

// Code traits: 30-bit codes in u32, 63-bit codes in u64.
template <typename K> struct CodeTraits;
template <> struct CodeTraits<uint32_t> {
    static constexpr int BITS = 30;
    static constexpr int AXIS_BITS = 10;
    TC_DEV static int delta_bits(uint32_t x) { return x ? __clz(x) - 2 : 30; }
    // The cell corner of integer coordinates q (code = interleave(q)).
    TC_DEV static float corner(uint32_t q, float min_coord, float range) {
        float f = (float)q;
        f = __fmul_rn(f, 0x1p-10f);  // exactly f / 1024: a power-of-two scale
        f = __fmul_rn(f, range);
        return __fadd_rn(f, min_coord);
    }
    TC_DEV static uint32_t axis(uint32_t code, int a) { return d_compact3_10(code >> a); }
    TC_DEV static void decode(uint32_t code, float min_coord, float range, float* out) {
#pragma unroll
        for (int a = 0; a < 3; ++a) out[a] = corner(axis(code, a), min_coord, range);
        out[3] = 1.0f;
    }
};
template <> struct CodeTraits<uint64_t> {
    static constexpr int BITS = 63;
    static constexpr int AXIS_BITS = 21;
    TC_DEV static int delta_bits(uint64_t x) { return x ? __clzll(x) - 1 : 63; }
    TC_DEV static float corner(uint32_t q, float min_coord, float range) {
        float f = (float)q;
        f = __fmul_rn(f, 0x1p-21f);  // exactly f / 2^21
        f = __fmul_rn(f, range);
        return __fadd_rn(f, min_coord);
    }
    TC_DEV static uint32_t axis(uint64_t code, int a) { return (uint32_t)d_compact3_21(code >> a); }
    TC_DEV static void decode(uint64_t code, float min_coord, float range, float* out) {
#pragma unroll
        for (int a = 0; a < 3; ++a) out[a] = corner(axis(code, a), min_coord, range);
        out[3] = 1.0f;
    }
};
