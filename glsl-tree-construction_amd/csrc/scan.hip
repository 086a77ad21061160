// scan.hip -- single-pass tile scans with decoupled look-back: unique
// compaction, exclusive scan, the fused edge-count + exclusive scan, and the
// segmented (many clouds per launch) forms.
//
// Reference: k_CountUnique, shader/unique.comp:31-45 and k_PartialSum,
// shader/partial_sum.comp:23-39 are single-thread loops ([1,1,1] dispatches,
// src/unique.rs:245, src/partial_sum.rs:246); countEdges,
// shader/edge_count.comp:30-41.  Published algorithm: Merrill & Garland,
// "Single-pass Parallel Prefix Scan with Decoupled Look-back" (2016).
//
// gfx950 design.  Tile = workgroup id (a claim counter would serialise at
// about 11 ns per claim on MI355X).  Tile status words are u64 [value:32 |
// flag:2 | epoch:30], one per 64-byte line: adjacent tiles run on different
// XCDs, and sharing lines between them made each look-back poll slower (2^26
// sorted codes, 16K-key tiles: 92.8 -> 77.9 us).  A one-lane kernel advances
// the device-side epoch before each scan, so status words never need clearing
// (also under graph replay).  Wave 0 looks back over 64 predecessors per poll
// (measured: reading 2-4 words per lane, or spreading the poll over 2-8 waves,
// halves the polls but makes each slower, a net loss).  The wait is bounded: a
// look-back that has polled spin_limit times without progress stops at the
// predecessor that has not published, the workgroup reduces that tile's
// elements itself and the look-back goes on below it.  So progress never
// depends on the order in which workgroups start, and a predecessor that has
// not started (its XCD may be held by another kernel's workgroups) costs one
// tile read, not a wait on that other kernel.
//
// Unique (k_unique_tiles): 512 threads x 8 rows of 16-byte vectors (16384 u32
// or 8192 u64 keys per tile) loaded straight into registers; the key before
// each wave's run of a row comes from a scalar load, the others from the lane
// below.  Run-start flags are counted per row in 16-bit fields of two u64s,
// so one block scan gives every row's offsets.  The tile publishes its count,
// compacts its first four rows into LDS, then looks back; the stores after it
// are contiguous.  (Measured on 2^26 sorted 30-bit codes of the Gaussian
// mixture: 132 us for the LDS-staged 8192-key tiles this replaces; 61.8 us with
// exact prefixes given and no look-back at all.)
//
// Scans (k_scan_tiles): one 8192-element tile per workgroup, loaded with 16 B
// vector loads into LDS and read back so each thread owns 16 consecutive
// elements; a block scan gives tile-local offsets.  Results go back through
// LDS so stores are contiguous too.
//
// Segmented modes run one scan over many concatenated clouds: an element that
// starts a cloud always counts as a new unique key, and the exclusive prefix at
// each cloud's first element is written to seg_out[cloud] (its base in the
// compacted output).
#include "tc_internal.hpp"

namespace {

constexpr int SCAN_BLOCK = 512;
constexpr int SCAN_IPT = 16;
constexpr uint32_t SCAN_TILE = SCAN_BLOCK * SCAN_IPT;
// u64 words between two tiles' status words (64 bytes).
constexpr uint32_t STATUS_STRIDE = 8;

// Scan control words (in ctx->small).
struct ScanCtl {
    uint32_t reserved0;
    uint32_t reserved1;
    uint32_t seq;
};

struct ScanArgs {
    const void* in0;
    const void* in1;
    void* out0;
    void* out1;
    uint32_t* count_out;
    uint32_t n_cap;
    const uint32_t* d_n_keys;
    const uint32_t* seg_off;  // segmented: [n_seg + 1] starts; total = seg_off[n_seg]
    uint32_t n_seg_cap;
    const uint32_t* d_n_seg;  // segmented: device segment count (NULL: n_seg_cap)
    uint32_t* seg_out;        // segmented: [n_seg] exclusive prefix at each segment start
    uint64_t* status;
    ScanCtl* ctl;
    uint32_t spin_limit;
};

constexpr bool is_seg(int m) { return m == MODE_SEG_UNIQUE32 || m == MODE_SEG_UNIQUE64 || m == MODE_SEG_SCAN; }
constexpr bool is_unique(int m) {
    return m == MODE_UNIQUE32 || m == MODE_UNIQUE64 || m == MODE_SEG_UNIQUE32 || m == MODE_SEG_UNIQUE64;
}
constexpr bool is_u64(int m) { return m == MODE_UNIQUE64 || m == MODE_SEG_UNIQUE64; }
constexpr bool is_plain_values(int m) { return m == MODE_SCAN || m == MODE_SEG_SCAN; }

template <int MODE>
using ItemT = typename std::conditional<is_u64(MODE), uint64_t, uint32_t>::type;

TC_DEV uint64_t st_word(uint32_t flag, uint32_t epoch, uint32_t value) {
    return ((uint64_t)((epoch << 2) | flag) << 32) | value;
}

TC_DEV uint32_t wave_sum(uint32_t v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Look-back by wave 0 from predecessor `top` down, adding into `excl` (both
// uniform over the wave).  Returns true once the prefix is complete (an
// inclusive predecessor, or the start of the array); false after spin_limit
// polls without progress: the predecessors above `top` are then added and
// tile `top` is the one that has not published.
TC_DEV bool wave_lookback(const uint64_t* status, uint32_t epoch, uint32_t spin_limit, int64_t& top,
                          uint32_t& excl) {
    const uint32_t lane = lane_id();
    uint32_t spins = 0;
    while (true) {
        const int64_t idx = top - (int64_t)lane;
        uint32_t f, v;
        if (idx >= 0) {
            const uint64_t w = ld_agent64(&status[idx * STATUS_STRIDE]);
            const uint32_t hi = (uint32_t)(w >> 32);
            f = ((hi >> 2) == epoch) ? (hi & 3u) : 0u;
            v = (uint32_t)w;
        } else {
            f = 2u;  // before tile 0: an inclusive zero
            v = 0u;
        }
        const uint64_t incl_mask = __ballot(f == 2u);
        const uint64_t nr = __ballot(f == 0u);
        // Lanes up to and including the nearest inclusive predecessor.
        const uint64_t upto = incl_mask ? ((incl_mask & (~incl_mask + 1)) << 1) - 1 : ~0ull;
        const uint64_t wait = nr & upto;
        if (wait == 0) {
            excl += wave_sum(((upto >> lane) & 1ull) ? v : 0u);
            if (incl_mask) return true;
            top -= 64;
            continue;
        }
        if (++spins > spin_limit) {
            // The lanes below the first unpublished one hold aggregates only
            // (an inclusive one among them would have ended the walk).
            const uint64_t ready = (wait & (~wait + 1)) - 1;
            excl += wave_sum(((ready >> lane) & 1ull) ? v : 0u);
            top -= (int64_t)__popcll(ready);
            return false;
        }
        __builtin_amdgcn_s_sleep(1);
    }
}

// Largest c in [0, nseg-1] with off[c] <= i (skips empty segments).
TC_DEV uint32_t find_seg(const uint32_t* off, uint32_t nseg, uint32_t i) {
    uint32_t lo = 0, hi = nseg;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (off[mid] <= i) lo = mid;
        else hi = mid;
    }
    return lo;
}

// Scan value of element i (the fall-back path's reduction over [0, base)).
template <int MODE>
TC_DEV uint32_t elem_value(const ScanArgs& a, uint32_t n, uint32_t i, uint32_t nseg, uint32_t& seg,
                           uint32_t& seg_start, uint32_t& seg_next) {
    using T = ItemT<MODE>;
    if (is_plain_values(MODE)) return static_cast<const uint32_t*>(a.in0)[i];
    if (is_unique(MODE)) {
        const T* in = static_cast<const T*>(a.in0);
        bool starts = (i == 0);
        if (is_seg(MODE)) {
            while (i >= seg_next && seg + 1 < nseg) {
                ++seg;
                seg_start = a.seg_off[seg];
                seg_next = a.seg_off[seg + 1];
            }
            starts = starts || (i == seg_start);
        }
        return (starts || in[i] != in[i - 1]) ? 1u : 0u;
    }
    // MODE_ECSCAN
    const uint8_t* pn = static_cast<const uint8_t*>(a.in0);
    uint32_t p = (uint32_t)static_cast<const int32_t*>(a.in1)[i];
    p = p < n ? p : 0u;
    return i == 0 ? 1u : (uint32_t)(pn[i] / 3 - pn[p] / 3);
}

// The look-back's fall-back: the sum of the scan values of [lo, hi), by the
// whole block (each thread sums a contiguous chunk).  Every thread calls it.
template <int MODE>
TC_DEV uint32_t reduce_range(const ScanArgs& a, uint32_t n, uint32_t nseg, uint32_t lo, uint32_t hi,
                             uint32_t* s_tmp) {
    const uint32_t chunk = ceil_div_u32(hi - lo, SCAN_BLOCK);
    const uint32_t t0 = lo + threadIdx.x * chunk;
    const uint32_t t1 = t0 + chunk < hi ? t0 + chunk : hi;
    uint32_t sum = 0, seg = 0, seg_start = 0, seg_next = 0xFFFFFFFFu;
    if (is_seg(MODE) && nseg > 0 && t0 < t1) {
        seg = find_seg(a.seg_off, nseg, t0);
        seg_start = a.seg_off[seg];
        seg_next = a.seg_off[seg + 1];
    }
    for (uint32_t i = t0; i < t1; ++i) sum += elem_value<MODE>(a, n, i, nseg, seg, seg_start, seg_next);
    uint32_t total;
    block_exclusive_scan<SCAN_BLOCK>(sum, s_tmp, &total);
    return total;
}

// Exclusive prefix of `tile` (> 0, tiles of tile_size elements), whose
// aggregate is published: wave 0
// looks back; a predecessor that does not publish within spin_limit polls is
// reduced by the whole block (spin_limit 0: every predecessor, used to test
// that path).  Publishes the tile's inclusive value.  Every thread calls it;
// the other waves may still be busy with their own work when they do.
template <int MODE>
TC_DEV uint32_t tile_prefix(const ScanArgs& a, uint32_t n, uint32_t nseg, uint32_t tile, uint32_t tile_size,
                            uint32_t epoch, uint32_t block_total, uint32_t* s_tmp, uint32_t* s_stall,
                            uint32_t* s_excl) {
    const bool w0 = (threadIdx.x >> 6) == 0;
    int64_t top = (int64_t)tile - 1;
    uint32_t excl = 0;
    while (true) {
        if (w0) {
            const bool done = top < 0 || (a.spin_limit != 0u && wave_lookback(a.status, epoch, a.spin_limit, top, excl));
            if (lane_id() == 0) {
                *s_stall = done ? 0u : (uint32_t)top + 1u;
                if (done) {
                    *s_excl = excl;
                    st_agent64(&a.status[(size_t)tile * STATUS_STRIDE], st_word(2u, epoch, excl + block_total));
                }
            }
        }
        __syncthreads();
        const uint32_t stalled = *s_stall;
        if (stalled == 0u) break;  // uniform
        const uint32_t lo = (stalled - 1u) * tile_size;
        const uint32_t r = reduce_range<MODE>(a, n, nseg, lo, lo + tile_size < n ? lo + tile_size : n, s_tmp);
        // (reduce_range's barriers order every read of *s_stall before wave 0 writes it again.)
        if (w0) {
            excl += r;
            top = (int64_t)stalled - 2;
        }
    }
    return *s_excl;
}

// Coalesced tile load of u32 scan values into LDS (tile-relative positions).
TC_DEV void load_tile(const uint32_t* __restrict__ in, uint32_t base, uint32_t cnt, uint32_t* s, bool aligned) {
    if (aligned && cnt == SCAN_TILE) {
        const uint4* vin = reinterpret_cast<const uint4*>(in + base);
        uint4* vs = reinterpret_cast<uint4*>(s);
#pragma unroll
        for (int r = 0; r < (int)(SCAN_TILE / 4 / SCAN_BLOCK); ++r) vs[r * SCAN_BLOCK + threadIdx.x] = vin[r * SCAN_BLOCK + threadIdx.x];
    } else {
        for (uint32_t k = threadIdx.x; k < cnt; k += SCAN_BLOCK) s[k] = in[base + k];
    }
}

__global__ void k_scan_prep(ScanCtl* ctl) {
    // Advance the epoch tag: status words of earlier scans read as not ready.
    ctl->seq = (ctl->seq + 1u) & 0x3FFFFFFFu;
}

// Exclusive scans (plain, segmented, and the fused edge count).
template <int MODE>
__global__ __launch_bounds__(SCAN_BLOCK) void k_scan_tiles(ScanArgs a) {
    static_assert(!is_unique(MODE), "unique modes run k_unique_tiles");
    __shared__ uint32_t s_tmp[SCAN_BLOCK / 64 + 2];
    __shared__ uint32_t s_excl;
    __shared__ uint32_t s_stall;  // look-back: the stalled predecessor + 1 (0: none)
    __shared__ uint32_t s_seg0;  // segmented modes: the segment of the tile's first element
    __shared__ __attribute__((aligned(16))) uint32_t s_items[SCAN_TILE];

    const uint32_t tid = threadIdx.x;
    const uint32_t wave = tid >> 6;

    uint32_t n = a.n_cap;
    uint32_t nseg = 0;
    if (is_seg(MODE)) {
        nseg = a.d_n_seg ? *a.d_n_seg : a.n_seg_cap;
        if (nseg > a.n_seg_cap) nseg = a.n_seg_cap;
        n = nseg ? a.seg_off[nseg] : 0u;
        if (n > a.n_cap) n = a.n_cap;
    } else if (a.d_n_keys) {
        const uint32_t u = *a.d_n_keys;
        const uint32_t m = u > 0 ? u - 1 : 0;  // the scans run over the n_keys - 1 radix nodes
        n = m < a.n_cap ? m : a.n_cap;
    }
    const uint32_t num_tiles = n ? ceil_div_u32(n, SCAN_TILE) : 1u;
    const uint32_t tile = blockIdx.x;
    if (tile >= num_tiles) return;  // capacity grids: workgroups past the data leave at once
    const uint32_t epoch = a.ctl->seq & 0x3FFFFFFFu;
    // Unaligned inputs (tile starts not 16 B aligned) take the element path.
    const bool aligned = (reinterpret_cast<uintptr_t>(a.in0) & 15u) == 0;

    const uint32_t base = tile * SCAN_TILE;
    const uint32_t cnt = n > base ? (n - base < SCAN_TILE ? n - base : SCAN_TILE) : 0u;
    const uint32_t r0 = tid * SCAN_IPT;  // tile-relative index of this thread's first element
    const uint32_t i0 = base + r0;
    uint32_t val[SCAN_IPT];
    auto V = [&](int j) -> uint32_t { return val[j]; };
    uint32_t start_mask = 0;  // bit j: element i0 + j starts a segment
    uint32_t seg_first = 0;

    // Wave 0 finds the tile's first segment (64 candidates per load round)
    // while the other waves load the tile; each thread then steps forward
    // from it instead of searching all segments itself.
    if (is_seg(MODE) && wave == 0 && nseg > 0 && base < n) {
        const uint32_t c = seg_find_wave(a.seg_off, nseg, base);
        if (lane_id() == 0) s_seg0 = c;
    }
    if (MODE != MODE_ECSCAN) {  // (segmented modes always take this sync)
        load_tile(static_cast<const uint32_t*>(a.in0), base, cnt, s_items, aligned);
        __syncthreads();
    }
    if (is_seg(MODE) && nseg > 0 && i0 < n) {
        uint32_t seg = seg_find_near(a.seg_off, nseg, i0, s_seg0);
        seg_first = seg;
        uint32_t seg_start = a.seg_off[seg], seg_next = a.seg_off[seg + 1];
#pragma unroll
        for (int j = 0; j < SCAN_IPT; ++j) {
            const uint32_t i = i0 + j;
            if (i >= n) break;
            while (i >= seg_next && seg + 1 < nseg) {
                ++seg;
                seg_start = a.seg_off[seg];
                seg_next = a.seg_off[seg + 1];
            }
            if (i == seg_start) start_mask |= 1u << j;
        }
    }

    if (is_plain_values(MODE)) {
#pragma unroll
        for (int j = 0; j < SCAN_IPT; ++j) val[j] = (r0 + j < cnt) ? s_items[r0 + j] : 0u;
    } else {  // MODE_ECSCAN: edge count of radix node i, root entry 1
        uint32_t z0 = 0, z1 = 0, z2 = 0;
#pragma unroll
        for (int j = 0; j < SCAN_IPT; ++j) val[j] = (i0 + j < n) ? elem_value<MODE>(a, n, i0 + j, 0, z0, z1, z2) : 0u;
    }

    uint32_t tsum = 0;
#pragma unroll
    for (int j = 0; j < SCAN_IPT; ++j) tsum += V(j);
    uint32_t block_total;
    const uint32_t texcl = block_exclusive_scan<SCAN_BLOCK>(tsum, s_tmp, &block_total);

    uint32_t tile_excl = 0;
    if (tile == 0) {
        if (tid == 0) st_agent64(&a.status[0], st_word(2u, epoch, block_total));
    } else {
        if (tid == 0) st_agent64(&a.status[(size_t)tile * STATUS_STRIDE], st_word(1u, epoch, block_total));
        tile_excl = tile_prefix<MODE>(a, n, nseg, tile, SCAN_TILE, epoch, block_total, s_tmp, &s_stall, &s_excl);
    }

    // Segment bases: the exclusive prefix at each segment's first element.
    if (is_seg(MODE) && start_mask) {
        uint32_t seg = seg_first;
        uint32_t run = tile_excl + texcl;
#pragma unroll
        for (int j = 0; j < SCAN_IPT; ++j) {
            if ((start_mask >> j) & 1u) {
                // the last segment starting here (earlier equal offsets are empty clouds)
                while (seg + 1 < nseg && a.seg_off[seg + 1] <= i0 + j) ++seg;
                a.seg_out[seg] = run;
            }
            run += V(j);
        }
    }

    if (is_plain_values(MODE)) {
        // Exclusive values through LDS, then contiguous stores.
        uint32_t run = tile_excl + texcl;
        uint32_t* s_out = reinterpret_cast<uint32_t*>(s_items);
        __syncthreads();  // everyone has read its items
#pragma unroll
        for (int j = 0; j < SCAN_IPT; ++j) {
            s_out[r0 + j] = run;
            run += V(j);
        }
        __syncthreads();
        uint32_t* out = static_cast<uint32_t*>(a.out0);
        for (uint32_t k = tid; k < cnt; k += SCAN_BLOCK) out[base + k] = s_out[k];
    } else {
        uint32_t run = tile_excl + texcl;
        uint32_t* ec = static_cast<uint32_t*>(a.out0);
        uint32_t* off = static_cast<uint32_t*>(a.out1);
#pragma unroll
        for (int j = 0; j < SCAN_IPT; ++j) {
            if (i0 + j < n) {
                ec[i0 + j] = V(j);
                off[i0 + j] = run;
            }
            run += V(j);
        }
    }
    if (tile == num_tiles - 1 && tid == SCAN_BLOCK - 1) {
        const uint32_t total = tile_excl + block_total;
        if (is_plain_values(MODE)) static_cast<uint32_t*>(a.out0)[n] = total;
        if (MODE == MODE_ECSCAN) static_cast<uint32_t*>(a.out1)[n] = total;
        // MODE_SCAN reports [element count, total] (the build's n_brt and
        // octree node count); the other modes their total.
        if (a.count_out) {
            if (MODE == MODE_SCAN) {
                a.count_out[0] = n;
                a.count_out[1] = total;
            } else {
                *a.count_out = total;
            }
        }
    }
}


// ------------------------------------------------------------------ unique
constexpr int UQ_ROWS = 8;
constexpr int SCAN_WAVES = SCAN_BLOCK / 64;

template <typename T>
struct UqShape {
    static constexpr uint32_t VEC = 16 / sizeof(T);       // keys per 16-byte load
    static constexpr uint32_t ROW = SCAN_BLOCK * VEC;      // keys per row
    static constexpr uint32_t TILE = ROW * UQ_ROWS;        // 16384 u32 / 8192 u64
    static constexpr uint32_t HALF = ROW * (UQ_ROWS / 2);  // keys compacted per LDS round
};

uint32_t unique_tile_keys(int mode) {
    return (mode == MODE_UNIQUE64 || mode == MODE_SEG_UNIQUE64) ? UqShape<uint64_t>::TILE : UqShape<uint32_t>::TILE;
}

// Inclusive wave scan of four 16-bit counters packed in a u64 (no field
// overflows: a row holds at most 2048 keys).
TC_DEV uint64_t wave_inclusive_scan64(uint64_t v) {
    const uint32_t lane = lane_id();
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const uint64_t t = __shfl_up(v, off, 64);
        if (lane >= (uint32_t)off) v += t;
    }
    return v;
}

TC_DEV void unpack16(const uint4& v, uint32_t (&k)[4]) {
    k[0] = v.x;
    k[1] = v.y;
    k[2] = v.z;
    k[3] = v.w;
}
TC_DEV void unpack16(const ulonglong2& v, uint64_t (&k)[2]) {
    k[0] = v.x;
    k[1] = v.y;
}

// Unique compaction of sorted keys (plain and segmented).  Slot q of thread t
// in row r holds key base + r * ROW + t * VEC + q, so a tile is visited in key
// order row by row; a run starts at a key that differs from the one before it
// or that starts a cloud.
template <int MODE>
__global__ __launch_bounds__(SCAN_BLOCK) void k_unique_tiles(ScanArgs a) {
    using T = ItemT<MODE>;
    using Sh = UqShape<T>;
    using V = typename std::conditional<sizeof(T) == 4, uint4, ulonglong2>::type;
    constexpr bool SEG = is_seg(MODE);
    constexpr uint32_t VEC = Sh::VEC, ROW = Sh::ROW, TILE = Sh::TILE;
    constexpr int NP = UQ_ROWS / 4;  // u64 counter words per thread
    constexpr int HALF_ROWS = UQ_ROWS / 2;
    constexpr uint32_t VMASK = (1u << VEC) - 1u;
    static_assert(UQ_ROWS * VEC <= 32, "one flag word per thread");
    static_assert(TILE <= 65536, "tile-local offsets fit 16 bits");

    __shared__ uint64_t s_w[NP * SCAN_WAVES];
    __shared__ uint32_t s_tmp[SCAN_WAVES + 2];
    __shared__ uint32_t s_excl;
    __shared__ uint32_t s_stall;  // look-back: the stalled predecessor + 1 (0: none)
    __shared__ uint32_t s_seg[2];  // segmented: the clouds starting in this tile, [s_seg[0], s_seg[1])
    __shared__ __attribute__((aligned(16))) T s_keys[Sh::HALF];
    // Segmented: for each thread's slots of each row, the tile-local output
    // index of its first kept key (low 16 bits) and its run-start flags (bits
    // 16 and up); and a bitmap of the tile's cloud starts.
    __shared__ uint32_t s_chunk[SEG ? UQ_ROWS * SCAN_BLOCK : 1];
    __shared__ uint32_t s_start[SEG ? TILE / 32 : 1];

    const uint32_t tid = threadIdx.x;
    const uint32_t wave = tid >> 6;
    const uint32_t lane = lane_id();

    uint32_t n = a.n_cap;
    uint32_t nseg = 0;
    if (SEG) {
        nseg = a.d_n_seg ? *a.d_n_seg : a.n_seg_cap;
        if (nseg > a.n_seg_cap) nseg = a.n_seg_cap;
        n = nseg ? a.seg_off[nseg] : 0u;
        if (n > a.n_cap) n = a.n_cap;
    } else if (a.d_n_keys) {
        const uint32_t u = *a.d_n_keys;
        n = u < a.n_cap ? u : a.n_cap;
    }
    const uint32_t num_tiles = n ? ceil_div_u32(n, TILE) : 1u;
    const uint32_t tile = blockIdx.x;
    if (tile >= num_tiles) return;  // capacity grids: workgroups past the data leave at once
    const uint32_t epoch = a.ctl->seq & 0x3FFFFFFFu;
    const uint32_t base = tile * TILE;
    const uint32_t cnt = n > base ? (n - base < TILE ? n - base : TILE) : 0u;
    const T* in = static_cast<const T*>(a.in0);

    if (SEG) {
        for (uint32_t w = tid; w < TILE / 32; w += SCAN_BLOCK) s_start[w] = 0u;
        // Waves 0 and 1: the first cloud starting at or after the tile's first
        // key, and at or after its end.
        if (wave < 2) {
            const uint32_t x = wave == 0 ? base : base + cnt;
            uint32_t c = 0;
            if (nseg > 0 && cnt > 0 && x > 0) c = seg_find_wave(a.seg_off, nseg, x - 1) + 1u;
            if (lane == 0) s_seg[wave] = c;
        }
    }

    // The tile, straight into registers (16-byte loads of keys read once).
    T k[UQ_ROWS][VEC];
    const bool aligned = (reinterpret_cast<uintptr_t>(in) & 15u) == 0;
    if (aligned && cnt == TILE) {
        const V* vin = reinterpret_cast<const V*>(in + base);
#pragma unroll
        for (int r = 0; r < UQ_ROWS; ++r) unpack16(ld_stream(&vin[r * SCAN_BLOCK + tid]), k[r]);
    } else {
#pragma unroll
        for (int r = 0; r < UQ_ROWS; ++r)
#pragma unroll
            for (uint32_t q = 0; q < VEC; ++q) {
                const uint32_t e = r * ROW + tid * VEC + q;
                k[r][q] = e < cnt ? in[base + e] : (T)0;
            }
    }
    // The key before each wave's slots of a row (a wave-uniform index).
    const uint32_t wv = __builtin_amdgcn_readfirstlane(wave);
    T before_wave[UQ_ROWS];
#pragma unroll
    for (int r = 0; r < UQ_ROWS; ++r) {
        const uint32_t e = base + r * ROW + wv * 64 * VEC;
        before_wave[r] = (e > 0 && e - 1 < n) ? in[e - 1] : (T)0;
    }
    if (SEG) {
        __syncthreads();  // bitmap cleared, s_seg written
        const uint32_t c_lo = s_seg[0], c_hi = s_seg[1];
        for (uint32_t c = c_lo + tid; c < c_hi; c += SCAN_BLOCK) {
            const uint32_t q = a.seg_off[c] - base;
            atomicOr(&s_start[q >> 5], 1u << (q & 31u));
        }
        __syncthreads();
    }

    // Run-start flags, bit r * VEC + q.
    uint32_t keep = 0;
#pragma unroll
    for (int r = 0; r < UQ_ROWS; ++r) {
        const T up = __shfl_up(k[r][VEC - 1], 1, 64);
        T prev = lane == 0 ? before_wave[r] : up;
        const uint32_t e0 = r * ROW + tid * VEC;
        const uint32_t starts = SEG ? (s_start[e0 >> 5] >> (e0 & 31u)) & VMASK : 0u;
#pragma unroll
        for (uint32_t q = 0; q < VEC; ++q) {
            const bool run = (base + e0 + q == 0) || ((starts >> q) & 1u) || k[r][q] != prev;
            keep |= ((e0 + q < cnt && run) ? 1u : 0u) << (r * VEC + q);
            prev = k[r][q];
        }
    }
    auto row_flags = [&](int r) -> uint32_t { return (keep >> (r * VEC)) & VMASK; };

    // Per-row counts in 16-bit fields, four rows per u64: one block scan gives
    // every row's offsets.
    uint64_t pk[NP], inc[NP];
#pragma unroll
    for (int p = 0; p < NP; ++p) {
        pk[p] = 0;
#pragma unroll
        for (int q = 0; q < 4; ++q) pk[p] |= (uint64_t)__popc(row_flags(4 * p + q)) << (16 * q);
        inc[p] = wave_inclusive_scan64(pk[p]);
        if (lane == 63) s_w[p * SCAN_WAVES + wave] = inc[p];
    }
    __syncthreads();
    uint32_t row_base[UQ_ROWS];  // tile-local output index of each row's first kept key
    uint32_t slot[UQ_ROWS];      // ... and of this thread's first kept key in each row
    uint32_t block_total = 0;
#pragma unroll
    for (int p = 0; p < NP; ++p) {
        uint64_t below = 0, tot = 0;
#pragma unroll
        for (int w = 0; w < SCAN_WAVES; ++w) {
            const uint64_t x = s_w[p * SCAN_WAVES + w];
            below += ((uint32_t)w < wave) ? x : 0ull;
            tot += x;
        }
        const uint64_t ex = below + inc[p] - pk[p];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            row_base[4 * p + q] = block_total;
            slot[4 * p + q] = block_total + ((uint32_t)(ex >> (16 * q)) & 0xFFFFu);
            block_total += (uint32_t)(tot >> (16 * q)) & 0xFFFFu;
        }
    }
    if (wave == 0 && lane == 0)
        st_agent64(&a.status[(size_t)tile * STATUS_STRIDE], st_word(tile == 0 ? 2u : 1u, epoch, block_total));
    if (SEG) {
#pragma unroll
        for (int r = 0; r < UQ_ROWS; ++r) s_chunk[r * SCAN_BLOCK + tid] = slot[r] | (row_flags(r) << 16);
    }

    // Kept keys (or their positions) of rows [r0, r0 + HALF_ROWS) into LDS,
    // indexed from the first of them.
    auto compact_keys = [&](int r0) {
#pragma unroll
        for (int r = r0; r < r0 + HALF_ROWS; ++r) {
            uint32_t pos = slot[r] - row_base[r0];
#pragma unroll
            for (uint32_t q = 0; q < VEC; ++q) {
                const uint32_t b = (row_flags(r) >> q) & 1u;
                if (b) s_keys[pos] = k[r][q];
                pos += b;
            }
        }
    };
    uint32_t* s_pos = reinterpret_cast<uint32_t*>(s_keys);
    auto compact_positions = [&](int r0) {
#pragma unroll
        for (int r = r0; r < r0 + HALF_ROWS; ++r) {
            uint32_t pos = slot[r] - row_base[r0];
#pragma unroll
            for (uint32_t q = 0; q < VEC; ++q) {
                const uint32_t b = (row_flags(r) >> q) & 1u;
                if (b) s_pos[pos] = base + r * ROW + tid * VEC + q;
                pos += b;
            }
        }
    };

    // First half into LDS while wave 0 looks back.
    compact_keys(0);
    uint32_t tile_excl = 0;
    if (tile == 0) __syncthreads();  // the staged keys are read below (tile_prefix's barriers do this otherwise)
    else tile_excl = tile_prefix<MODE>(a, n, nseg, tile, TILE, epoch, block_total, s_tmp, &s_stall, &s_excl);

    T* out = static_cast<T*>(a.out0);
    uint32_t* run_first = SEG ? nullptr : static_cast<uint32_t*>(a.out1);
    const uint32_t h0 = row_base[HALF_ROWS];
    for (uint32_t j = tid; j < h0; j += SCAN_BLOCK) out[tile_excl + j] = s_keys[j];
    if (run_first) {
        __syncthreads();
        compact_positions(0);
        __syncthreads();
        for (uint32_t j = tid; j < h0; j += SCAN_BLOCK) run_first[tile_excl + j] = s_pos[j];
    }
    __syncthreads();
    compact_keys(HALF_ROWS);
    __syncthreads();
    for (uint32_t j = tid; j < block_total - h0; j += SCAN_BLOCK) out[tile_excl + h0 + j] = s_keys[j];
    if (run_first) {
        __syncthreads();
        compact_positions(HALF_ROWS);
        __syncthreads();
        for (uint32_t j = tid; j < block_total - h0; j += SCAN_BLOCK) run_first[tile_excl + h0 + j] = s_pos[j];
    }
    if (SEG) {
        // Each cloud's base: the output index of its first key.
        const uint32_t c_lo = s_seg[0], c_hi = s_seg[1];
        for (uint32_t c = c_lo + tid; c < c_hi; c += SCAN_BLOCK) {
            const uint32_t q = a.seg_off[c] - base;
            const uint32_t v = s_chunk[(q / ROW) * SCAN_BLOCK + (q % ROW) / VEC];
            a.seg_out[c] = tile_excl + (v & 0xFFFFu) + __popc((v >> 16) & ((1u << (q % VEC)) - 1u));
        }
    }
    if (tile == num_tiles - 1 && tid == SCAN_BLOCK - 1) {
        const uint32_t total = tile_excl + block_total;
        if (run_first) run_first[total] = n;  // end of the last run
        if (a.count_out) *a.count_out = total;
    }
}

template <int MODE>
void launch(uint32_t grid, hipStream_t s, const ScanArgs& a) {
    if constexpr (is_unique(MODE)) hipLaunchKernelGGL(k_unique_tiles<MODE>, dim3(grid), dim3(SCAN_BLOCK), 0, s, a);
    else hipLaunchKernelGGL(k_scan_tiles<MODE>, dim3(grid), dim3(SCAN_BLOCK), 0, s, a);
}

}  // namespace

uint32_t* tc_scan_epoch(tc_ctx* ctx) { return &reinterpret_cast<ScanCtl*>(ctx->small + SMALL_SCAN_CTR)->seq; }

size_t tc_scan_status_bytes() { return ((size_t)TC_SCAN_MAX_TILES + 1) * STATUS_STRIDE * sizeof(uint64_t); }

int tc_persistent_grid(tc_ctx* ctx, uint32_t work_items, int block, int per_cu) {
    uint32_t blocks = ceil_div_u32(work_items ? work_items : 1, (uint32_t)block);
    const uint32_t cap = (uint32_t)ctx->num_cus * (uint32_t)per_cu;
    return (int)(blocks < cap ? blocks : cap);
}

int tc_launch_scan(tc_ctx* ctx, int mode, const void* in0, const void* in1, void* out0, void* out1,
                   uint32_t* count_out, uint32_t n, const uint32_t* d_n_keys, const uint32_t* seg_off,
                   uint32_t n_seg_cap, const uint32_t* d_n_seg, uint32_t* seg_out, hipStream_t s,
                   bool epoch_advanced) {
    const uint32_t tile_keys = is_unique(mode) ? unique_tile_keys(mode) : SCAN_TILE;
    const uint32_t tiles = ceil_div_u32(n ? n : 1, tile_keys);
    if (!ctx->scan_status) return TC_ERR_NOMEM;
    // Status words for TC_SCAN_MAX_TILES tiles (2^30 / the smallest tile) are preallocated with the context.
    if (tiles > TC_SCAN_MAX_TILES) return TC_ERR_TOO_LARGE;
    ScanCtl* ctl = reinterpret_cast<ScanCtl*>(ctx->small + SMALL_SCAN_CTR);
    ScanArgs a{in0, in1, out0, out1, count_out, n, d_n_keys, seg_off, n_seg_cap, d_n_seg, seg_out,
               ctx->scan_status, ctl, ctx->spin_limit};
    ProfScope prof(ctx, s, is_seg(mode) ? K_SEG : mode == MODE_SCAN ? K_SCAN : mode == MODE_ECSCAN ? K_ECSCAN : K_UNIQUE);
    if (!epoch_advanced) {
        hipLaunchKernelGGL(k_scan_prep, dim3(1), dim3(1), 0, s, ctl);
        TC_LAUNCH_CHECK();
    }
    switch (mode) {
        case MODE_SCAN: launch<MODE_SCAN>(tiles, s, a); break;
        case MODE_UNIQUE32: launch<MODE_UNIQUE32>(tiles, s, a); break;
        case MODE_UNIQUE64: launch<MODE_UNIQUE64>(tiles, s, a); break;
        case MODE_ECSCAN: launch<MODE_ECSCAN>(tiles, s, a); break;
        case MODE_SEG_UNIQUE32: launch<MODE_SEG_UNIQUE32>(tiles, s, a); break;
        case MODE_SEG_UNIQUE64: launch<MODE_SEG_UNIQUE64>(tiles, s, a); break;
        case MODE_SEG_SCAN: launch<MODE_SEG_SCAN>(tiles, s, a); break;
        default: return TC_ERR_INVALID;
    }
    TC_LAUNCH_CHECK();
    return TC_OK;
}

extern "C" int tc_unique_u32(tc_ctx* ctx, const uint32_t* sorted, uint32_t* out, uint32_t n,
                             uint32_t* d_n_unique, tc_stream stream) {
    if (!ctx || !d_n_unique || (n && (!sorted || !out))) return TC_ERR_INVALID;
    if (n > TC_MAX_SORT_N) return TC_ERR_TOO_LARGE;
    if (tc_overlaps(sorted, (size_t)n * sizeof(uint32_t), out, (size_t)n * sizeof(uint32_t))) return TC_ERR_INVALID;
    return tc_launch_scan(ctx, MODE_UNIQUE32, sorted, nullptr, out, nullptr, d_n_unique, n, nullptr, nullptr, 0,
                          nullptr, nullptr, as_stream(stream));
}

extern "C" int tc_unique_u64(tc_ctx* ctx, const uint64_t* sorted, uint64_t* out, uint32_t n,
                             uint32_t* d_n_unique, tc_stream stream) {
    if (!ctx || !d_n_unique || (n && (!sorted || !out))) return TC_ERR_INVALID;
    if (n > TC_MAX_SORT_N) return TC_ERR_TOO_LARGE;
    if (tc_overlaps(sorted, (size_t)n * sizeof(uint64_t), out, (size_t)n * sizeof(uint64_t))) return TC_ERR_INVALID;
    return tc_launch_scan(ctx, MODE_UNIQUE64, sorted, nullptr, out, nullptr, d_n_unique, n, nullptr, nullptr, 0,
                          nullptr, nullptr, as_stream(stream));
}

extern "C" int tc_unique_runs_u32(tc_ctx* ctx, const uint32_t* sorted, uint32_t* out, uint32_t* run_first,
                                  uint32_t n, uint32_t* d_n_unique, tc_stream stream) {
    if (!ctx || !d_n_unique || !run_first || (n && (!sorted || !out))) return TC_ERR_INVALID;
    if (n > TC_MAX_SORT_N) return TC_ERR_TOO_LARGE;
    if (tc_overlaps(sorted, (size_t)n * sizeof(uint32_t), out, (size_t)n * sizeof(uint32_t))) return TC_ERR_INVALID;
    if (tc_overlaps(sorted, (size_t)n * sizeof(uint32_t), run_first, ((size_t)n + 1) * 4)) return TC_ERR_INVALID;
    return tc_launch_scan(ctx, MODE_UNIQUE32, sorted, nullptr, out, run_first, d_n_unique, n, nullptr, nullptr, 0,
                          nullptr, nullptr, as_stream(stream));
}

extern "C" int tc_unique_runs_u64(tc_ctx* ctx, const uint64_t* sorted, uint64_t* out, uint32_t* run_first,
                                  uint32_t n, uint32_t* d_n_unique, tc_stream stream) {
    if (!ctx || !d_n_unique || !run_first || (n && (!sorted || !out))) return TC_ERR_INVALID;
    if (n > TC_MAX_SORT_N) return TC_ERR_TOO_LARGE;
    if (tc_overlaps(sorted, (size_t)n * sizeof(uint64_t), out, (size_t)n * sizeof(uint64_t))) return TC_ERR_INVALID;
    if (tc_overlaps(sorted, (size_t)n * sizeof(uint64_t), run_first, ((size_t)n + 1) * 4)) return TC_ERR_INVALID;
    return tc_launch_scan(ctx, MODE_UNIQUE64, sorted, nullptr, out, run_first, d_n_unique, n, nullptr, nullptr, 0,
                          nullptr, nullptr, as_stream(stream));
}

extern "C" int tc_exclusive_scan_u32(tc_ctx* ctx, const uint32_t* in, uint32_t* out, uint32_t n,
                                     const uint32_t* d_n_keys, tc_stream stream) {
    if (!ctx || !out || (n && !in)) return TC_ERR_INVALID;
    if (n > TC_MAX_SORT_N) return TC_ERR_TOO_LARGE;
    if (tc_overlaps(in, (size_t)n * 4, out, ((size_t)n + 1) * 4)) return TC_ERR_INVALID;
    return tc_launch_scan(ctx, MODE_SCAN, in, nullptr, out, nullptr, nullptr, n, d_n_keys, nullptr, 0, nullptr,
                          nullptr, as_stream(stream));
}

extern "C" int tc_edge_count_scan(tc_ctx* ctx, const uint8_t* prefix_n, const int32_t* parent,
                                  uint32_t* edge_count, uint32_t* offsets, uint32_t n_brt,
                                  const uint32_t* d_n_keys, tc_stream stream) {
    if (!ctx || !offsets || (n_brt && (!prefix_n || !parent || !edge_count))) return TC_ERR_INVALID;
    if (n_brt > TC_MAX_SORT_N) return TC_ERR_TOO_LARGE;
    return tc_launch_scan(ctx, MODE_ECSCAN, prefix_n, parent, edge_count, offsets, nullptr, n_brt, d_n_keys,
                          nullptr, 0, nullptr, nullptr, as_stream(stream));
}

extern "C" int tc_set_lookback_spin_limit(tc_ctx* ctx, uint32_t polls) {
    if (!ctx) return TC_ERR_INVALID;
    ctx->spin_limit = polls;
    return TC_OK;
}
