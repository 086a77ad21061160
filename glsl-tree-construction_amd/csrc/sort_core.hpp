// sort_core.hpp -- the OneSweep tile step shared by the plain sort (sort.hip)
// and the segmented batch sort (batch.hip).
//
// One tile = BLOCK*KPT keys.  Wave w owns keys [w*64*KPT, (w+1)*64*KPT) of the
// tile, lane-striped (key i*64 + lane), so a wave's processing order (i, lane)
// is the keys' index order and the rank below is stable.
#pragma once

#include "tc_internal.hpp"

// Digit width: 8 bits per pass, except that the last pass over Morton codes,
// whose top digit has fewer live bits (30-bit codes: bits 24..29; 63-bit
// codes: bits 56..62), ranks only those bits.  Fewer bits mean fewer ballots
// per key and fewer digits to publish and look back over (measured on MI355X,
// 2^26 keys: 0.189 ms for an 8-bit pass, 0.175 ms for 7 bits, 0.164 ms for 6).
// The digit histogram and the look-back rows keep 256 words per digit
// position / tile whatever the width; a narrow pass uses the first 2^RB.
constexpr uint32_t LB_ROW = 256;

// Live bits of the top digit (shift 8*(bytes-1)) of a Morton code: 30 - 24
// for 30-bit codes, 63 - 56 for 63-bit codes.
template <typename K>
constexpr int morton_top_digit_bits() {
    return sizeof(K) == 4 ? 6 : 7;
}

// Peer mask: lanes of this wave whose RB-bit digit equals mine, among `live`
// lanes (RB ballots; the reference's 32-lane 8-bit version is
// shader/new_radix_sort.comp:147-152).  Must be called with all lanes active.
//
// Per bit b: rep = 0 or ~0 (my bit b, sign-extended), m = ballot of bit b;
// rep ^ m has a 1 for every lane whose bit b differs from mine.  OR-ing those
// over the RB bits gives the lanes with a different digit: 4 VALU ops per bit
// (bit extract, compare, one v_bitop3 per half word) instead of
// select-and-and chains.  The digit is bits [off, off + RB) of `word` (off is uniform, so off + b is a
// scalar operand of each bit extract and the key needs no shift per key).
template <int RB, bool ALL_LIVE = false>
TC_DEV uint64_t match_digit(uint32_t word, uint32_t off, uint64_t live) {
    uint32_t diff_lo = 0u, diff_hi = 0u;
#pragma unroll
    for (int b = 0; b < RB - 1; ++b) {
        uint32_t rep = (uint32_t)__builtin_amdgcn_sbfe((int32_t)word, off + b, 1);
        // Opaque to the optimiser, which would otherwise test the bit again
        // with a shift + compare instead of comparing rep (no instruction).
        asm("" : "+v"(rep));
        const uint64_t m = __ballot(rep != 0u);
        // (rep ^ m) | diff in one v_bitop3 per half (truth table 0xde, operands a=rep, b=diff, c=m).
        diff_lo = __builtin_amdgcn_bitop3_b32(rep, diff_lo, (uint32_t)m, 0xde);
        diff_hi = __builtin_amdgcn_bitop3_b32(rep, diff_hi, (uint32_t)(m >> 32), 0xde);
    }
    uint32_t rep = (uint32_t)__builtin_amdgcn_sbfe((int32_t)word, off + RB - 1, 1);
    asm("" : "+v"(rep));
    const uint64_t m = __ballot(rep != 0u);
    if constexpr (ALL_LIVE) {
        // Every lane live: the last bit's step also complements, ~((rep ^ m) | diff) (table 0x21).
        const uint32_t lo = __builtin_amdgcn_bitop3_b32(rep, diff_lo, (uint32_t)m, 0x21);
        const uint32_t hi = __builtin_amdgcn_bitop3_b32(rep, diff_hi, (uint32_t)(m >> 32), 0x21);
        return ((uint64_t)hi << 32) | lo;
    } else {
        diff_lo = __builtin_amdgcn_bitop3_b32(rep, diff_lo, (uint32_t)m, 0xde);
        diff_hi = __builtin_amdgcn_bitop3_b32(rep, diff_hi, (uint32_t)(m >> 32), 0xde);
        return live & ~(((uint64_t)diff_hi << 32) | diff_lo);
    }
}

template <typename K, int RB = 8>
TC_DEV uint32_t digit_of(K k, uint32_t shift) {
    return (uint32_t)((k >> shift) & ((1u << RB) - 1u));
}

// The 32-bit half of a key that holds the digit at `shift` (a multiple of 8,
// so with RB <= 8 the digit never straddles the halves).
template <typename K>
TC_DEV uint32_t digit_word(K k, uint32_t shift) {
    if constexpr (sizeof(K) == 4) return (uint32_t)k;
    else return shift >= 32u ? (uint32_t)(k >> 32) : (uint32_t)k;
}

template <typename K, int BLOCK, int KPT, int RB = 8>
struct OnesweepShared {
    static constexpr int WAVES = BLOCK / 64;
    static constexpr uint32_t TILE = BLOCK * KPT;
    static constexpr int D = 1 << RB;
    // Per-wave digit counters, then tile-local positions (< TILE <= 65535,
    // so 16 bits); the fall-back digit counts reuse the space afterwards.
    union {
        uint16_t whist[WAVES * D];
        uint32_t fb[D];
    };
    uint32_t gbase[D];
    uint32_t tmp[WAVES + 2];
    uint32_t tile;
    uint32_t aux;
    uint32_t fallback;
    alignas(16) unsigned char stage[TILE * sizeof(K)];
};

// One digit-binning tile step (shader/new_radix_sort.comp:116-314), digit
// d = (key >> shift) & (2^RB - 1):
//   keys sk[base, base+cnt) (cnt <= TILE) go to
//   dk[prefix[d] + add + (#digit-d keys of earlier tiles of this segment) + rank].
// The caller guarantees that the key bits above shift + RB are equal within
// the segment (zero for Morton codes), so RB < 8 sorts like an 8-bit digit.
// `tile` indexes the look-back rows (LB_ROW u32 words per tile) of lb_cur; the
// look-back walks down to `first`, the first tile of the segment, which
// publishes its counts as inclusive.  lb_next (may be NULL) is the next pass's
// look-back row of this tile, cleared here.
// seg_base is the index of the segment's first key (the fall-back counts the
// keys in [seg_base, base)); spin_limit bounds the look-back wait.
template <typename K, bool PAIRS, int BLOCK, int KPT, int RB, bool FULL, bool STREAM>
TC_DEV void onesweep_tile_impl(OnesweepShared<K, BLOCK, KPT, RB>& sm, const K* __restrict__ sk, K* __restrict__ dk,
                               const uint32_t* __restrict__ sv, uint32_t* __restrict__ dv, uint32_t base,
                               uint32_t cnt, uint32_t shift, uint32_t tile, uint32_t first, uint32_t seg_base,
                               const uint32_t* __restrict__ prefix, uint32_t add, uint32_t* __restrict__ lb_cur,
                               uint32_t* __restrict__ lb_next, uint32_t spin_limit) {
    constexpr int WAVES = BLOCK / 64;
    constexpr uint32_t WAVE_TILE = 64 * KPT;
    constexpr uint32_t D = 1u << RB;
    static_assert(RB >= 1 && RB <= 8 && BLOCK >= (int)D, "one thread per digit");
    K* s_keys = reinterpret_cast<K*>(sm.stage);
    uint32_t* s_vals = reinterpret_cast<uint32_t*>(sm.stage);  // reused once the keys are out

    const uint32_t tid = threadIdx.x;
    const uint32_t lane = tid & 63u;
    const uint32_t wave = tid >> 6;
    constexpr bool full = FULL;  // full tiles (cnt == TILE) skip every bounds test
    const uint32_t wofs = wave * WAVE_TILE + lane;  // tile-relative index of key 0 of this lane

    static_assert(BLOCK * KPT <= 65535, "tile positions must fit the 16-bit counters");
    for (uint32_t i = tid; i < WAVES * D; i += BLOCK) sm.whist[i] = 0u;

    K keys[KPT];
    uint32_t vals[KPT];
    uint32_t pos[KPT];
    if constexpr (FULL) {
#pragma unroll
        for (int i = 0; i < KPT; ++i) keys[i] = STREAM ? ld_stream(&sk[base + wofs + i * 64]) : sk[base + wofs + i * 64];
        if (PAIRS) {
#pragma unroll
            for (int i = 0; i < KPT; ++i) vals[i] = STREAM ? ld_stream(&sv[base + wofs + i * 64]) : sv[base + wofs + i * 64];
        }
    } else {
#pragma unroll
        for (int i = 0; i < KPT; ++i) {
            const uint32_t r = wofs + i * 64;
            keys[i] = r < cnt ? sk[base + r] : (K)0;
            if (PAIRS) vals[i] = r < cnt ? sv[base + r] : 0u;
        }
    }
    if (lb_next) {
        for (uint32_t i = tid; i < LB_ROW; i += BLOCK) lb_next[tile * LB_ROW + i] = 0u;
    }
    __syncthreads();  // counters cleared

    // Stable wave-level ranking.  Each wave touches only its own counters; LDS
    // operations of one wave execute in program order, so iteration i+1
    // observes the leader updates of iteration i.  The signal fence keeps the
    // compiler from moving the LDS accesses across iterations (it emits no
    // instruction; a volatile pointer here would turn them into flat accesses).
    uint16_t* wh = sm.whist + wave * D;
#pragma unroll
    for (int i = 0; i < KPT; ++i) {
        const uint32_t word = digit_word<K>(keys[i], shift);
        const uint32_t d = __builtin_amdgcn_ubfe(word, shift & 31u, RB);
        const bool valid = full || (wofs + i * 64 < cnt);
        const uint64_t peers = match_digit<RB, FULL>(word, shift & 31u, FULL ? ~0ull : __ballot(valid));
        const uint32_t lt = mbcnt64(peers);
        const uint32_t before = wh[d];
        pos[i] = before + lt;
        // Materialise pos now: left to itself the compiler keeps before, lt and
        // the counter's address live per key until the staging loop (3
        // registers per key instead of 1, which halves the occupancy).
        asm volatile("" : "+v"(pos[i]));
        if (valid && lt == 0) wh[d] = (uint16_t)(before + (uint32_t)__popcll(peers));
        __atomic_signal_fence(__ATOMIC_SEQ_CST);
    }
    __syncthreads();

    // Per digit: exclusive prefix over waves, tile total, publish aggregate.
    uint32_t total_d = 0;
    if (tid < D) {
        uint32_t run = 0;
#pragma unroll
        for (int w = 0; w < WAVES; ++w) {
            const uint32_t c = sm.whist[w * D + tid];
            sm.whist[w * D + tid] = (uint16_t)run;
            run += c;
        }
        total_d = run;
        st_agent(&lb_cur[tile * LB_ROW + tid], (tile == first ? LB_INCLUSIVE : LB_AGGREGATE) | (total_d << 2));
    }
    uint32_t tile_total;
    const uint32_t start = block_exclusive_scan<BLOCK>(total_d, sm.tmp, &tile_total);
    if (tid < D) {
        // whist[w][d] becomes the tile-local position of wave w's first digit-d key.
#pragma unroll
        for (int w = 0; w < WAVES; ++w) sm.whist[w * D + tid] = (uint16_t)(sm.whist[w * D + tid] + start);
    }
    __syncthreads();

    // Stage keys in LDS in (digit, original order) order.  The digit is
    // recomputed from the key (opaque here, so the ranking loop's counter
    // addresses are not kept alive for this loop).
#pragma unroll
    for (int i = 0; i < KPT; ++i) {
        const bool valid = full || (wofs + i * 64 < cnt);
        if (valid) {
            asm volatile("" : "+v"(keys[i]));
            const uint32_t d = digit_of<K, RB>(keys[i], shift);
            const uint32_t lp = sm.whist[wave * D + d] + pos[i];
            s_keys[lp] = keys[i];
            pos[i] = lp;
        }
    }

    // Decoupled look-back, one thread per digit, down to the segment's first tile
    // (which always publishes inclusive counts, so the walk never passes it).
    // Four predecessors are fetched per round trip: agent-scope loads bypass
    // the XCD-local L2, so each round trip is long.  The wait is bounded: after
    // spin_limit empty polls the whole workgroup counts the digits of the
    // segment's earlier keys itself (exact, just slower), so progress never
    // depends on the order in which the hardware starts workgroups.
    if (tid == 0) sm.fallback = (spin_limit == 0u && tile != first) ? 1u : 0u;  // 0: always count directly
    uint32_t excl = 0;
    __syncthreads();
    if (tid < D && tile != first && spin_limit != 0u) {
        uint32_t k = tile - 1;
        uint32_t spins = 0;
        while (true) {
            uint32_t v[4];
#pragma unroll
            for (int j = 0; j < 4; ++j) v[j] = (k >= first + j) ? ld_agent(&lb_cur[(k - j) * LB_ROW + tid]) : 0u;
            bool found = false;
            uint32_t used = 0;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                if (found || used != (uint32_t)j) break;
                const uint32_t f = v[j] & LB_FLAG_MASK;
                if (f == LB_NOT_READY) break;
                excl += v[j] >> 2;
                if (f == LB_INCLUSIVE) found = true;
                else ++used;
            }
            if (found) break;
            k -= used;
            if (used < 4) {
                if (++spins > spin_limit) {
                    sm.fallback = 1u;
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
            }
        }
    }
    __syncthreads();
    if (sm.fallback) {
        for (uint32_t i = tid; i < D; i += BLOCK) sm.fb[i] = 0u;
        __syncthreads();
        for (uint32_t i = seg_base + tid; i < base; i += BLOCK) atomicAdd(&sm.fb[digit_of<K, RB>(sk[i], shift)], 1u);
        __syncthreads();
        if (tid < D) excl = sm.fb[tid];
    }
    if (tid < D) {
        if (tile != first) st_agent(&lb_cur[tile * LB_ROW + tid], LB_INCLUSIVE | ((excl + total_d) << 2));
        sm.gbase[tid] = prefix[tid] + add + excl - start;
    }
    __syncthreads();

    uint32_t dst[KPT];
#pragma unroll
    for (int i = 0; i < KPT; ++i) {
        const uint32_t j = tid + i * BLOCK;
        if (full || j < cnt) {
            const K key = s_keys[j];
            dst[i] = sm.gbase[digit_of<K, RB>(key, shift)] + j;
            dk[dst[i]] = key;
        }
    }
    if (PAIRS) {
        __syncthreads();
#pragma unroll
        for (int i = 0; i < KPT; ++i) {
            const bool valid = full || (wofs + i * 64 < cnt);
            if (valid) s_vals[pos[i]] = vals[i];
        }
        __syncthreads();
#pragma unroll
        for (int i = 0; i < KPT; ++i) {
            const uint32_t j = tid + i * BLOCK;
            if (full || j < cnt) dv[dst[i]] = s_vals[j];
        }
    }
}

// Full tiles take the variant without bounds tests (cnt is uniform per workgroup).
// STREAM: full tiles load the source non-temporally (ld_stream).  Measured:
// faster for plain key sorts (2^26 keys, about the MALL's size: 0.154 ->
// 0.145 ms per pass), slower for the segmented batch sort (564 MB of keys:
// 0.310 -> 0.336 ms per pass); key-value sorts showed no gain (indexed 2^26
// build 1.76-1.81 ms either way) and keep plain loads.
template <typename K, bool PAIRS, int BLOCK, int KPT, int RB, bool STREAM = false>
TC_DEV void onesweep_tile(OnesweepShared<K, BLOCK, KPT, RB>& sm, const K* __restrict__ sk, K* __restrict__ dk,
                          const uint32_t* __restrict__ sv, uint32_t* __restrict__ dv, uint32_t base,
                          uint32_t cnt, uint32_t shift, uint32_t tile, uint32_t first, uint32_t seg_base,
                          const uint32_t* __restrict__ prefix, uint32_t add, uint32_t* __restrict__ lb_cur,
                          uint32_t* __restrict__ lb_next, uint32_t spin_limit) {
    if (cnt == OnesweepShared<K, BLOCK, KPT, RB>::TILE)
        onesweep_tile_impl<K, PAIRS, BLOCK, KPT, RB, true, STREAM>(sm, sk, dk, sv, dv, base, cnt, shift, tile, first,
                                                           seg_base, prefix, add, lb_cur, lb_next, spin_limit);
    else
        onesweep_tile_impl<K, PAIRS, BLOCK, KPT, RB, false, STREAM>(sm, sk, dk, sv, dv, base, cnt, shift, tile, first,
                                                            seg_base, prefix, add, lb_cur, lb_next, spin_limit);
}
