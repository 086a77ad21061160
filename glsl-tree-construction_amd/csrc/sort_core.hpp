// sort_core.hpp -- the OneSweep tile step shared by the plain sort (sort.hip)
// and the segmented batch sort (batch.hip).
//
// One tile = BLOCK*KPT keys.  Wave w owns keys [w*64*KPT, (w+1)*64*KPT) of the
// tile, lane-striped (key i*64 + lane), so a wave's processing order (i, lane)
// is the keys' index order and the rank below is stable.
//
// Digit width RB (1..11 bits) is a template parameter: a sort of KB-bit keys
// runs ceil(KB / RB_max) passes whose widths differ by at most one bit (for
// 30-bit Morton codes: three 10-bit passes, or 8, 8, 7, 7).  Wider digits mean
// fewer passes over HBM but more ballots per key and longer look-back rows;
// the autotuner picks the width with the launch shape.
#pragma once

#include "tc_internal.hpp"

// Look-back row stride (u32 words per tile) of an RB-bit pass: one word per
// digit, at least 256 (the segmented batch sort keeps 256-word rows for its
// 8-bit and narrower passes).
constexpr uint32_t LB_ROW = 256;
template <int RB>
constexpr uint32_t lb_row() {
    return (1u << RB) < LB_ROW ? LB_ROW : (1u << RB);
}

// Live bits of the top digit (shift 8*(bytes-1)) of a Morton code: 30 - 24
// for 30-bit codes, 63 - 56 for 63-bit codes (the segmented batch sort's
// narrow last pass).
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
    // All RB bit extracts and ballots first, then the bitop3 chain: each ballot
    // gets its own SGPR pair, so no v_bitop3 waits on a compare that just wrote
    // VCC (the per-bit form serialised every bit through VCC with wait states;
    // measured 0.1399 -> 0.1375 ms per pass at 2^26 with the early counter read
    // of the ranking loop).
    uint32_t rep[RB];
    uint64_t m[RB];
#pragma unroll
    for (int b = 0; b < RB; ++b) {
        rep[b] = (uint32_t)__builtin_amdgcn_sbfe((int32_t)word, off + b, 1);
        // Opaque to the optimiser, which would otherwise test the bit again
        // with a shift + compare instead of comparing rep (no instruction).
        asm("" : "+v"(rep[b]));
    }
#pragma unroll
    for (int b = 0; b < RB; ++b) m[b] = __ballot(rep[b] != 0u);
    // (rep ^ m) | diff in one v_bitop3 per half (truth table 0xde, operands a=rep, b=diff, c=m).
    uint32_t diff_lo = 0u, diff_hi = 0u;
#pragma unroll
    for (int b = 0; b < RB - 1; ++b) {
        diff_lo = __builtin_amdgcn_bitop3_b32(rep[b], diff_lo, (uint32_t)m[b], 0xde);
        diff_hi = __builtin_amdgcn_bitop3_b32(rep[b], diff_hi, (uint32_t)(m[b] >> 32), 0xde);
    }
    if constexpr (ALL_LIVE) {
        // Every lane live: the last bit's step also complements, ~((rep ^ m) | diff) (table 0x21).
        const uint32_t lo = __builtin_amdgcn_bitop3_b32(rep[RB - 1], diff_lo, (uint32_t)m[RB - 1], 0x21);
        const uint32_t hi = __builtin_amdgcn_bitop3_b32(rep[RB - 1], diff_hi, (uint32_t)(m[RB - 1] >> 32), 0x21);
        return ((uint64_t)hi << 32) | lo;
    } else {
        diff_lo = __builtin_amdgcn_bitop3_b32(rep[RB - 1], diff_lo, (uint32_t)m[RB - 1], 0xde);
        diff_hi = __builtin_amdgcn_bitop3_b32(rep[RB - 1], diff_hi, (uint32_t)(m[RB - 1] >> 32), 0xde);
        return live & ~(((uint64_t)diff_hi << 32) | diff_lo);
    }
}

template <typename K, int RB = 8>
TC_DEV uint32_t digit_of(K k, uint32_t shift) {
    return (uint32_t)((k >> shift) & (K)((1u << RB) - 1u));
}

// A 32-bit word holding the RB-bit digit at `shift` at bit offset *off
// (uniform): u32 keys are their own word; u64 keys are shifted once, since a
// digit may straddle the two halves.
template <typename K>
TC_DEV uint32_t digit_word(K k, uint32_t shift, uint32_t* off) {
    if constexpr (sizeof(K) == 4) {
        *off = shift;
        return (uint32_t)k;
    } else {
        *off = 0u;
        return (uint32_t)(k >> shift);
    }
}

template <typename K, int BLOCK, int KPT, int RB = 8>
struct OnesweepShared {
    static constexpr int WAVES = BLOCK / 64;
    static constexpr uint32_t TILE = BLOCK * KPT;
    static constexpr int D = 1 << RB;
    // Digits owned per thread in the per-digit steps (thread t: digits
    // [t*DPT, t*DPT + DPT)); two are read and published as one 64-bit word.
    static constexpr int DPT = D > BLOCK ? D / BLOCK : 1;
    static constexpr size_t KEY_BYTES = (size_t)TILE * sizeof(K);
    static constexpr size_t WH_BYTES = (size_t)WAVES * D * sizeof(uint16_t);
    // Per-wave digit counters during the ranking (then tile-local positions,
    // < TILE <= 65535, so 16 bits), then the keys staged in digit order, then
    // the values: one region, since their lifetimes do not overlap.
    alignas(16) unsigned char stage[KEY_BYTES > WH_BYTES ? KEY_BYTES : WH_BYTES];
    uint32_t gbase[D];  // digit bases of the scatter; the fall-back's digit counts before that
    uint32_t tmp[WAVES + 2];
    uint32_t aux;
    uint32_t fallback;  // look-back: the stalled predecessor + 1 (0: none)
};

// One digit-binning tile step (shader/new_radix_sort.comp:116-314), digit
// d = (key >> shift) & (2^RB - 1):
//   keys sk[base, base+cnt) (cnt <= TILE) go to
//   dk[prefix[d] + add + (#digit-d keys of earlier tiles of this segment) + rank].
// The caller guarantees that the key bits above shift + RB are equal within
// the segment (zero for Morton codes, or the bits of later passes), which the
// LSD order makes irrelevant here.
// `tile` indexes the look-back rows (lb_row<RB>() u32 words per tile) of
// lb_cur; the look-back walks down to `first`, the first tile of the segment,
// which publishes its counts as inclusive.  lb_next (may be NULL) is the next
// pass's look-back row of this tile, cleared here (next_row words).
// seg_base is the index of the segment's first key (the fall-back counts the
// keys in [seg_base, base)); spin_limit bounds the look-back wait.
template <typename K, bool PAIRS, int BLOCK, int KPT, int RB, bool FULL, bool STREAM>
TC_DEV void onesweep_tile_impl(OnesweepShared<K, BLOCK, KPT, RB>& sm, const K* __restrict__ sk, K* __restrict__ dk,
                               const uint32_t* __restrict__ sv, uint32_t* __restrict__ dv, uint32_t base,
                               uint32_t cnt, uint32_t shift, uint32_t tile, uint32_t first, uint32_t seg_base,
                               const uint32_t* __restrict__ prefix, uint32_t add, uint32_t* __restrict__ lb_cur,
                               uint32_t* __restrict__ lb_next, uint32_t next_row, uint32_t spin_limit) {
    using Sh = OnesweepShared<K, BLOCK, KPT, RB>;
    constexpr int WAVES = Sh::WAVES;
    constexpr uint32_t WAVE_TILE = 64 * KPT;
    constexpr uint32_t D = 1u << RB;
    constexpr int DPT = Sh::DPT;
    constexpr uint32_t ROW = lb_row<RB>();
    static_assert(RB >= 1 && RB <= 11, "digit width");
    static_assert(DPT == 1 || DPT == 2, "at most two digits per thread (one 64-bit look-back word)");
    static_assert(BLOCK * KPT <= 65535, "tile positions must fit the 16-bit counters");
    uint16_t* whist = reinterpret_cast<uint16_t*>(sm.stage);
    K* s_keys = reinterpret_cast<K*>(sm.stage);
    uint32_t* s_vals = reinterpret_cast<uint32_t*>(sm.stage);  // reused once the keys are out

    const uint32_t tid = threadIdx.x;
    const uint32_t lane = tid & 63u;
    const uint32_t wave = tid >> 6;
    constexpr bool full = FULL;  // full tiles (cnt == TILE) skip every bounds test
    const uint32_t wofs = wave * WAVE_TILE + lane;  // tile-relative index of key 0 of this lane
    const bool owner = tid * DPT < D;              // owns digits d0 .. d0 + DPT - 1
    const uint32_t d0 = tid * DPT;

    for (uint32_t i = tid; i < WAVES * D; i += BLOCK) whist[i] = 0u;

    K keys[KPT];
    uint32_t vals[KPT];
    uint32_t pos[KPT];
    if constexpr (FULL) {
        // One 64-bit base per lane: the KPT loads then differ by immediate
        // offsets (indexing with base + wofs + i*64 in 32 bits cost two VALU
        // instructions per key for the address).
        const K* src = sk + base + wofs;
#pragma unroll
        for (int i = 0; i < KPT; ++i) keys[i] = STREAM ? ld_stream(src + i * 64) : src[i * 64];
        if (PAIRS) {
            const uint32_t* vsrc = sv + base + wofs;
#pragma unroll
            for (int i = 0; i < KPT; ++i) vals[i] = STREAM ? ld_stream(vsrc + i * 64) : vsrc[i * 64];
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
        for (uint32_t i = tid; i < next_row; i += BLOCK) lb_next[tile * next_row + i] = 0u;
    }
    __syncthreads();  // counters cleared

    // Stable wave-level ranking.  Each wave touches only its own counters; LDS
    // operations of one wave execute in program order, so iteration i+1
    // observes the leader updates of iteration i.  The signal fence keeps the
    // compiler from moving the LDS accesses across iterations (it emits no
    // instruction; a volatile pointer here would turn them into flat accesses).
    uint16_t* wh = whist + wave * D;
#pragma unroll
    for (int i = 0; i < KPT; ++i) {
        uint32_t off;
        const uint32_t word = digit_word<K>(keys[i], shift, &off);
        const uint32_t d = __builtin_amdgcn_ubfe(word, off & 31u, RB);
        const bool valid = full || (wofs + i * 64 < cnt);
        // The wave's counter is read before the ballots and kept there by a
        // scheduling barrier, so its LDS latency overlaps the match (it was
        // issued after seven of the eight ballots, its latency exposed per key).
        const uint32_t before = wh[d];
        __builtin_amdgcn_sched_barrier(0);
        const uint64_t peers = match_digit<RB, FULL>(word, off & 31u, FULL ? ~0ull : __ballot(valid));
        const uint32_t lt = mbcnt64(peers);
        pos[i] = before + lt;
        // Materialise pos now: left to itself the compiler keeps before, lt and
        // the counter's address live per key until the staging loop (3
        // registers per key instead of 1, which halves the occupancy).
        asm volatile("" : "+v"(pos[i]));
        if (valid && lt == 0) wh[d] = (uint16_t)(before + (uint32_t)__popcll(peers));
        __atomic_signal_fence(__ATOMIC_SEQ_CST);
    }
    __syncthreads();

    // Per owned digit: exclusive prefix over waves, tile total, publish the
    // aggregate (one 32- or 64-bit word per thread).
    uint32_t tot[DPT];
    uint32_t own_sum = 0;
#pragma unroll
    for (int j = 0; j < DPT; ++j) {
        tot[j] = 0;
        if (owner) {
            uint32_t run = 0;
#pragma unroll
            for (int w = 0; w < WAVES; ++w) {
                const uint32_t c = whist[w * D + d0 + j];
                whist[w * D + d0 + j] = (uint16_t)run;
                run += c;
            }
            tot[j] = run;
            own_sum += run;
        }
    }
    const uint32_t flag0 = tile == first ? LB_INCLUSIVE : LB_AGGREGATE;
    if (owner) {
        if constexpr (DPT == 1) {
            st_agent(&lb_cur[tile * ROW + d0], flag0 | (tot[0] << 2));
        } else {
            const uint64_t w = ((uint64_t)(flag0 | (tot[1] << 2)) << 32) | (flag0 | (tot[0] << 2));
            st_agent64(reinterpret_cast<uint64_t*>(lb_cur + tile * ROW + d0), w);
        }
    }
    uint32_t tile_total;
    const uint32_t own_start = block_exclusive_scan<BLOCK>(own_sum, sm.tmp, &tile_total);
    uint32_t start[DPT];
#pragma unroll
    for (int j = 0; j < DPT; ++j) start[j] = own_start + (j > 0 ? tot[0] : 0u);
    if (owner) {
        // whist[w][d] becomes the tile-local position of wave w's first digit-d key.
#pragma unroll
        for (int j = 0; j < DPT; ++j) {
#pragma unroll
            for (int w = 0; w < WAVES; ++w)
                whist[w * D + d0 + j] = (uint16_t)(whist[w * D + d0 + j] + start[j]);
        }
    }
    __syncthreads();

    // Tile-local position of every key; the digit is recomputed from the key
    // (opaque here, so the ranking loop's counter addresses are not kept alive
    // for this loop).  The counters share LDS with the staging area, so all
    // positions are read before any key is staged.
#pragma unroll
    for (int i = 0; i < KPT; ++i) {
        const bool valid = full || (wofs + i * 64 < cnt);
        if (valid) {
            asm volatile("" : "+v"(keys[i]));
            pos[i] += whist[wave * D + digit_of<K, RB>(keys[i], shift)];
        }
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < KPT; ++i) {
        const bool valid = full || (wofs + i * 64 < cnt);
        if (valid) s_keys[pos[i]] = keys[i];
    }

    // Decoupled look-back, one thread per digit (or per pair of digits, read
    // as one 64-bit word: both halves are always published together, so they
    // carry the same flag), down to the segment's first tile (which always
    // publishes inclusive counts, so the walk never passes it).  Four
    // predecessors are fetched per round trip: agent-scope loads bypass the
    // XCD-local L2, so each round trip is long.
    //
    // The wait is bounded: a walker that has polled spin_limit times without
    // progress stops at the predecessor that has not published; the workgroup
    // then counts that tile's digits itself (it is a full tile: only a
    // segment's last tile is partial, and it comes after this one) and the
    // walkers stopped there step below it and go on.  Tiles are workgroup ids,
    // and the hardware may start a workgroup before a lower-numbered one (each
    // XCD starts its own share in order, but an XCD held by another kernel's
    // workgroups starts nothing): a predecessor that has not started costs one
    // tile read here, never a wait on that other kernel.  spin_limit 0 counts
    // every predecessor that way (used to test it).
    // sm.fallback: the stalled predecessor + 1 (0: none).
    uint32_t excl[DPT];
#pragma unroll
    for (int j = 0; j < DPT; ++j) excl[j] = 0;
    bool walking = owner && tile != first;
    uint32_t k = tile - 1;  // the next predecessor to account for
    if (tid == 0) sm.fallback = 0u;
    __syncthreads();
    while (true) {
        if (walking && spin_limit == 0u) {
            atomicMax(&sm.fallback, k + 1u);
        } else if (walking) {
            uint32_t spins = 0;
            while (true) {
                uint64_t v[4];
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    if (k >= first + q) {
                        if constexpr (DPT == 1) v[q] = ld_agent(&lb_cur[(k - q) * ROW + d0]);
                        else v[q] = ld_agent64(reinterpret_cast<const uint64_t*>(lb_cur + (k - q) * ROW + d0));
                    } else {
                        v[q] = 0;
                    }
                }
                bool found = false;
                uint32_t used = 0;
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    if (found || used != (uint32_t)q) break;
                    const uint32_t lo = (uint32_t)v[q];
                    const uint32_t f = lo & LB_FLAG_MASK;
                    if (f == LB_NOT_READY) break;
                    excl[0] += lo >> 2;
                    if constexpr (DPT == 2) excl[1] += (uint32_t)(v[q] >> 32) >> 2;
                    if (f == LB_INCLUSIVE) found = true;
                    else ++used;
                }
                if (found) {
                    walking = false;
                    break;
                }
                k -= used;
                if (used < 4) {
                    if (++spins > spin_limit) {
                        atomicMax(&sm.fallback, k + 1u);
                        break;
                    }
                    __builtin_amdgcn_s_sleep(1);
                }
            }
        }
        __syncthreads();
        const uint32_t stalled = sm.fallback;
        if (stalled == 0u) break;  // uniform: read by every thread after the barrier
        // Digit counts of predecessor stalled - 1, into gbase (free until the
        // scatter bases are written below).
        uint32_t* fb = sm.gbase;
        for (uint32_t i = tid; i < D; i += BLOCK) fb[i] = 0u;
        __syncthreads();
        if (tid == 0) sm.fallback = 0u;  // every thread has read it
        const K* tk = sk + seg_base + (size_t)(stalled - 1u - first) * Sh::TILE;
        for (uint32_t i = tid; i < Sh::TILE; i += BLOCK) atomicAdd(&fb[digit_of<K, RB>(tk[i], shift)], 1u);
        __syncthreads();
        if (walking && k + 1u == stalled) {
#pragma unroll
            for (int j = 0; j < DPT; ++j) excl[j] += fb[d0 + j];
            if (k == first) walking = false;
            else --k;
        }
        // The next round's barrier orders these reads before fb is cleared
        // again; the final round's orders them before gbase is written.
    }
    if (owner) {
        if (tile != first) {
            if constexpr (DPT == 1) {
                st_agent(&lb_cur[tile * ROW + d0], LB_INCLUSIVE | ((excl[0] + tot[0]) << 2));
            } else {
                const uint64_t w = ((uint64_t)(LB_INCLUSIVE | ((excl[1] + tot[1]) << 2)) << 32) |
                                   (LB_INCLUSIVE | ((excl[0] + tot[0]) << 2));
                st_agent64(reinterpret_cast<uint64_t*>(lb_cur + tile * ROW + d0), w);
            }
        }
#pragma unroll
        for (int j = 0; j < DPT; ++j) sm.gbase[d0 + j] = prefix[d0 + j] + add + excl[j] - start[j];
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
                          uint32_t* __restrict__ lb_next, uint32_t next_row, uint32_t spin_limit) {
    if (cnt == OnesweepShared<K, BLOCK, KPT, RB>::TILE)
        onesweep_tile_impl<K, PAIRS, BLOCK, KPT, RB, true, STREAM>(sm, sk, dk, sv, dv, base, cnt, shift, tile, first,
                                                                   seg_base, prefix, add, lb_cur, lb_next, next_row,
                                                                   spin_limit);
    else
        onesweep_tile_impl<K, PAIRS, BLOCK, KPT, RB, false, STREAM>(sm, sk, dk, sv, dv, base, cnt, shift, tile, first,
                                                                    seg_base, prefix, add, lb_cur, lb_next, next_row,
                                                                    spin_limit);
}
