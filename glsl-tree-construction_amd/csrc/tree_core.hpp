// tree_core.hpp -- per-node device functions of the radix tree and octree
// stages, shared by the plain kernels (tree.hip) and the segmented batch
// kernels (batch.hip).  Every function works on one cloud's arrays (pointers
// already offset to the cloud) with cloud-local indices; `node_base` is the
// global index of the cloud's first octree node, subtracted from stored links
// so a cloud's octree is the same whether built alone or inside a batch.
#pragma once

#include "tc_internal.hpp"

// Codes of one cloud, read through an optional LDS copy of c[w0, w0 + wlen)
// (the window around a workgroup's nodes, where nearly every search probe
// lands).  The window pointer is typed as LDS (address space 3) so that window
// reads compile to ds_read and the rest to global loads, not both to flat
// loads.  Indices are 32-bit: a cloud has at most 2^30 keys.
template <typename K>
using lds_ptr = const __attribute__((address_space(3))) K*;

template <typename K>
struct CodeView {
    const K* __restrict__ c;
    lds_ptr<K> win = nullptr;
    int32_t w0 = 0;
    uint32_t wlen = 0;
    TC_DEV K operator[](uint32_t j) const {
        const uint32_t r = j - (uint32_t)w0;
        if (r < wlen) return win[r];
        return c[j];
    }
};

// LDS view of a __shared__ array.
template <typename K>
TC_DEV lds_ptr<K> as_lds(const K* p) {
    return (lds_ptr<K>)p;
}

// The radix tree of one cloud (cloud-local indices) as the octree stage reads
// it, through an optional LDS copy of the window [w0, w0 + wlen): left child,
// code, octree node offset, and a packed meta word edge_count | has_leaf_left
// << 8 | has_leaf_right << 9 | prefix_n << 10 (edge counts are at most 21,
// prefix lengths at most 63).  Outside the window, local index x is read at
// gb + x of the arrays (gb: the cloud's first slot in concatenated arrays,
// else 0; the pointers stay uniform).
template <typename K>
struct TreeView {
    const int32_t* __restrict__ left;
    const uint8_t* __restrict__ hll;
    const uint8_t* __restrict__ hlr;
    const uint32_t* __restrict__ ec;
    const uint8_t* __restrict__ prefix_n;
    const K* __restrict__ codes;
    const uint32_t* __restrict__ offsets;
    lds_ptr<int32_t> wleft = nullptr;
    lds_ptr<uint16_t> wmeta = nullptr;
    lds_ptr<K> wcode = nullptr;
    lds_ptr<uint32_t> woff = nullptr;
    int32_t w0 = 0;
    uint32_t wlen = 0;
    uint32_t gb = 0;

    static TC_DEV uint16_t pack(uint32_t e, uint8_t l, uint8_t r, uint8_t pn) {
        return (uint16_t)((e < 255u ? e : 255u) | ((uint32_t)(l != 0) << 8) | ((uint32_t)(r != 0) << 9) |
                          ((uint32_t)(pn & 63u) << 10));
    }
    static TC_DEV uint32_t edge_count(uint32_t m) { return m & 255u; }
    static TC_DEV int prefix(uint32_t m) { return (int)(m >> 10); }
    TC_DEV int32_t left_child(uint32_t x) const {
        const uint32_t r = x - (uint32_t)w0;
        if (r < wlen) return wleft[r];
        return left[gb + x];
    }
    TC_DEV uint32_t meta(uint32_t x) const {
        const uint32_t r = x - (uint32_t)w0;
        if (r < wlen) return wmeta[r];
        return pack(ec[gb + x], hll[gb + x], hlr[gb + x], prefix_n[gb + x]);
    }
    TC_DEV K code(uint32_t x) const {
        const uint32_t r = x - (uint32_t)w0;
        if (r < wlen) return wcode[r];
        return codes[gb + x];
    }
    TC_DEV uint32_t offset(uint32_t x) const {
        const uint32_t r = x - (uint32_t)w0;
        if (r < wlen) return woff[r];
        return offsets[gb + x];
    }
};

// Window staging.  Thread t takes entries t, t + BLOCK, ... of the W-entry
// window starting at j0; all of its loads are issued into registers before the
// first LDS store, so a tile pays one memory round trip for its window (a
// load-then-store loop waits once per iteration).  Indexes are clamped into
// the valid range (which must not be empty) instead of branching around the
// loads; entries outside [lo, hi) then get `fill`.
// The lane index is opaque for the reason given at stage_links.
template <int BLOCK, int W, typename K>
TC_DEV void stage_codes(K* dst, const K* __restrict__ src, int64_t j0, int64_t lo, int64_t hi, K fill) {
    constexpr int PER = (W + BLOCK - 1) / BLOCK;
    K v[PER];
    uint32_t lane_t = threadIdx.x;
    asm volatile("" : "+v"(lane_t));
#pragma unroll
    for (int r = 0; r < PER; ++r) {
        const int64_t j = j0 + r * BLOCK + (int)lane_t;
        v[r] = src[j < lo ? lo : (j < hi ? j : hi - 1)];
    }
#pragma unroll
    for (int r = 0; r < PER; ++r) {
        const int64_t j = j0 + r * BLOCK + (int)lane_t;
        if (j < lo || j >= hi) v[r] = fill;
    }
#pragma unroll
    for (int r = 0; r < PER; ++r) {
        const int q = r * BLOCK + (int)lane_t;
        if (q < W) dst[q] = v[r];
    }
}

// Window [w0, w1) (0 < w1 - w0 <= W) of the tree arrays for TreeView.  The
// lane index is made opaque here: otherwise the compiler hoists the per-lane
// part of all the PER x 7 load addresses (64-bit each) out of the caller's
// tile loop, keeps them live across it and spills them.
template <int BLOCK, int W, typename K>
TC_DEV void stage_tree(int32_t* s_left, uint16_t* s_meta, K* s_code, uint32_t* s_off,
                       const int32_t* __restrict__ left, const uint32_t* __restrict__ ec,
                       const uint8_t* __restrict__ hll, const uint8_t* __restrict__ hlr,
                       const uint8_t* __restrict__ prefix_n, const K* __restrict__ codes,
                       const uint32_t* __restrict__ offsets, int64_t w0, int64_t w1) {
    constexpr int PER = (W + BLOCK - 1) / BLOCK;
    int32_t vl[PER];
    uint32_t ve[PER], vo[PER];
    uint8_t va[PER], vb[PER], vp[PER];
    K vc[PER];
    uint32_t lane_t = threadIdx.x;
    asm volatile("" : "+v"(lane_t));
    // 32-bit byte offsets from the uniform base pointers (indices are below
    // 2^30), so each load is one base register plus one offset register.
    auto at = [](const auto* base, uint32_t j) {
        using T = typename std::remove_cv<typename std::remove_pointer<decltype(base)>::type>::type;
        return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + j * (uint32_t)sizeof(T));
    };
#pragma unroll
    for (int r = 0; r < PER; ++r) {
        int64_t jj = w0 + r * BLOCK + (int)lane_t;
        const uint32_t j = (uint32_t)(jj < w1 ? jj : w1 - 1);
        vl[r] = at(left, j);
        ve[r] = at(ec, j);
        va[r] = at(hll, j);
        vb[r] = at(hlr, j);
        vp[r] = at(prefix_n, j);
        vc[r] = at(codes, j);
        vo[r] = at(offsets, j);
    }
#pragma unroll
    for (int r = 0; r < PER; ++r) {
        const int q = r * BLOCK + (int)lane_t;
        if (w0 + q < w1) {
            s_left[q] = vl[r];
            s_meta[q] = TreeView<K>::pack(ve[r], va[r], vb[r], vp[r]);
            s_code[q] = vc[r];
            s_off[q] = vo[r];
        }
    }
}

// Common-prefix length of codes[i] and codes[j] over the code width; -1 for
// j outside [0, nk) (j is taken modulo 2^32, so i - x for x > i is outside).
template <typename K, typename View = CodeView<K>>
struct Delta {
    View c;
    uint32_t nk;
    K ci;
    TC_DEV int operator()(uint32_t j) const {
        if (j >= nk) return -1;
        return CodeTraits<K>::delta_bits(ci ^ c[j]);
    }
};

TC_DEV int log2_ceil_u32(uint32_t x) {
    const int fl = 31 - __clz(x);
    return fl + (((uint32_t)1 << fl) < x ? 1 : 0);
}

// Outputs of one radix node: its prefix length, split, leaf flags and the edge
// counts of its internal children.
struct BrtResult {
    int dnode;
    uint32_t gamma;
    bool leaf_l, leaf_r;
    uint32_t ec_l, ec_r;
};

// The radix-tree arrays of one cloud (cloud-local indices); ec may be NULL.
struct BrtArrays {
    uint8_t* prefix_n;
    uint8_t* hll;
    uint8_t* hlr;
    int32_t* left_child;
    int32_t* parent;
    uint32_t* ec;
};

// Store node i: its own entries, the parent links of its internal children,
// and (with ec) their edge counts; node 0 also writes the root's parent slot
// (the reference's zero-initialised entry, src/main.rs:103) and count, 1.
// Node i's results (cloud-local indices and values) stored at slot base + index
// of the arrays (base: the cloud's first slot in concatenated arrays, else 0).
TC_DEV void brt_store(const BrtArrays& o, uint32_t i, const BrtResult& r, uint32_t base = 0) {
    o.prefix_n[base + i] = (uint8_t)r.dnode;
    o.left_child[base + i] = (int32_t)r.gamma;
    o.hll[base + i] = r.leaf_l;
    o.hlr[base + i] = r.leaf_r;
    if (!r.leaf_l) o.parent[base + r.gamma] = (int32_t)i;
    if (!r.leaf_r) o.parent[base + r.gamma + 1] = (int32_t)i;
    if (i == 0) o.parent[base] = 0;
    if (o.ec) {
        if (!r.leaf_l) o.ec[base + r.gamma] = r.ec_l;
        if (!r.leaf_r) o.ec[base + r.gamma + 1] = r.ec_r;
        if (i == 0) o.ec[base] = 1u;
    }
}

// k_BuildRadixTree, shader/build_radix_tree.comp:71-146 (Karras 2012), for
// internal node i of a cloud of nk strictly increasing codes.
//
// With want_ec the node also computes the edge counts of its internal children
// (countEdges, shader/edge_count.comp:36-38): a child covering [lo, hi] has
// prefix length lcp(codes[lo], codes[hi]), so its count
// prefix_n[child]/3 - prefix_n[i]/3 is known here without the child's output.
//
// With LIMIT > 0 the node gives up (returns false) when its range reaches
// LIMIT keys: every probe then stays within LIMIT of i, so a caller can serve
// them from an LDS window and handle the few long ranges (nodes near the root)
// separately, instead of letting one long search hold up a whole wave.
template <typename K, uint32_t LIMIT = 0, typename View = CodeView<K>>
TC_DEV bool brt_node(View codes, uint32_t nk, uint32_t i, bool want_ec, BrtResult* out) {
    const Delta<K, View> delta{codes, nk, codes[i]};
    // Direction (Karras 2012 eq. 1); ties (duplicate codes only) take +1.
    const bool fwd = delta(i + 1) - delta(i - 1) >= 0;
    // Index i + d*x, modulo 2^32 (x < 2^31, nk <= 2^30: wrapped values are >= nk).
    auto at = [&](uint32_t x) { return fwd ? i + x : i - x; };
    const int dmin = delta(fwd ? i - 1 : i + 1);
    uint32_t lmax = 2;
    while (lmax <= nk && delta(at(lmax)) > dmin) {
        lmax *= 2;
        if (LIMIT && lmax > LIMIT) return false;
    }
    uint32_t l = 0;
    for (uint32_t t = lmax / 2; t >= 1; t /= 2)
        if (delta(at(l + t)) > dmin) l += t;
    const uint32_t j = at(l);
    const int dnode = delta(j);
    // Split: binary search with ceil(l / 2^k) steps (shader/build_radix_tree.comp:119-129).
    uint32_t s = 0;
    const int kmax = l <= 1 ? 0 : log2_ceil_u32(l);
    for (int k = 1; k <= kmax; ++k) {
        const uint32_t t = (l + (1u << k) - 1) >> k;
        if (delta(at(s + t)) > dnode) s += t;
    }
    BrtResult r;
    r.dnode = dnode;
    r.gamma = fwd ? i + s : i - s - 1;
    const uint32_t lo = i < j ? i : j, hi = i < j ? j : i;
    r.leaf_l = lo == r.gamma;
    r.leaf_r = hi == r.gamma + 1;
    r.ec_l = r.ec_r = 0;
    if (want_ec) {
        const int level = dnode / 3;
        if (!r.leaf_l) {
            const Delta<K, View> dl{codes, nk, codes[lo]};
            r.ec_l = (uint32_t)(dl(r.gamma) / 3 - level);
        }
        if (!r.leaf_r) {
            const Delta<K, View> dr{codes, nk, codes[r.gamma + 1]};
            r.ec_r = (uint32_t)(dr(hi) / 3 - level);
        }
    }
    *out = r;
    return true;
}

// brt_node for one node by a group of G lanes (G divides 64; every lane of the
// group passes the same i and gets the same result).  Each round of a search
// probes G positions at once: the doubling step tests G powers of two, the two
// binary searches become (G + 1)-ary.  With strictly increasing codes
// delta(i, at(x)) does not increase with x, so every search predicate is true
// up to a boundary and false after it, and each search returns the boundary
// the one-lane search finds (the largest x whose probe is true).  A node whose
// range spans 2^22 keys waits for 3 + 7 + 7 rounds of probes instead of
// 22 + 22 + 22, which is what the nodes near the root of a cluster spend.
// Lanes per long-range node in the radix-tree kernels.
constexpr int BRT_LONG_GROUP = 8;

template <typename K, int G, typename View = CodeView<K>>
TC_DEV void brt_node_group(View codes, uint32_t nk, uint32_t i, bool want_ec, BrtResult* out, uint32_t e_start = 1) {
    static_assert(G >= 2 && G <= 32 && (64 % G) == 0, "groups tile the wave");
    const uint32_t lane = lane_id();
    const uint32_t gl = lane % G;         // lane within the group
    const uint32_t gshift = lane - gl;    // the group's first lane
    const uint32_t gmask = (1u << G) - 1u;
    auto group_bits = [&](bool p) { return (uint32_t)(__ballot(p) >> gshift) & gmask; };
    const Delta<K, View> delta{codes, nk, codes[i]};
    const bool fwd = delta(i + 1) - delta(i - 1) >= 0;
    auto at = [&](uint32_t x) { return fwd ? i + x : i - x; };
    const int dmin = delta(fwd ? i - 1 : i + 1);
    // lmax = 2^e for the first e >= 1 with delta(at(2^e)) <= dmin (e <= 31:
    // 2^31 > nk, outside the keys).  A caller that has already seen the probes
    // up to 2^(e_start - 1) hold starts there (one doubling round less for the
    // ranges that round would have spent below the window's bound).
    uint32_t e0 = e_start, efail;
    while (true) {
        const uint32_t e = e0 + gl;
        const uint32_t m = group_bits(e <= 31 && delta(at(1u << e)) > dmin);
        if (m != gmask) {
            efail = e0 + (uint32_t)__builtin_ctz(~m);
            break;
        }
        e0 += G;
    }
    // Largest x in [lo, hi) whose probe is true, given it is true at lo and
    // false at hi.
    auto search = [&](uint32_t lo, uint32_t hi, int thr) {
        while (hi - lo > 1) {
            const uint32_t step = (hi - lo + G) / (G + 1);
            const uint32_t x = lo + (gl + 1) * step;
            const uint32_t m = group_bits(x < hi && delta(at(x)) > thr);
            lo += (uint32_t)__builtin_ctz(~m) * step;  // the probes true below the first false one
            hi = lo + step < hi ? lo + step : hi;
        }
        return lo;
    };
    // l lies in [2^(efail - 1), 2^efail) when that lower probe is known to
    // hold: tested here (efail > e_start) or by the caller (e_start > 1).  From
    // e_start = 1 with efail = 1 the search starts at 0, as the probe at 1 can
    // fail on equal codes.
    const uint32_t lo0 = (efail > e_start || e_start > 1) ? 1u << (efail - 1) : 0u;
    const uint32_t l = search(lo0, 1u << efail, dmin);
    const uint32_t j = at(l);
    const int dnode = delta(j);
    const uint32_t s = search(0, l, dnode);
    BrtResult r;
    r.dnode = dnode;
    r.gamma = fwd ? i + s : i - s - 1;
    const uint32_t lo = i < j ? i : j, hi = i < j ? j : i;
    r.leaf_l = lo == r.gamma;
    r.leaf_r = hi == r.gamma + 1;
    r.ec_l = r.ec_r = 0;
    if (want_ec) {
        const int level = dnode / 3;
        if (!r.leaf_l) {
            const Delta<K, View> dl{codes, nk, codes[lo]};
            r.ec_l = (uint32_t)(dl(r.gamma) / 3 - level);
        }
        if (!r.leaf_r) {
            const Delta<K, View> dr{codes, nk, codes[r.gamma + 1]};
            r.ec_r = (uint32_t)(dr(hi) / 3 - level);
        }
    }
    *out = r;
}

// NPL radix-tree nodes per lane (brt_node) for ranges below LIMIT keys (a
// power of two), every probe in an LDS window: w[j - w0] = codes[j] for
// |j - idx[n]| < 2 LIMIT, with slots outside the cloud's keys (and slot -1)
// holding an all-ones sentinel, whose XOR with any code is at least 2^B for
// codes of B bits, i.e. "no common prefix".  Each probe "delta(j) > t" is one
// XOR and one compare against a threshold: for codes of B bits in a word of W
// bits, clz(x) - (W - B) > t exactly when x < 2^(B - 1 - t).  The searches
// have fixed lengths so the NPL nodes run
// interleaved, NPL independent LDS reads per step: the doubling step probes
// 2, 4, ..., LIMIT at once (the first false one is where the one-by-one
// doubling stops), and each binary search takes log2(LIMIT) bounded steps
// (the largest x below the bound whose probe is true, as the halving and the
// ceil(l / 2^k) steps of brt_node find).  A node's chain is 1 + 2 log2(LIMIT)
// dependent reads, and the NPL chains overlap.  Returns bit n set for node n
// resolved here; the others reach LIMIT keys (or are not nodes: the caller
// masks them) and take the full search.
template <typename K, uint32_t LIMIT, int NPL, bool UNBOUNDED = false>
TC_DEV uint32_t brt_nodes_near(lds_ptr<K> w, int32_t w0, const uint32_t (&idx)[NPL], bool want_ec,
                               BrtResult (&res)[NPL]) {
    static_assert(LIMIT >= 2 && (LIMIT & (LIMIT - 1)) == 0, "LIMIT is a power of two");
    constexpr int B = CodeTraits<K>::BITS;
    constexpr int LOG = __builtin_ctz(LIMIT);
    auto code = [&](uint32_t j) { return w[(int32_t)j - w0]; };
    auto dlt = [&](K a, K b) {  // -1 for "outside" (sentinel)
        const int d = CodeTraits<K>::delta_bits(a ^ b);
        return d > -1 ? d : -1;
    };
    auto thr = [&](int t) { return t >= B ? (K)0 : (K)((K)1 << (B - 1 - t)); };  // delta > t  <=>  x < thr(t)
    K ci[NPL], tmin[NPL], tnode[NPL];
    bool fwd[NPL];
    uint32_t lmax[NPL], l[NPL], s[NPL], j[NPL];
    int dnode[NPL];
    uint32_t resolved = 0;
    // UNBOUNDED: the searches run without bound tests.  delta(i, i + d x) does not increase
    // with x (distinct codes; window slots outside the cloud hold the
    // sentinel), so every probe at or past the range end fails for the range
    // search, and at or past l for the split search (delta there is at most
    // dnode); every probe stays below LIMIT keys from the node, inside the
    // window, and the results of unresolved nodes are discarded.  A probe's
    // address is then one multiply-add from the node's slot and a signed byte
    // step.  The plain kernel takes this form for u32 codes (uniform 2^26
    // codes: 0.694 -> 0.679 ms, at 85 VGPRs, six waves per SIMD instead of
    // eight); with u64 codes it held 124 VGPRs, four waves per SIMD, and
    // slowed (2^24 63-bit codes 0.211 -> 0.227 ms), and the segmented kernel,
    // held to 64 VGPRs, would spill 30 of them.
    using lds_bytes = const __attribute__((address_space(3))) char*;
    lds_bytes base[NPL];
    int32_t step[NPL];
#pragma unroll
    for (int n = 0; n < NPL; ++n) {
        const uint32_t i = idx[n];
        ci[n] = code(i);
        const int dl = dlt(ci[n], code(i - 1));  // slot -1 holds the sentinel
        const int dr = dlt(ci[n], code(i + 1));
        fwd[n] = dr >= dl;
        tmin[n] = thr(fwd[n] ? dl : dr);
        base[n] = (lds_bytes)(w + ((int32_t)i - w0));
        step[n] = fwd[n] ? (int32_t)sizeof(K) : -(int32_t)sizeof(K);
    }
    auto at = [&](int n, uint32_t x) -> K {
        if constexpr (UNBOUNDED) return *(lds_ptr<K>)(base[n] + __mul24(step[n], (int32_t)x));
        else return code(fwd[n] ? idx[n] + x : idx[n] - x);
    };
    auto probe = [&](int n, uint32_t x, K t) { return (K)(ci[n] ^ at(n, x)) < t; };
#pragma unroll
    for (int n = 0; n < NPL; ++n) {
        int e = LOG + 1;
#pragma unroll
        for (int k = LOG; k >= 1; --k)
            if (!probe(n, 1u << k, tmin[n])) e = k;
        resolved |= (e <= LOG ? 1u : 0u) << n;
        lmax[n] = e <= LOG ? 1u << e : 1u;  // unresolved nodes search nothing more here
        l[n] = 0;
    }
    // Branch-free steps (bounded form: a probe past the bound reads slot 0 of
    // the search, the node itself, and is discarded), so the NPL reads of a
    // step issue together instead of each waiting behind a branch.
#pragma unroll
    for (uint32_t t = LIMIT / 2; t >= 1; t /= 2)
#pragma unroll
        for (int n = 0; n < NPL; ++n) {
            const uint32_t x = l[n] + t;
            const bool in = UNBOUNDED || x < lmax[n];
            const bool p = probe(n, in ? x : 0u, tmin[n]);
            l[n] = (in && p) ? x : l[n];
        }
#pragma unroll
    for (int n = 0; n < NPL; ++n) {
        j[n] = fwd[n] ? idx[n] + l[n] : idx[n] - l[n];
        dnode[n] = dlt(ci[n], at(n, l[n]));
        tnode[n] = thr(dnode[n]);
        s[n] = 0;
    }
#pragma unroll
    for (uint32_t t = LIMIT / 2; t >= 1; t /= 2)
#pragma unroll
        for (int n = 0; n < NPL; ++n) {
            const uint32_t x = s[n] + t;
            const bool in = UNBOUNDED || x < l[n];
            const bool p = probe(n, in ? x : 0u, tnode[n]);
            s[n] = (in && p) ? x : s[n];
        }
#pragma unroll
    for (int n = 0; n < NPL; ++n) {
        const uint32_t i = idx[n];
        BrtResult r;
        r.dnode = dnode[n];
        r.gamma = fwd[n] ? i + s[n] : i - s[n] - 1;
        const uint32_t lo = i < j[n] ? i : j[n], hi = i < j[n] ? j[n] : i;
        r.leaf_l = lo == r.gamma;
        r.leaf_r = hi == r.gamma + 1;
        r.ec_l = r.ec_r = 0;
        if (want_ec) {
            // All four reads unconditionally (in the window), then selects.
            const int level = dnode[n] / 3;
            const int el = dlt(code(lo), code(r.gamma)) / 3 - level;
            const int er = dlt(code(r.gamma + 1), code(hi)) / 3 - level;
            r.ec_l = r.leaf_l ? 0u : (uint32_t)el;
            r.ec_r = r.leaf_r ? 0u : (uint32_t)er;
        }
        res[n] = r;
    }
    return resolved;
}

template <typename K>
TC_DEV uint32_t child_slot(K code, int level) {
    constexpr int BITS = CodeTraits<K>::BITS;
    return (uint32_t)((code >> (BITS - 3 * level)) & 7u);
}

// One radix node's share of the octree (k_MakeOctNodes + k_LinkLeafNodes,
// shader/octree.comp:126-222), written without atomics: radix node a with
// ec[a] > 0 owns the chain offsets[a] .. offsets[a]+ec[a]-1 (bottom, i.e.
// deepest, first).  Chain node j > 0 has one child, node j-1.  The bottom node's
// children are found by walking down the radix tree from a through nodes with
// ec == 0 (they share a's octree level L): each internal node met with ec > 0
// contributes its chain top, each radix leaf met contributes itself (SetLeaf:
// children[slot] = leaf, leaf mask left as cleared, i.e. 0).  Every child of the
// bottom node sits one level below it, so its slot is the next 3 bits of its
// code.  Prefix lengths of same-level nodes lie in [3L, 3L+2] and grow by at
// least one per radix level, so the walk has at most three levels (a, two
// nodes, four nodes) and at most eight results; it is unrolled on those
// positions (position = path bits), which keeps every index in a register.
// The walk replaces the upward ancestor search of shader/octree.comp:157-165
// and the racy mask updates (each node is written once, whole, by one lane).
// offsets[] hold global node indices; stored links are local (minus node_base).
// Indices a, c are cloud-local; arrays are read at gb + index (lv.gb, see
// TreeView).
template <typename K>
TC_DEV void octree_build_node(tc_oct_node* __restrict__ nodes, uint32_t node_cap, uint32_t node_base,
                              const TreeView<K>& lv, uint32_t a, uint32_t nb, int root_level, int link_leaves,
                              float min_coord, float range, uint32_t write_base = 0) {
    const uint32_t ma = lv.meta(a);
    const uint32_t cnt = TreeView<K>::edge_count(ma);
    if (cnt == 0) return;
    const uint32_t off = lv.offset(a);
    const K code = lv.code(a);
    const int bottom = TreeView<K>::prefix(ma) / 3;

    // Walk (LDS reads only while inside the window).  Candidate p's kind is
    // two bits of `kinds`: 0 none, 1 radix leaf, 2 internal node with its own
    // chain (one register instead of eight keeps the kernel unspilled at
    // eight workgroups per CU).
    uint32_t cand[8];
    uint32_t kinds = 0;
    uint32_t mid[2];
    bool mid_ok[2] = {false, false};
    uint32_t low[4];
    bool low_ok[4] = {false, false, false, false};
#pragma unroll
    for (int p = 0; p < 8; ++p) cand[p] = 0;
    auto visit = [&](uint32_t x, int pos, int stride, uint32_t* next, bool* next_ok) {
        const uint32_t lc = (uint32_t)lv.left_child(x);
        const uint32_t m = lv.meta(x);
#pragma unroll
        for (int side = 0; side < 2; ++side) {
            const uint32_t c = lc + side;
            const int ps = pos + side * stride;
            if (c > nb) continue;  // a valid tree's children are in [0, nb]
            if ((m >> (8 + side)) & 1u) {
                cand[ps] = c;
                kinds |= (link_leaves ? 1u : 0u) << (2 * ps);
            } else if (c < nb) {
                const uint32_t e = TreeView<K>::edge_count(lv.meta(c));
                if (e > 0 || !next) {
                    cand[ps] = c;
                    kinds |= (e > 0 ? 2u : 0u) << (2 * ps);
                } else {
                    next[side] = c;
                    next_ok[side] = true;
                }
            }
        }
    };
    visit(a, 0, 4, mid, mid_ok);
#pragma unroll
    for (int k = 0; k < 2; ++k)
        if (mid_ok[k]) visit(mid[k], 4 * k, 2, low + 2 * k, low_ok + 2 * k);
#pragma unroll
    for (int k = 0; k < 4; ++k)
        if (low_ok[k]) visit(low[k], 2 * k, 1, nullptr, nullptr);

    // Children of the bottom node: slot from the child's code, link = leaf
    // index or the child's chain top.
    // Two groups of four candidates: the compiler barrier between them keeps
    // at most four candidates' code and offset loads in flight, which bounds
    // the kernel's registers (81 VGPRs with all eight hoisted together).
    int32_t bch[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    int32_t bmask = 0;
#pragma unroll
    for (int g = 0; g < 2; ++g) {
#pragma unroll
        for (int p = 4 * g; p < 4 * g + 4; ++p) {
            const uint32_t kind = (kinds >> (2 * p)) & 3u;
            if (kind == 0) continue;
            const uint32_t c = cand[p];
            const uint32_t slot = child_slot<K>(lv.code(c), bottom + 1);
            int32_t v = (int32_t)c;
            if (kind == 2) {
                v = (int32_t)(lv.offset(c) + TreeView<K>::edge_count(lv.meta(c)) - 1 - node_base);
                bmask |= 1 << slot;
            }
#pragma unroll
            for (int q = 0; q < 8; ++q) bch[q] = (slot == (uint32_t)q) ? v : bch[q];
        }
        asm volatile("" ::: "memory");
    }

    // Node idx (global) is stored at nodes[idx - write_base] (nodes may be one
    // cloud's own buffer, of node_cap entries).  The bottom node (j = 0) takes
    // the gathered children; every chain node above has one child, the node
    // one level deeper, whose slot and link are formed at the store.  A chain
    // node's corner is the decode of the code cut to its level; cutting the
    // code to its top 3 L bits keeps the top L bits of each axis, so the axes
    // are decoded once and cut per node (the same integers, the same floats).
    constexpr int AB = CodeTraits<K>::AXIS_BITS;
    uint32_t q[3];
#pragma unroll
    for (int ax = 0; ax < 3; ++ax) q[ax] = CodeTraits<K>::axis(code, ax);
    for (uint32_t j = 0; j < cnt; ++j) {
        const uint32_t idx = off + j;
        if (idx - write_base >= node_cap) break;
        const int level = bottom - (int)j;
        const int cut = AB - level;  // 0 <= cut <= AB
        float c4[4];
#pragma unroll
        for (int ax = 0; ax < 3; ++ax) {
            const uint32_t qa = (q[ax] >> cut) << cut;
            c4[ax] = CodeTraits<K>::corner(qa, min_coord, range);
        }
        c4[3] = 1.0f;
        // range / 2^(level - root_level), exact: a power-of-two scale.
        const float cell = ldexpf(range, root_level - level);
        int4* dst = reinterpret_cast<int4*>(nodes + (idx - write_base));
        if (j == 0) {
            dst[0] = make_int4(bch[0], bch[1], bch[2], bch[3]);
            dst[1] = make_int4(bch[4], bch[5], bch[6], bch[7]);
            dst[3] = make_int4(__float_as_int(cell), bmask, 0, 0);
        } else {
            const uint32_t slot = child_slot<K>(code, level + 1);
            const int32_t v = (int32_t)(idx - 1 - node_base);
            dst[0] = make_int4(slot == 0 ? v : 0, slot == 1 ? v : 0, slot == 2 ? v : 0, slot == 3 ? v : 0);
            dst[1] = make_int4(slot == 4 ? v : 0, slot == 5 ? v : 0, slot == 6 ? v : 0, slot == 7 ? v : 0);
            dst[3] = make_int4(__float_as_int(cell), 1 << slot, 0, 0);
        }
        dst[2] = make_int4(__float_as_int(c4[0]), __float_as_int(c4[1]), __float_as_int(c4[2]), __float_as_int(c4[3]));
    }
}
