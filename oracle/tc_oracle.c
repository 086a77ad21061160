/*
 * tc_oracle.c -- sequential CPU restatement of the reference's tree-construction
 * shaders.  TEST INFRASTRUCTURE ONLY (the checker; never part of the product).
 *
 * Reference paths are relative to vulkan/radixsort-vulkan/ in
 * ucsc-redwood/glsl-tree-construction.  Where the reference is undefined, racy
 * or self-inconsistent, the normalised rule is stated at the function and the
 * GPU implementation follows the same rule.
 *
 * Build with -ffp-contract=off: the quantise and decode steps must round each
 * fp32 operation separately, as the GPU path does.
 */
#include "tc_oracle.h"

#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------ Morton */

/* GLSL uint(float) truncates toward zero; out-of-range input is undefined in
 * GLSL (shader/morton.comp:37).  Normalised: NaN and negatives -> 0, values
 * >= 2^32 -> 0xFFFFFFFF, otherwise truncation. */
static uint32_t f2u_sat(float t) {
    if (!(t > 0.0f)) return 0u; /* NaN, -0, negatives */
    if (t >= 4294967296.0f) return 0xFFFFFFFFu;
    return (uint32_t)t;
}

/* single_point_to_code_v2, shader/morton.comp:50-56: ((v - min) / range) * scale,
 * each step a separately rounded fp32 operation. */
uint32_t tco_quantize(float v, float min_coord, float range, float bit_scale) {
    float t = v - min_coord;
    t = t / range;
    t = t * bit_scale;
    return f2u_sat(t);
}

/* morton3D_SplitBy3bits, shader/morton.comp:36-44 (10-bit input). */
static uint32_t split3_10(uint32_t b) {
    uint32_t x = b & 0x000003ffu;
    x = (x | x << 16) & 0x030000ffu;
    x = (x | x << 8) & 0x0300f00fu;
    x = (x | x << 4) & 0x030c30c3u;
    x = (x | x << 2) & 0x09249249u;
    return x;
}

/* 21-bit variant of the same spread for 63-bit codes (no reference counterpart;
 * the reference only has 30-bit codes, shader/morton.comp:51). */
static uint64_t split3_21(uint64_t b) {
    uint64_t x = b & 0x1fffffull;
    x = (x | x << 32) & 0x001f00000000ffffull;
    x = (x | x << 16) & 0x001f0000ff0000ffull;
    x = (x | x << 8) & 0x100f00f00f00f00full;
    x = (x | x << 4) & 0x10c30c30c30c30c3ull;
    x = (x | x << 2) & 0x1249249249249249ull;
    return x;
}

/* m3D_e_magicbits + single_point_to_code_v2, shader/morton.comp:46-56. */
uint32_t tco_morton30_point(float x, float y, float z, float min_coord, float range) {
    return split3_10(tco_quantize(x, min_coord, range, 1024.0f)) |
           (split3_10(tco_quantize(y, min_coord, range, 1024.0f)) << 1) |
           (split3_10(tco_quantize(z, min_coord, range, 1024.0f)) << 2);
}

uint64_t tco_morton63_point(float x, float y, float z, float min_coord, float range) {
    return split3_21(tco_quantize(x, min_coord, range, 2097152.0f)) |
           (split3_21(tco_quantize(y, min_coord, range, 2097152.0f)) << 1) |
           (split3_21(tco_quantize(z, min_coord, range, 2097152.0f)) << 2);
}

/* k_ComputeMortonCode, shader/morton.comp:60-67 (points are float[4], w ignored). */
void tco_morton30(const float* p, uint32_t* codes, size_t n, float min_coord, float range) {
    for (size_t i = 0; i < n; ++i)
        codes[i] = tco_morton30_point(p[4 * i], p[4 * i + 1], p[4 * i + 2], min_coord, range);
}

void tco_morton63(const float* p, uint64_t* codes, size_t n, float min_coord, float range) {
    for (size_t i = 0; i < n; ++i)
        codes[i] = tco_morton63_point(p[4 * i], p[4 * i + 1], p[4 * i + 2], min_coord, range);
}

/* morton3_d_get_third_bits, src/morton.rs:61-68 (GLSL twin shader/octree.comp:87-94). */
static uint32_t compact3_10(uint32_t m) {
    uint32_t x = m & 0x09249249u;
    x = (x ^ (x >> 2)) & 0x030c30c3u;
    x = (x ^ (x >> 4)) & 0x0300f00fu;
    x = (x ^ (x >> 8)) & 0x030000ffu;
    x = (x ^ (x >> 16)) & 0x000003ffu;
    return x;
}

static uint64_t compact3_21(uint64_t m) {
    uint64_t x = m & 0x1249249249249249ull;
    x = (x ^ (x >> 2)) & 0x10c30c30c30c30c3ull;
    x = (x ^ (x >> 4)) & 0x100f00f00f00f00full;
    x = (x ^ (x >> 8)) & 0x001f0000ff0000ffull;
    x = (x ^ (x >> 16)) & 0x001f00000000ffffull;
    x = (x ^ (x >> 32)) & 0x00000000001fffffull;
    return x;
}

/* morton32_to_xyz, src/morton.rs:43-59: ((float)axis / 1024) * range + min, w = 1. */
void tco_morton30_decode(uint32_t code, float min_coord, float range, float out[4]) {
    for (int a = 0; a < 3; ++a) {
        float f = (float)compact3_10(code >> a);
        f = f / 1024.0f;
        f = f * range;
        out[a] = f + min_coord;
    }
    out[3] = 1.0f;
}

void tco_morton63_decode(uint64_t code, float min_coord, float range, float out[4]) {
    for (int a = 0; a < 3; ++a) {
        float f = (float)compact3_21(code >> a);
        f = f / 2097152.0f;
        f = f * range;
        out[a] = f + min_coord;
    }
    out[3] = 1.0f;
}

/* --------------------------------------------------------------- Histogram */

/* k_GlobalHistogram, shader/histogram.comp:103-195: one read of the keys gives
 * every digit's histogram as an exclusive prefix, digit-major:
 *   hist[256*d + b] = #{ keys k : ((k >> 8d) & 255) < b }.
 * Normalised: the 0xF mask typo on the second digit of the .y key
 * (shader/histogram.comp:128) is not reproduced; every digit uses 8 bits. */
static void hist_to_prefix(uint32_t* h, int digits) {
    for (int d = 0; d < digits; ++d) {
        uint32_t run = 0;
        for (int b = 0; b < 256; ++b) {
            uint32_t c = h[256 * d + b];
            h[256 * d + b] = run;
            run += c;
        }
    }
}

void tco_global_histogram_u32(const uint32_t* keys, size_t n, uint32_t* hist) {
    memset(hist, 0, 4 * 256 * sizeof(uint32_t));
    for (size_t i = 0; i < n; ++i)
        for (int d = 0; d < 4; ++d) hist[256 * d + ((keys[i] >> (8 * d)) & 255u)]++;
    hist_to_prefix(hist, 4);
}

void tco_global_histogram_u64(const uint64_t* keys, size_t n, uint32_t* hist) {
    memset(hist, 0, 8 * 256 * sizeof(uint32_t));
    for (size_t i = 0; i < n; ++i)
        for (int d = 0; d < 8; ++d) hist[256 * d + (uint32_t)((keys[i] >> (8 * d)) & 255u)]++;
    hist_to_prefix(hist, 8);
}

/* ----------------------------------------------------------------- Binning */

/* One OneSweep digit-binning pass, shader/new_radix_sort.comp:116-314: key i
 * goes to globalHist[d] + (#keys of digit d before it), i.e. a stable
 * counting-sort scatter on digit d = (key >> radix_shift) & 255.  The
 * reference's tile decomposition, subgroup ranking and lookback only compute
 * that position; the position itself is this closed form. */
void tco_binning_pass_u32(const uint32_t* sk, uint32_t* dk, const uint32_t* sv, uint32_t* dv,
                          size_t n, const uint32_t* prefix, int shift) {
    uint32_t cur[256];
    memcpy(cur, prefix, sizeof(cur));
    for (size_t i = 0; i < n; ++i) {
        uint32_t d = (sk[i] >> shift) & 255u;
        uint32_t pos = cur[d]++;
        dk[pos] = sk[i];
        if (sv) dv[pos] = sv[i];
    }
}

void tco_binning_pass_u64(const uint64_t* sk, uint64_t* dk, const uint32_t* sv, uint32_t* dv,
                          size_t n, const uint32_t* prefix, int shift) {
    uint32_t cur[256];
    memcpy(cur, prefix, sizeof(cur));
    for (size_t i = 0; i < n; ++i) {
        uint32_t d = (uint32_t)((sk[i] >> shift) & 255u);
        uint32_t pos = cur[d]++;
        dk[pos] = sk[i];
        if (sv) dv[pos] = sv[i];
    }
}

/* Driver test_radix_sort, src/main.rs:185-800: histogram once, then one
 * binning pass per digit ping-ponging between the two buffers; the final
 * result ends in the caller's buffer (4 passes for u32, 8 for u64). */
void tco_sort_u32(uint32_t* keys, uint32_t* values, size_t n, uint32_t* tk, uint32_t* tv) {
    uint32_t hist[4 * 256];
    tco_global_histogram_u32(keys, n, hist);
    uint32_t *a = keys, *b = tk, *va = values, *vb = values ? tv : NULL;
    for (int p = 0; p < 4; ++p) {
        tco_binning_pass_u32(a, b, va, vb, n, hist + 256 * p, 8 * p);
        uint32_t* t = a; a = b; b = t;
        t = va; va = vb; vb = t;
    }
    /* 4 passes: the result is back in keys/values. */
}

void tco_sort_u64(uint64_t* keys, uint32_t* values, size_t n, uint64_t* tk, uint32_t* tv) {
    uint32_t hist[8 * 256];
    tco_global_histogram_u64(keys, n, hist);
    uint64_t *a = keys, *b = tk;
    uint32_t *va = values, *vb = values ? tv : NULL;
    for (int p = 0; p < 8; ++p) {
        tco_binning_pass_u64(a, b, va, vb, n, hist + 256 * p, 8 * p);
        uint64_t* t = a; a = b; b = t;
        uint32_t* tv2 = va; va = vb; vb = tv2;
    }
}

/* ------------------------------------------------------------------ Unique */

/* k_CountUnique, shader/unique.comp:31-45 (std::unique + count).  Normalised:
 * the result goes to a separate buffer instead of in place. */
uint32_t tco_unique_u32(const uint32_t* s, uint32_t* out, size_t n) {
    if (n == 0) return 0;
    size_t j = 0;
    out[0] = s[0];
    for (size_t i = 1; i < n; ++i)
        if (s[i] != out[j]) out[++j] = s[i];
    return (uint32_t)(j + 1);
}

uint32_t tco_unique_u64(const uint64_t* s, uint64_t* out, size_t n) {
    if (n == 0) return 0;
    size_t j = 0;
    out[0] = s[0];
    for (size_t i = 1; i < n; ++i)
        if (s[i] != out[j]) out[++j] = s[i];
    return (uint32_t)(j + 1);
}

/* Unique with run starts: the std::unique loop above, also recording the
 * position of each run's first key (first[j]) and first[U] = n.  No reference
 * counterpart (the reference keeps only the codes, src/unique.rs:278-283). */
uint32_t tco_unique_runs_u32(const uint32_t* s, uint32_t* out, uint32_t* first, size_t n) {
    size_t j = 0;
    if (n) {
        out[0] = s[0];
        first[0] = 0;
        for (size_t i = 1; i < n; ++i)
            if (s[i] != out[j]) {
                out[++j] = s[i];
                first[j] = (uint32_t)i;
            }
        ++j;
    }
    first[j] = (uint32_t)n;
    return (uint32_t)j;
}

uint32_t tco_unique_runs_u64(const uint64_t* s, uint64_t* out, uint32_t* first, size_t n) {
    size_t j = 0;
    if (n) {
        out[0] = s[0];
        first[0] = 0;
        for (size_t i = 1; i < n; ++i)
            if (s[i] != out[j]) {
                out[++j] = s[i];
                first[j] = (uint32_t)i;
            }
        ++j;
    }
    first[j] = (uint32_t)n;
    return (uint32_t)j;
}

/* ------------------------------------------------------------- Radix tree */

static int clz32_(uint32_t x) { return x ? __builtin_clz(x) : 32; }
static int clz64_(uint64_t x) { return x ? __builtin_clzll(x) : 64; }

/* delta_u32, shader/build_radix_tree.comp:48-58.  Normalised (reference Q9):
 * the common-prefix length is counted over the code width (30 or 63 bits),
 * equal codes give the full width, and an index outside [0, n_keys-1] gives -1
 * (the reference's explicit cut-offs at :97,101,122 play that role). */
static int delta30(const uint32_t* c, int64_t nk, int64_t i, int64_t j) {
    if (j < 0 || j > nk - 1) return -1;
    uint32_t x = c[i] ^ c[j];
    return x ? clz32_(x) - 2 : 30;
}

static int delta63(const uint64_t* c, int64_t nk, int64_t i, int64_t j) {
    if (j < 0 || j > nk - 1) return -1;
    uint64_t x = c[i] ^ c[j];
    return x ? clz64_(x) - 1 : 63;
}

/* log2_ceil_u32 / ceil_div_u32, shader/build_radix_tree.comp:44-68. */
static int log2_ceil(uint32_t x) {
    int fl = 31 - clz32_(x);
    return fl + (((uint32_t)1 << fl) < x ? 1 : 0);
}
static uint32_t ceil_div(uint32_t a, uint32_t b) { return (a + b - 1) / b; }

/* k_BuildRadixTree, shader/build_radix_tree.comp:71-146 (Karras 2012).
 * Normalised: direction d = sign(delta(i,i+1) - delta(i,i-1)) in signed int
 * (the uint8 subtraction at :79-82 cannot be negative, Q6); the root covers all
 * n_keys codes (Q7); parent[0] = 0, the value the reference's zero-initialised
 * host array keeps (src/main.rs:103). */
#define TCO_BRT_BODY(DELTA)                                                         \
    for (int64_t i = i0; i < i1; ++i) {                                             \
        int d = (DELTA(c, nk, i, i + 1) - DELTA(c, nk, i, i - 1)) >= 0 ? 1 : -1;    \
        int dmin = DELTA(c, nk, i, i - d);                                          \
        int64_t lmax = 2;                                                           \
        while (DELTA(c, nk, i, i + lmax * d) > dmin) lmax *= 2;                     \
        int64_t l = 0;                                                              \
        for (int64_t t = lmax / 2; t >= 1; t /= 2)                                  \
            if (DELTA(c, nk, i, i + (l + t) * d) > dmin) l += t;                    \
        int64_t j = i + l * d;                                                      \
        int dnode = DELTA(c, nk, i, j);                                             \
        prefix_n[i] = (uint8_t)dnode;                                               \
        int64_t s = 0;                                                              \
        uint32_t maxdiv = l <= 1 ? 1u : (uint32_t)1 << log2_ceil((uint32_t)l);      \
        for (uint32_t div = 2, t = ceil_div((uint32_t)l, 2); div <= maxdiv;         \
             div <<= 1, t = ceil_div((uint32_t)l, div))                             \
            if (DELTA(c, nk, i, i + (s + (int64_t)t) * d) > dnode) s += t;          \
        int64_t gamma = i + s * d + (d < 0 ? d : 0);                                \
        left_child[i] = (int32_t)gamma;                                             \
        hll[i] = (uint8_t)((i < j ? i : j) == gamma);                               \
        hlr[i] = (uint8_t)((i > j ? i : j) == gamma + 1);                           \
        if (!hll[i]) parent[gamma] = (int32_t)i;                                    \
        if (!hlr[i]) parent[gamma + 1] = (int32_t)i;                                \
    }

/* Nodes [i0, i1) only (every node is independent; each parent[] entry has
 * one writer), so index ranges can be built on separate threads. */
void tco_build_radix_tree_range_u32(const uint32_t* c, uint32_t n_keys, uint32_t i0, uint32_t i1,
                                    uint8_t* prefix_n, uint8_t* hll, uint8_t* hlr, int32_t* left_child,
                                    int32_t* parent) {
    int64_t nk = n_keys;
    /* d is never 0 for strictly increasing codes: the neighbours' prefix
     * lengths differ (Karras 2012, sec. 4).  Ties (duplicate codes) take +1. */
    TCO_BRT_BODY(delta30)
}

void tco_build_radix_tree_range_u64(const uint64_t* c, uint32_t n_keys, uint32_t i0, uint32_t i1,
                                    uint8_t* prefix_n, uint8_t* hll, uint8_t* hlr, int32_t* left_child,
                                    int32_t* parent) {
    int64_t nk = n_keys;
    TCO_BRT_BODY(delta63)
}

void tco_build_radix_tree_u32(const uint32_t* c, uint32_t n_keys, uint8_t* prefix_n, uint8_t* hll,
                              uint8_t* hlr, int32_t* left_child, int32_t* parent) {
    uint32_t nb = n_keys > 0 ? n_keys - 1 : 0;
    if (nb > 0) parent[0] = 0;
    tco_build_radix_tree_range_u32(c, n_keys, 0, nb, prefix_n, hll, hlr, left_child, parent);
}

void tco_build_radix_tree_u64(const uint64_t* c, uint32_t n_keys, uint8_t* prefix_n, uint8_t* hll,
                              uint8_t* hlr, int32_t* left_child, int32_t* parent) {
    uint32_t nb = n_keys > 0 ? n_keys - 1 : 0;
    if (nb > 0) parent[0] = 0;
    tco_build_radix_tree_range_u64(c, n_keys, 0, nb, prefix_n, hll, hlr, left_child, parent);
}

/* -------------------------------------------------------------- Edge count */

/* countEdges, shader/edge_count.comp:30-41: ec[i] = prefix_n[i]/3 -
 * prefix_n[parent[i]]/3.  Normalised (Q10/Q11): n_brt = n_keys-1 entries, and
 * the root's entry is 1, the slot that holds the octree root. */
void tco_edge_count(const uint8_t* prefix_n, const int32_t* parent, uint32_t* ec, uint32_t n_brt,
                    int code_bits) {
    (void)code_bits;
    for (uint32_t i = 0; i < n_brt; ++i)
        ec[i] = i == 0 ? 1u : (uint32_t)(prefix_n[i] / 3 - prefix_n[parent[i]] / 3);
}

/* ----------------------------------------------------------------- Scan */

/* k_PartialSum, shader/partial_sum.comp:23-39 with the caller's fix-up
 * src/main.rs:127-129.  Normalised (Q11): exclusive scan, out[n] = total. */
uint32_t tco_exclusive_scan_u32(const uint32_t* in, uint32_t* out, size_t n) {
    uint32_t run = 0;
    for (size_t i = 0; i < n; ++i) {
        out[i] = run;
        run += in[i];
    }
    out[n] = run;
    return run;
}

/* ---------------------------------------------------------------- Octree */

/* SetChild, shader/octree.comp:22-25.  The mask is or-ed atomically (Q15):
 * ranges of radix nodes may be emitted on separate threads, and several of
 * them set bits of one parent's mask; each children[] slot has one writer. */
static void set_child(tco_oct_node* node, int which, int32_t idx) {
    node->children[which] = idx;
    __atomic_fetch_or(&node->child_node_mask, 1 << which, __ATOMIC_RELAXED);
}

/* SetLeaf, shader/octree.comp:27-30, reproduced literally (Q16): the leaf
 * mask is cleared, so with zero-initialised nodes it stays 0. */
static void set_leaf(tco_oct_node* node, int which, int32_t idx) {
    node->children[which] = idx;
    __atomic_fetch_and(&node->child_leaf_mask, ~(1 << which), __ATOMIC_RELAXED);
}

/* Nearest radix-tree ancestor-or-self with a non-zero edge count.  With the
 * root's count normalised to 1 every walk ends at the root at the latest; it
 * replaces the counter > 30 cut-off of shader/octree.comp:157-165. */
static uint32_t contributing(uint32_t i, const uint32_t* ec, const int32_t* parent) {
    while (ec[i] == 0) i = (uint32_t)parent[i];
    return i;
}

/* k_MakeOctNodes + k_LinkLeafNodes, shader/octree.comp:126-222, and the host
 * root set-up src/main.rs:140-153.  Normalised:
 *  - the root is node 0 at level root_level = prefix_n[0]/3 (Q12: root_level is
 *    the root's, not the per-node value of :127), cell_size = range;
 *  - radix node i (i >= 1, ec[i] > 0) owns nodes offsets[i] .. offsets[i]+ec[i]-1,
 *    bottom (deepest, level prefix_n[i]/3) first; a node at level L has corner
 *    decode(prefix_L(code_i) << (bits - 3L)) and cell_size range / 2^(L - root_level);
 *  - child slots are set with an atomic-or mask (Q15), written here sequentially;
 *  - leaves are linked when link_leaves != 0 (the reference's disabled pass). */
#define TCO_OCT_BODY(BITS, DECODE)                                                            \
    int root_level = prefix_n[0] / 3;                                                         \
    for (uint32_t i = i0; i < i1; ++i) {                                                      \
        uint32_t cnt = ec[i];                                                                 \
        if (cnt == 0) continue;                                                               \
        int bottom = prefix_n[i] / 3;                                                         \
        for (uint32_t j = 0; j < cnt; ++j) {                                                  \
            int level = bottom - (int)j;                                                      \
            tco_oct_node* nd = &nodes[offsets[i] + j];                                        \
            DECODE((codes[i] >> ((BITS) - 3 * level)) << ((BITS) - 3 * level), min_coord, range, \
                   nd->corner);                                                               \
            nd->cell_size = range / (float)(1u << (level - root_level));                      \
            if (j + 1 < cnt) {                                                                \
                int which = (int)((codes[i] >> ((BITS) - 3 * level)) & 7u);                   \
                set_child(&nodes[offsets[i] + j + 1], which, (int32_t)(offsets[i] + j));      \
            }                                                                                 \
        }                                                                                     \
        if (i > 0) {                                                                          \
            int top_level = bottom - (int)cnt + 1;                                            \
            uint32_t a = contributing((uint32_t)parent[i], ec, parent);                       \
            int which = (int)((codes[i] >> ((BITS) - 3 * top_level)) & 7u);                   \
            set_child(&nodes[offsets[a]], which, (int32_t)(offsets[i] + cnt - 1));            \
        }                                                                                     \
    }                                                                                         \
    if (link_leaves) {                                                                        \
        for (uint32_t i = i0; i < i1; ++i) {                                                  \
            int leaf_level = prefix_n[i] / 3 + 1;                                             \
            for (int side = 0; side < 2; ++side) {                                            \
                if (!(side == 0 ? hll[i] : hlr[i])) continue;                                 \
                int32_t leaf = left_child[i] + side;                                          \
                int which = (int)((codes[leaf] >> ((BITS) - 3 * leaf_level)) & 7u);           \
                uint32_t a = contributing(i, ec, parent);                                     \
                set_leaf(&nodes[offsets[a]], which, leaf);                                    \
            }                                                                                 \
        }                                                                                     \
    }

/* Radix nodes [i0, i1) only, into zeroed nodes (the result does not depend on
 * how the range is split). */
void tco_build_octree_range_u32(tco_oct_node* nodes, const uint32_t* offsets, const uint32_t* ec,
                                const uint32_t* codes, const uint8_t* prefix_n, const int32_t* parent,
                                const int32_t* left_child, const uint8_t* hll, const uint8_t* hlr,
                                uint32_t i0, uint32_t i1, float min_coord, float range, int link_leaves) {
    TCO_OCT_BODY(30, tco_morton30_decode)
}

void tco_build_octree_range_u64(tco_oct_node* nodes, const uint32_t* offsets, const uint32_t* ec,
                                const uint64_t* codes, const uint8_t* prefix_n, const int32_t* parent,
                                const int32_t* left_child, const uint8_t* hll, const uint8_t* hlr,
                                uint32_t i0, uint32_t i1, float min_coord, float range, int link_leaves) {
    TCO_OCT_BODY(63, tco_morton63_decode)
}

uint32_t tco_build_octree_u32(tco_oct_node* nodes, const uint32_t* offsets, const uint32_t* ec,
                              const uint32_t* codes, const uint8_t* prefix_n, const int32_t* parent,
                              const int32_t* left_child, const uint8_t* hll, const uint8_t* hlr,
                              uint32_t n_brt, float min_coord, float range, int link_leaves) {
    if (n_brt == 0) return 0;
    uint32_t total = offsets[n_brt];
    memset(nodes, 0, (size_t)total * sizeof(tco_oct_node));
    tco_build_octree_range_u32(nodes, offsets, ec, codes, prefix_n, parent, left_child, hll, hlr, 0, n_brt,
                               min_coord, range, link_leaves);
    return total;
}

uint32_t tco_build_octree_u64(tco_oct_node* nodes, const uint32_t* offsets, const uint32_t* ec,
                              const uint64_t* codes, const uint8_t* prefix_n, const int32_t* parent,
                              const int32_t* left_child, const uint8_t* hll, const uint8_t* hlr,
                              uint32_t n_brt, float min_coord, float range, int link_leaves) {
    if (n_brt == 0) return 0;
    uint32_t total = offsets[n_brt];
    memset(nodes, 0, (size_t)total * sizeof(tco_oct_node));
    tco_build_octree_range_u64(nodes, offsets, ec, codes, prefix_n, parent, left_child, hll, hlr, 0, n_brt,
                               min_coord, range, link_leaves);
    return total;
}

uint64_t tco_octree_node_bound(uint64_t n_keys, int code_bits) {
    if (n_keys < 2) return 0;
    uint64_t total = 1, p = 1, half = n_keys / 2;
    int levels = (code_bits - 1) / 3;
    for (int L = 1; L <= levels; ++L) {
        p = p < (1ull << 60) ? p * 8 : p;
        total += p < half ? p : half;
    }
    return total;
}
