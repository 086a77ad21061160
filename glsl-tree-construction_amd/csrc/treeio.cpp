// treeio.cpp -- host-side tree validation and export (tc_tree_validate_host,
// tc_tree_write_host).  No reference counterpart: the reference never reads
// its octree back (src/octree.rs:404-430) and its only writer is the
// commented-out dump of the sorted keys (src/main.rs:761-768).
//
// The validator re-derives every property from the definitions, not from the
// construction: Karras radix tree over the unique codes, edge counts and
// offsets as the node count per radix node, octree cells as the distinct
// 3-bit-digit prefixes of the codes below the root level.
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/tc.h"

// ------------------------------------------- host decode (no GPU needed)
namespace {
uint32_t h_compact3_10(uint32_t m) {
    uint32_t x = m & 0x09249249u;
    x = (x ^ (x >> 2)) & 0x030c30c3u;
    x = (x ^ (x >> 4)) & 0x0300f00fu;
    x = (x ^ (x >> 8)) & 0x030000ffu;
    x = (x ^ (x >> 16)) & 0x000003ffu;
    return x;
}
uint64_t h_compact3_21(uint64_t m) {
    uint64_t x = m & 0x1249249249249249ull;
    x = (x ^ (x >> 2)) & 0x10c30c30c30c30c3ull;
    x = (x ^ (x >> 4)) & 0x100f00f00f00f00full;
    x = (x ^ (x >> 8)) & 0x001f0000ff0000ffull;
    x = (x ^ (x >> 16)) & 0x001f00000000ffffull;
    x = (x ^ (x >> 32)) & 0x00000000001fffffull;
    return x;
}
}  // namespace

// morton32_to_xyz, src/morton.rs:43-59 (volatile keeps each fp32 step rounded).
extern "C" void tc_morton32_to_xyz(float ret[4], uint32_t code, float min_coord, float range) {
    for (int a = 0; a < 3; ++a) {
        volatile float f = (float)h_compact3_10(code >> a);
        f = f / 1024.0f;
        f = f * range;
        ret[a] = f + min_coord;
    }
    ret[3] = 1.0f;
}

extern "C" void tc_morton64_to_xyz(float ret[4], uint64_t code, float min_coord, float range) {
    for (int a = 0; a < 3; ++a) {
        volatile float f = (float)h_compact3_21(code >> a);
        f = f / 2097152.0f;
        f = f * range;
        ret[a] = f + min_coord;
    }
    ret[3] = 1.0f;
}


namespace {

struct Fail {
    char* msg;
    size_t len;
    int operator()(const char* fmt, ...) const {
        if (msg && len) {
            va_list ap;
            va_start(ap, fmt);
            vsnprintf(msg, len, fmt, ap);
            va_end(ap);
        }
        return TC_ERR_BAD_TREE;
    }
};

template <typename K> int code_bits();
template <> int code_bits<uint32_t>() { return 30; }
template <> int code_bits<uint64_t>() { return 63; }

// Common prefix length of two codes over the code width (the normalised
// delta of shader/build_radix_tree.comp:48-58).
template <typename K>
int lcp(K a, K b) {
    const K x = a ^ b;
    if (!x) return code_bits<K>();
    if (sizeof(K) == 4) return __builtin_clz((uint32_t)x) - 2;
    return __builtin_clzll((uint64_t)x) - 1;
}

void decode(int bits, uint64_t code, float min_coord, float range, float out[4]) {
    if (bits == 30) tc_morton32_to_xyz(out, (uint32_t)code, min_coord, range);
    else tc_morton64_to_xyz(out, code, min_coord, range);
}

template <typename K>
int validate(const tc_tree_host* t, const Fail& fail) {
    const int B = code_bits<K>();
    const uint32_t U = t->n_unique;
    const K* u = static_cast<const K*>(t->unique);
    if (U > 0 && !u) return fail("unique codes missing");
    for (uint32_t j = 1; j < U; ++j)
        if (!(u[j - 1] < u[j])) return fail("unique[%u] is not above unique[%u]", j, j - 1);
    for (uint32_t j = 0; j < U; ++j)
        if ((uint64_t)u[j] >> B) return fail("unique[%u] has bits above the %d-bit code", j, B);

    // Point -> leaf map and sorted codes.
    if (t->leaf_first) {
        const uint32_t* f = t->leaf_first;
        if (f[0] != 0 || f[U] != t->n_points) return fail("leaf_first must start at 0 and end at n_points");
        for (uint32_t j = 0; j < U; ++j)
            if (f[j] >= f[j + 1]) return fail("leaf_first[%u] is not below leaf_first[%u]", j, j + 1);
        if (t->sorted) {
            const K* s = static_cast<const K*>(t->sorted);
            for (uint32_t j = 0; j < U; ++j)
                for (uint32_t k = f[j]; k < f[j + 1]; ++k)
                    if (s[k] != u[j]) return fail("sorted[%u] is not unique[%u] of its leaf", k, j);
        }
    } else if (t->sorted && U) {
        const K* s = static_cast<const K*>(t->sorted);
        uint32_t j = 0;
        if (s[0] != u[0]) return fail("sorted[0] differs from unique[0]");
        for (uint32_t k = 1; k < t->n_points; ++k) {
            if (s[k] < s[k - 1]) return fail("sorted[%u] is below sorted[%u]", k, k - 1);
            if (s[k] != s[k - 1] && (++j >= U || s[k] != u[j])) return fail("unique codes differ from the sorted runs");
        }
        if (j + 1 != U) return fail("unique has %u codes, the sorted codes %u runs", U, j + 1);
    }
    if (t->point_index) {
        std::vector<uint8_t> seen(t->n_points, 0);
        for (uint32_t k = 0; k < t->n_points; ++k) {
            const uint32_t p = t->point_index[k];
            if (p >= t->n_points || seen[p]) return fail("point_index is not a permutation (entry %u)", k);
            seen[p] = 1;
        }
    }

    if (U < 2) {
        if (t->n_nodes != 0) return fail("%u octree nodes for fewer than two unique codes", t->n_nodes);
        return TC_OK;
    }
    const uint32_t NB = U - 1;
    const uint8_t* pn = t->prefix_n;
    const uint8_t* hll = t->has_leaf_left;
    const uint8_t* hlr = t->has_leaf_right;
    const int32_t* lc = t->left_child;
    const int32_t* par = t->parent;
    if (!pn || !hll || !hlr || !lc || !par || !t->edge_count || !t->offsets || (t->n_nodes && !t->nodes))
        return fail("radix tree or octree arrays missing");

    // Radix tree links: children in range, every leaf and non-root internal
    // node referenced exactly once, parents consistent.
    std::vector<uint8_t> leaf_refs(U, 0), node_refs(NB, 0);
    for (uint32_t i = 0; i < NB; ++i) {
        if (hll[i] > 1 || hlr[i] > 1) return fail("has_leaf flags of node %u are not 0/1", i);
        const int64_t g = lc[i];
        if (g < 0 || g >= (int64_t)NB) return fail("left_child[%u] = %lld out of range", i, (long long)g);
        for (int side = 0; side < 2; ++side) {
            const uint32_t c = (uint32_t)g + side;
            if (side == 0 ? hll[i] : hlr[i]) {
                if (leaf_refs[c]++) return fail("leaf %u has two parents", c);
            } else {
                if (c >= NB || c == 0) return fail("node %u: internal child %u out of range", i, c);
                if (node_refs[c]++) return fail("node %u has two parents", c);
                if (par[c] != (int32_t)i) return fail("parent[%u] = %d, expected %u", c, par[c], i);
            }
        }
    }
    for (uint32_t c = 1; c < NB; ++c)
        if (!node_refs[c]) return fail("node %u is not a child of any node", c);
    for (uint32_t c = 0; c < U; ++c)
        if (!leaf_refs[c]) return fail("leaf %u is not a child of any node", c);

    // Key ranges top-down from the root [0, U-1]: node i covers [lo, hi] with i
    // at one end, the split is left_child, prefix_n is the range's common
    // prefix and the split falls at its first differing bit.
    std::vector<uint32_t> lo(NB), hi(NB);
    std::vector<uint32_t> stack{0};
    lo[0] = 0;
    hi[0] = U - 1;
    while (!stack.empty()) {
        const uint32_t i = stack.back();
        stack.pop_back();
        const uint32_t g = (uint32_t)lc[i];
        if (!(lo[i] <= g && g < hi[i])) return fail("node %u: split %u outside its range [%u, %u]", i, g, lo[i], hi[i]);
        if (lo[i] != i && hi[i] != i) return fail("node %u is not an end of its range [%u, %u]", i, lo[i], hi[i]);
        const int p = lcp<K>(u[lo[i]], u[hi[i]]);
        if (pn[i] != p) return fail("prefix_n[%u] = %u, the range's common prefix is %d", i, pn[i], p);
        if (lcp<K>(u[g], u[g + 1]) != p) return fail("node %u does not split at its first differing bit", i);
        if (!hll[i]) {
            lo[g] = lo[i];
            hi[g] = g;
            stack.push_back(g);
        } else if (g != lo[i]) {
            return fail("node %u: left leaf %u but range starts at %u", i, g, lo[i]);
        }
        if (!hlr[i]) {
            lo[g + 1] = g + 1;
            hi[g + 1] = hi[i];
            stack.push_back(g + 1);
        } else if (g + 1 != hi[i]) {
            return fail("node %u: right leaf %u but range ends at %u", i, g + 1, hi[i]);
        }
    }

    // Edge counts (root slot reserved: ec[0] = 1) and offsets.
    const uint32_t* ec = t->edge_count;
    const uint32_t* off = t->offsets;
    if (ec[0] != 1) return fail("edge_count[0] = %u, the root counts 1", ec[0]);
    for (uint32_t i = 1; i < NB; ++i) {
        const uint32_t want = pn[i] / 3 - pn[par[i]] / 3;
        if (ec[i] != want) return fail("edge_count[%u] = %u, expected %u", i, ec[i], want);
    }
    if (off[0] != 0) return fail("offsets[0] = %u", off[0]);
    for (uint32_t i = 0; i < NB; ++i)
        if (off[i + 1] != off[i] + ec[i]) return fail("offsets[%u] is not the exclusive scan of edge_count", i + 1);
    if (off[NB] != t->n_nodes) return fail("offsets end at %u but there are %u octree nodes", off[NB], t->n_nodes);

    // Octree: walk from the root; each node's cell prefix (level, digits)
    // fixes its corner and size exactly (the library's decode).
    const tc_oct_node* nd = t->nodes;
    const uint32_t NN = t->n_nodes;
    const int root_level = pn[0] / 3;
    struct Item {
        uint32_t node;
        int level;
        uint64_t prefix;
    };
    std::vector<uint8_t> reached(NN, 0);
    std::vector<Item> todo{{0, root_level, (uint64_t)u[0] >> (B - 3 * root_level)}};
    reached[0] = 1;
    uint32_t n_reached = 1;
    std::vector<uint8_t> leaf_seen(U, 0);
    while (!todo.empty()) {
        const Item it = todo.back();
        todo.pop_back();
        const tc_oct_node& o = nd[it.node];
        float c[4];
        decode(B, it.prefix << (B - 3 * it.level), t->min_coord, t->range, c);
        const float size = t->range / (float)(1u << (it.level - root_level));
        if (std::memcmp(c, o.corner, 12) != 0 || o.cell_size != size)
            return fail("octree node %u: corner or cell size differs from its cell (level %d)", it.node, it.level);
        if (o.child_node_mask & ~0xFF) return fail("octree node %u: child mask above 8 bits", it.node);
        for (int s = 0; s < 8; ++s) {
            const int32_t ch = o.children[s];
            if ((o.child_node_mask >> s) & 1) {
                if (ch <= 0 || (uint32_t)ch >= NN) return fail("octree node %u slot %d: child %d out of range", it.node, s, ch);
                if (reached[ch]) return fail("octree node %d is reached twice", ch);
                if (it.level + 1 > B / 3) return fail("octree node %u has children below the last level", it.node);
                reached[ch] = 1;
                ++n_reached;
                todo.push_back({(uint32_t)ch, it.level + 1, (it.prefix << 3) | (uint64_t)s});
            } else if (t->link_leaves && ch != 0) {
                // A leaf slot: the leaf's code has this cell's prefix and digit s.
                if (ch < 0 || (uint32_t)ch >= U) return fail("octree node %u slot %d: leaf %d out of range", it.node, s, ch);
                if (it.level + 1 > B / 3) return fail("octree node %u has a leaf slot below the last level", it.node);
                const int L = it.level + 1;
                if (((uint64_t)u[ch] >> (B - 3 * L)) != ((it.prefix << 3) | (uint64_t)s))
                    return fail("leaf %d is linked into a cell that does not contain it", ch);
                if (leaf_seen[ch]++) return fail("leaf %d is linked twice", ch);
            }
        }
        if (!t->link_leaves) {
            for (int s = 0; s < 8; ++s)
                if (!((o.child_node_mask >> s) & 1) && o.children[s] != 0)
                    return fail("octree node %u slot %d written without leaf linking", it.node, s);
        }
    }
    if (n_reached != NN) return fail("%u of %u octree nodes are reachable from the root", n_reached, NN);
    if (t->link_leaves) {
        // Every code descends from the root to a leaf slot holding it (leaf 0
        // is the slot value 0, which cannot be told from an empty slot above).
        for (uint32_t j = 0; j < U; ++j) {
            uint32_t node = 0;
            int level = root_level;
            for (;;) {
                if (level + 1 > B / 3) return fail("code %u descends below the last level", j);
                const int s = (int)(((uint64_t)u[j] >> (B - 3 * (level + 1))) & 7u);
                const tc_oct_node& o = nd[node];
                if ((o.child_node_mask >> s) & 1) {
                    node = (uint32_t)o.children[s];
                    ++level;
                    continue;
                }
                if ((uint32_t)o.children[s] != j) return fail("code %u is not in its leaf slot (node %u slot %d)", j, node, s);
                break;
            }
        }
    }
    return TC_OK;
}

// --------------------------------------------------------------- writer
struct Arr {
    const char* name;
    const char* dtype;
    const void* data;
    uint64_t count, elem;
};

}  // namespace

extern "C" int tc_tree_validate_host(const tc_tree_host* t, char* msg, size_t msg_len) {
    if (!t || (t->code_bits != 30 && t->code_bits != 63)) return TC_ERR_INVALID;
    if (msg && msg_len) msg[0] = '\0';
    const Fail fail{msg, msg_len};
    return t->code_bits == 30 ? validate<uint32_t>(t, fail) : validate<uint64_t>(t, fail);
}

extern "C" int tc_tree_write_host(const tc_tree_host* t, const char* path) {
    if (!t || !path || (t->code_bits != 30 && t->code_bits != 63)) return TC_ERR_INVALID;
    const uint64_t ks = t->code_bits == 30 ? 4 : 8;
    const char* kdt = t->code_bits == 30 ? "<u4" : "<u8";
    const uint64_t U = t->n_unique, NB = U > 1 ? U - 1 : 0, N = t->n_points;
    std::vector<Arr> arrs;
    if (t->sorted) arrs.push_back({"sorted", kdt, t->sorted, N, ks});
    arrs.push_back({"unique", kdt, t->unique, U, ks});
    arrs.push_back({"prefix_n", "|u1", t->prefix_n, NB, 1});
    arrs.push_back({"has_leaf_left", "|u1", t->has_leaf_left, NB, 1});
    arrs.push_back({"has_leaf_right", "|u1", t->has_leaf_right, NB, 1});
    arrs.push_back({"left_child", "<i4", t->left_child, NB, 4});
    arrs.push_back({"parent", "<i4", t->parent, NB, 4});
    arrs.push_back({"edge_count", "<u4", t->edge_count, NB, 4});
    arrs.push_back({"offsets", "<u4", t->offsets, NB + 1, 4});
    arrs.push_back({"nodes", "tc_oct_node", t->nodes, t->n_nodes, 64});
    if (t->point_index) arrs.push_back({"point_index", "<u4", t->point_index, N, 4});
    if (t->leaf_first) arrs.push_back({"leaf_first", "<u4", t->leaf_first, U + 1, 4});
    for (const Arr& a : arrs)
        if (a.count && !a.data) return TC_ERR_INVALID;

    // Header first with placeholder offsets of fixed width, so its length is
    // known before the offsets are.
    auto header = [&](const std::vector<uint64_t>& offs) {
        char buf[256];
        std::string h = "{\"format\": \"tc-tree\", \"version\": 1";
        snprintf(buf, sizeof buf,
                 ", \"code_bits\": %d, \"min_coord\": %.9g, \"range\": %.9g, \"link_leaves\": %d, "
                 "\"n_points\": %u, \"n_unique\": %u, \"n_nodes\": %u",
                 t->code_bits, (double)t->min_coord, (double)t->range, t->link_leaves ? 1 : 0, t->n_points,
                 t->n_unique, t->n_nodes);
        h += buf;
        h += ", \"node_record\": {\"bytes\": 64, \"fields\": [[\"children\", \"<i4\", 8], [\"corner\", \"<f4\", 4], "
             "[\"cell_size\", \"<f4\", 1], [\"child_node_mask\", \"<i4\", 1], [\"child_leaf_mask\", \"<i4\", 1], "
             "[\"pad\", \"<i4\", 1]]}, \"arrays\": [";
        for (size_t i = 0; i < arrs.size(); ++i) {
            snprintf(buf, sizeof buf, "%s{\"name\": \"%s\", \"dtype\": \"%s\", \"count\": %llu, \"offset\": %-20llu}",
                     i ? ", " : "", arrs[i].name, arrs[i].dtype, (unsigned long long)arrs[i].count,
                     (unsigned long long)offs[i]);
            h += buf;
        }
        h += "]}";
        return h;
    };
    std::vector<uint64_t> offs(arrs.size(), 0);
    const uint64_t hlen = header(offs).size();
    uint64_t pos = (16 + hlen + 63) / 64 * 64;
    for (size_t i = 0; i < arrs.size(); ++i) {
        offs[i] = pos;
        pos = (pos + arrs[i].count * arrs[i].elem + 63) / 64 * 64;
    }
    std::string h = header(offs);
    h.resize((16 + hlen + 63) / 64 * 64 - 16, ' ');
    FILE* f = std::fopen(path, "wb");
    if (!f) return TC_ERR_IO;
    static const unsigned char magic[8] = {'T', 'C', 'T', 'R', 'E', 'E', 0, 1};
    unsigned char len8[8];
    for (int b = 0; b < 8; ++b) len8[b] = (unsigned char)(hlen >> (8 * b));
    bool ok = std::fwrite(magic, 1, 8, f) == 8 && std::fwrite(len8, 1, 8, f) == 8 &&
              std::fwrite(h.data(), 1, h.size(), f) == h.size();
    static const char zeros[64] = {0};
    uint64_t at = 16 + h.size();
    for (size_t i = 0; i < arrs.size() && ok; ++i) {
        if (offs[i] > at) ok = std::fwrite(zeros, 1, offs[i] - at, f) == offs[i] - at;
        const uint64_t bytes = arrs[i].count * arrs[i].elem;
        if (ok && bytes) ok = std::fwrite(arrs[i].data, 1, bytes, f) == bytes;
        at = offs[i] + bytes;
    }
    if (std::fclose(f) != 0) ok = false;
    return ok ? TC_OK : TC_ERR_IO;
}
