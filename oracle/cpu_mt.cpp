// cpu_mt.cpp -- multi-threaded CPU build of the whole pipeline: the CPU
// baseline timed beside the GPU (test infrastructure, like the rest of this
// directory).  Same outputs as the single-threaded oracle functions, which it
// calls on index ranges where the stage is per element (Morton code, radix
// node, octree emission); the sort is a per-thread-histogram stable LSD radix
// sort and unique and the scan are two-pass (count per chunk, scan the
// counts, write).
#include <algorithm>
#include <chrono>
#include <cstring>
#include <thread>
#include <vector>

#include "tc_oracle.h"

namespace {

// fn(t, lo, hi) on `threads` contiguous chunks of [0, n).
template <typename F>
void par(int threads, size_t n, F fn) {
    if (threads <= 1 || n < 4096) {
        fn(0, (size_t)0, n);
        return;
    }
    std::vector<std::thread> pool;
    pool.reserve(threads);
    for (int t = 0; t < threads; ++t) {
        const size_t lo = n * t / threads, hi = n * (t + 1) / threads;
        pool.emplace_back([=] { fn(t, lo, hi); });
    }
    for (auto& th : pool) th.join();
}

double ms_since(std::chrono::steady_clock::time_point t0) {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

// Stable LSD radix sort, 8-bit digits, in place (an even pass count leaves
// the result in keys).
template <typename K>
void lsd_sort(K* keys, K* tmp, size_t n, int threads) {
    constexpr int DIG = (int)sizeof(K);
    std::vector<uint32_t> cnt((size_t)threads * 256);
    K* src = keys;
    K* dst = tmp;
    for (int p = 0; p < DIG; ++p) {
        const int sh = 8 * p;
        std::fill(cnt.begin(), cnt.end(), 0u);
        par(threads, n, [&](int t, size_t lo, size_t hi) {
            uint32_t* c = &cnt[(size_t)t * 256];
            for (size_t i = lo; i < hi; ++i) ++c[(src[i] >> sh) & 255u];
        });
        const int used = n < 4096 ? 1 : threads;
        uint32_t run = 0;
        for (int d = 0; d < 256; ++d)
            for (int t = 0; t < used; ++t) {
                const uint32_t c = cnt[(size_t)t * 256 + d];
                cnt[(size_t)t * 256 + d] = run;
                run += c;
            }
        par(threads, n, [&](int t, size_t lo, size_t hi) {
            uint32_t* o = &cnt[(size_t)t * 256];
            for (size_t i = lo; i < hi; ++i) dst[o[(src[i] >> sh) & 255u]++] = src[i];
        });
        std::swap(src, dst);
    }
}

// The same with a u32 value carried beside each key (the key-value sort of
// the sort-dominated workload, 2^27 u64 keys with u32 payloads).
template <typename K>
void lsd_sort_pairs(K* keys, uint32_t* vals, K* ktmp, uint32_t* vtmp, size_t n, int threads) {
    constexpr int DIG = (int)sizeof(K);
    std::vector<uint32_t> cnt((size_t)threads * 256);
    K* src = keys;
    K* dst = ktmp;
    uint32_t* vsrc = vals;
    uint32_t* vdst = vtmp;
    for (int p = 0; p < DIG; ++p) {
        const int sh = 8 * p;
        std::fill(cnt.begin(), cnt.end(), 0u);
        par(threads, n, [&](int t, size_t lo, size_t hi) {
            uint32_t* c = &cnt[(size_t)t * 256];
            for (size_t i = lo; i < hi; ++i) ++c[(src[i] >> sh) & 255u];
        });
        const int used = n < 4096 ? 1 : threads;
        uint32_t run = 0;
        for (int d = 0; d < 256; ++d)
            for (int t = 0; t < used; ++t) {
                const uint32_t c = cnt[(size_t)t * 256 + d];
                cnt[(size_t)t * 256 + d] = run;
                run += c;
            }
        par(threads, n, [&](int t, size_t lo, size_t hi) {
            uint32_t* o = &cnt[(size_t)t * 256];
            for (size_t i = lo; i < hi; ++i) {
                const uint32_t j = o[(src[i] >> sh) & 255u]++;
                dst[j] = src[i];
                vdst[j] = vsrc[i];
            }
        });
        std::swap(src, dst);
        std::swap(vsrc, vdst);
    }
}

template <typename K>
uint32_t par_unique(const K* s, K* out, size_t n, int threads) {
    if (n == 0) return 0;
    const int used = n < 4096 ? 1 : threads;
    std::vector<uint32_t> cnt(used + 1, 0);
    par(threads, n, [&](int t, size_t lo, size_t hi) {
        uint32_t c = 0;
        for (size_t i = lo; i < hi; ++i) c += (i == 0 || s[i] != s[i - 1]) ? 1u : 0u;
        cnt[t + 1] = c;
    });
    for (int t = 0; t < used; ++t) cnt[t + 1] += cnt[t];
    par(threads, n, [&](int t, size_t lo, size_t hi) {
        uint32_t o = cnt[t];
        for (size_t i = lo; i < hi; ++i)
            if (i == 0 || s[i] != s[i - 1]) out[o++] = s[i];
    });
    return cnt[used];
}

uint32_t par_exclusive_scan(const uint32_t* in, uint32_t* out, size_t n, int threads) {
    const int used = n < 4096 ? 1 : threads;
    std::vector<uint32_t> sums(used + 1, 0);
    par(threads, n, [&](int t, size_t lo, size_t hi) {
        uint32_t s = 0;
        for (size_t i = lo; i < hi; ++i) s += in[i];
        sums[t + 1] = s;
    });
    for (int t = 0; t < used; ++t) sums[t + 1] += sums[t];
    par(threads, n, [&](int t, size_t lo, size_t hi) {
        uint32_t run = sums[t];
        for (size_t i = lo; i < hi; ++i) {
            out[i] = run;
            run += in[i];
        }
    });
    out[n] = sums[used];
    return sums[used];
}

template <typename K>
uint32_t mt_build(const float* pts, uint32_t n, float min_coord, float range, int threads, K* codes, K* sorted,
                  K* tmp, K* uniq, uint32_t* n_unique_out, uint8_t* prefix_n, uint8_t* hll, uint8_t* hlr,
                  int32_t* left, int32_t* parent, uint32_t* ec, uint32_t* offsets, tco_oct_node* nodes,
                  uint64_t node_cap, int link_leaves, double* stage_ms) {
    constexpr bool W64 = sizeof(K) == 8;
    if (threads < 1) threads = 1;
    double st[7] = {0, 0, 0, 0, 0, 0, 0};
    auto t0 = std::chrono::steady_clock::now();
    // k_ComputeMortonCode, shader/morton.comp:60-67 (per point).
    par(threads, n, [&](int, size_t lo, size_t hi) {
        for (size_t i = lo; i < hi; ++i) {
            const float* p = pts + 4 * i;
            if constexpr (W64) codes[i] = tco_morton63_point(p[0], p[1], p[2], min_coord, range);
            else codes[i] = tco_morton30_point(p[0], p[1], p[2], min_coord, range);
        }
    });
    st[0] = ms_since(t0);
    t0 = std::chrono::steady_clock::now();
    std::memcpy(sorted, codes, (size_t)n * sizeof(K));
    lsd_sort<K>(sorted, tmp, n, threads);
    st[1] = ms_since(t0);
    t0 = std::chrono::steady_clock::now();
    const uint32_t u = par_unique<K>(sorted, uniq, n, threads);
    *n_unique_out = u;
    st[2] = ms_since(t0);
    const uint32_t nb = u > 0 ? u - 1 : 0;
    uint32_t total = 0;
    if (nb > 0) {
        t0 = std::chrono::steady_clock::now();
        parent[0] = 0;
        par(threads, nb, [&](int, size_t lo, size_t hi) {
            if constexpr (W64)
                tco_build_radix_tree_range_u64(uniq, u, (uint32_t)lo, (uint32_t)hi, prefix_n, hll, hlr, left, parent);
            else
                tco_build_radix_tree_range_u32(uniq, u, (uint32_t)lo, (uint32_t)hi, prefix_n, hll, hlr, left, parent);
        });
        st[3] = ms_since(t0);
        // countEdges, shader/edge_count.comp:30-41, root slot = 1.
        t0 = std::chrono::steady_clock::now();
        par(threads, nb, [&](int, size_t lo, size_t hi) {
            for (size_t i = lo; i < hi; ++i)
                ec[i] = i == 0 ? 1u : (uint32_t)(prefix_n[i] / 3 - prefix_n[parent[i]] / 3);
        });
        st[4] = ms_since(t0);
        t0 = std::chrono::steady_clock::now();
        total = par_exclusive_scan(ec, offsets, nb, threads);
        st[5] = ms_since(t0);
        if (total > node_cap) return 0xFFFFFFFFu;
        t0 = std::chrono::steady_clock::now();
        par(threads, total, [&](int, size_t lo, size_t hi) {
            std::memset(nodes + lo, 0, (hi - lo) * sizeof(tco_oct_node));
        });
        par(threads, nb, [&](int, size_t lo, size_t hi) {
            if constexpr (W64)
                tco_build_octree_range_u64(nodes, offsets, ec, uniq, prefix_n, parent, left, hll, hlr, (uint32_t)lo,
                                           (uint32_t)hi, min_coord, range, link_leaves);
            else
                tco_build_octree_range_u32(nodes, offsets, ec, uniq, prefix_n, parent, left, hll, hlr, (uint32_t)lo,
                                           (uint32_t)hi, min_coord, range, link_leaves);
        });
        st[6] = ms_since(t0);
    }
    if (stage_ms) std::memcpy(stage_ms, st, sizeof st);
    return total;
}

}  // namespace

extern "C" {

uint32_t tco_mt_build_tree_u32(const float* pts, uint32_t n, float min_coord, float range, int threads,
                               uint32_t* codes, uint32_t* sorted, uint32_t* tmp, uint32_t* uniq,
                               uint32_t* n_unique, uint8_t* prefix_n, uint8_t* hll, uint8_t* hlr, int32_t* left,
                               int32_t* parent, uint32_t* ec, uint32_t* offsets, tco_oct_node* nodes,
                               uint64_t node_cap, int link_leaves, double* stage_ms) {
    return mt_build<uint32_t>(pts, n, min_coord, range, threads, codes, sorted, tmp, uniq, n_unique, prefix_n, hll,
                              hlr, left, parent, ec, offsets, nodes, node_cap, link_leaves, stage_ms);
}

uint32_t tco_mt_build_tree_u64(const float* pts, uint32_t n, float min_coord, float range, int threads,
                               uint64_t* codes, uint64_t* sorted, uint64_t* tmp, uint64_t* uniq,
                               uint32_t* n_unique, uint8_t* prefix_n, uint8_t* hll, uint8_t* hlr, int32_t* left,
                               int32_t* parent, uint32_t* ec, uint32_t* offsets, tco_oct_node* nodes,
                               uint64_t node_cap, int link_leaves, double* stage_ms) {
    return mt_build<uint64_t>(pts, n, min_coord, range, threads, codes, sorted, tmp, uniq, n_unique, prefix_n, hll,
                              hlr, left, parent, ec, offsets, nodes, node_cap, link_leaves, stage_ms);
}

// std::sort of a copy of the codes (the reference's only CPU stage is a sort
// of the codes, src/main.rs:85); returns the milliseconds it took.
double tco_std_sort_ms(const uint32_t* codes, uint32_t n, uint32_t* work) {
    std::memcpy(work, codes, (size_t)n * 4);
    const auto t0 = std::chrono::steady_clock::now();
    std::sort(work, work + n);
    return ms_since(t0);
}

// Stable LSD radix sort of n u64 keys with u32 values on `threads` threads, in
// place (tmp buffers allocated here, outside the timing); returns its ms.
double tco_mt_sort_pairs_u64_ms(uint64_t* keys, uint32_t* vals, uint32_t n, int threads) {
    std::vector<uint64_t> kt(n);
    std::vector<uint32_t> vt(n);
    const auto t0 = std::chrono::steady_clock::now();
    lsd_sort_pairs<uint64_t>(keys, vals, kt.data(), vt.data(), n, threads < 1 ? 1 : threads);
    return ms_since(t0);
}

// std::stable_sort of n (u64 key, u32 value) pairs by key, written back in
// place; returns the milliseconds of the sort itself.
double tco_std_stable_sort_pairs_u64_ms(uint64_t* keys, uint32_t* vals, uint32_t n) {
    std::vector<std::pair<uint64_t, uint32_t>> kv(n);
    for (uint32_t i = 0; i < n; ++i) kv[i] = {keys[i], vals[i]};
    const auto t0 = std::chrono::steady_clock::now();
    std::stable_sort(kv.begin(), kv.end(), [](const auto& a, const auto& b) { return a.first < b.first; });
    const double ms = ms_since(t0);
    for (uint32_t i = 0; i < n; ++i) {
        keys[i] = kv[i].first;
        vals[i] = kv[i].second;
    }
    return ms;
}

}  // extern "C"
