// gen.cpp -- seeded point-cloud generators (host only).
//
// The reference draws its inputs from an unseeded thread_rng
// (init_random, src/init.rs:3-23: uniform [min, min+range) per axis, w = 1), so
// its inputs cannot be reproduced.  These generators are counter based: the
// values of point i depend only on (seed, i), computed as SplitMix64 outputs
// (Steele, Lea & Flood, "Fast splittable pseudorandom number generators",
// 2014) of a per-seed stream, so any thread count yields the same arrays.
#include <cmath>
#include <cstdint>
#include <thread>
#include <vector>

#include "../../include/tc.h"

namespace {

constexpr uint64_t GOLDEN = 0x9E3779B97F4A7C15ull;

inline uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// k-th output of the SplitMix64 stream that starts at `state`.
inline uint64_t stream_at(uint64_t state, uint64_t k) { return mix64(state + (k + 1) * GOLDEN); }

// Exact uniform float in [0, 1): 24 random bits times 2^-24.
inline float unit24(uint64_t r) { return (float)(r >> 40) * (1.0f / 16777216.0f); }

// Uniform double in (0, 1] and [0, 1) from 53 bits.
inline double open01(uint64_t r) { return ((double)(r >> 11) + 1.0) * (1.0 / 9007199254740992.0); }
inline double half01(uint64_t r) { return (double)(r >> 11) * (1.0 / 9007199254740992.0); }

template <typename F>
void parallel_for(uint64_t n, F&& body) {
    unsigned hw = std::thread::hardware_concurrency();
    unsigned nt = hw == 0 ? 1 : (hw > 16 ? 16 : hw);
    if (n < (1u << 16)) nt = 1;
    if (nt == 1) {
        body(0, n);
        return;
    }
    std::vector<std::thread> ts;
    const uint64_t chunk = (n + nt - 1) / nt;
    for (unsigned t = 0; t < nt; ++t) {
        const uint64_t lo = t * chunk, hi = lo + chunk < n ? lo + chunk : n;
        if (lo >= hi) break;
        ts.emplace_back([&body, lo, hi] { body(lo, hi); });
    }
    for (auto& th : ts) th.join();
}

}  // namespace

// Uniform xyz in [0, extent)^3 (exact for power-of-two extents), w = 1.
extern "C" void tc_gen_uniform(float* pts, uint64_t n, uint64_t seed, float extent) {
    const uint64_t base = mix64(seed ^ 0x5EEDu) * 4;
    parallel_for(n, [&](uint64_t lo, uint64_t hi) {
        for (uint64_t i = lo; i < hi; ++i) {
            const uint64_t st = base + i * 4 * GOLDEN;
            pts[4 * i + 0] = unit24(stream_at(st, 0)) * extent;
            pts[4 * i + 1] = unit24(stream_at(st, 1)) * extent;
            pts[4 * i + 2] = unit24(stream_at(st, 2)) * extent;
            pts[4 * i + 3] = 1.0f;
        }
    });
}

// Gaussian mixture: n_centres centres uniform in [0, extent)^3; each point picks
// a centre uniformly and adds N(0, sigma^2) per axis (Box-Muller), clipped to
// [0, extent) (the largest float below extent at the top), w = 1.
extern "C" void tc_gen_gaussian_mixture(float* pts, uint64_t n, uint64_t seed, uint32_t n_centres, float sigma,
                                        float extent) {
    if (n_centres == 0) n_centres = 1;
    std::vector<double> centre(3 * (size_t)n_centres);
    const uint64_t cbase = mix64(seed ^ 0xCE47E5ull);
    for (uint32_t j = 0; j < n_centres; ++j)
        for (int a = 0; a < 3; ++a)
            centre[3 * j + a] = (double)(unit24(stream_at(cbase, 3 * (uint64_t)j + a)) * extent);
    const float top = std::nextafter(extent, 0.0f);
    const uint64_t base = mix64(seed ^ 0x6A055u);
    const double two_pi = 6.283185307179586476925286766559;
    parallel_for(n, [&](uint64_t lo, uint64_t hi) {
        for (uint64_t i = lo; i < hi; ++i) {
            const uint64_t st = base + i * 8 * GOLDEN;
            const uint32_t c = (uint32_t)(stream_at(st, 0) % n_centres);
            const double r0 = std::sqrt(-2.0 * std::log(open01(stream_at(st, 1))));
            const double t0 = two_pi * half01(stream_at(st, 2));
            const double r1 = std::sqrt(-2.0 * std::log(open01(stream_at(st, 3))));
            const double t1 = two_pi * half01(stream_at(st, 4));
            const double z[3] = {r0 * std::cos(t0), r0 * std::sin(t0), r1 * std::cos(t1)};
            for (int a = 0; a < 3; ++a) {
                float f = (float)(centre[3 * c + a] + (double)sigma * z[a]);
                if (!(f >= 0.0f)) f = 0.0f;
                if (f >= extent) f = top;
                pts[4 * i + a] = f;
            }
            pts[4 * i + 3] = 1.0f;
        }
    });
}
