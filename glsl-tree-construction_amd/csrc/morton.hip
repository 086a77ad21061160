// morton.hip -- Morton encoding / decoding kernels.
//
// Reference: k_ComputeMortonCode, shader/morton.comp:60-67 (30-bit codes from
// float[4] points; grid-stride, 512 threads).  Here one point per lane per
// iteration: a 16 B float4 load (a wave reads 1 KiB contiguous) and a 4 B (or
// 8 B) code store, 20 B (24 B) of HBM traffic per point.
#include "tc_internal.hpp"

namespace {

constexpr int MORTON_BLOCK = 256;

__global__ __launch_bounds__(MORTON_BLOCK) void k_morton30(const float4* __restrict__ pts,
                                                           uint32_t* __restrict__ codes, uint32_t n,
                                                           float min_coord, float range) {
    const uint32_t stride = gridDim.x * MORTON_BLOCK;
    for (uint32_t i = blockIdx.x * MORTON_BLOCK + threadIdx.x; i < n; i += stride) {
        const float4 p = pts[i];
        codes[i] = d_encode30(p.x, p.y, p.z, min_coord, range);
    }
}

__global__ __launch_bounds__(MORTON_BLOCK) void k_morton63(const float4* __restrict__ pts,
                                                           uint64_t* __restrict__ codes, uint32_t n,
                                                           float min_coord, float range) {
    const uint32_t stride = gridDim.x * MORTON_BLOCK;
    for (uint32_t i = blockIdx.x * MORTON_BLOCK + threadIdx.x; i < n; i += stride) {
        const float4 p = pts[i];
        codes[i] = d_encode63(p.x, p.y, p.z, min_coord, range);
    }
}

__global__ __launch_bounds__(MORTON_BLOCK) void k_morton30_decode(const uint32_t* __restrict__ codes,
                                                                  float4* __restrict__ out, uint32_t n,
                                                                  float min_coord, float range) {
    const uint32_t stride = gridDim.x * MORTON_BLOCK;
    for (uint32_t i = blockIdx.x * MORTON_BLOCK + threadIdx.x; i < n; i += stride) {
        float c[4];
        CodeTraits<uint32_t>::decode(codes[i], min_coord, range, c);
        out[i] = make_float4(c[0], c[1], c[2], c[3]);
    }
}

uint32_t morton_grid(tc_ctx* ctx, uint32_t n) {
    uint32_t blocks = ceil_div_u32(n, MORTON_BLOCK);
    uint32_t cap = (uint32_t)ctx->num_cus * 16;
    return blocks < cap ? blocks : cap;
}

}  // namespace

extern "C" int tc_morton30(tc_ctx* ctx, const float* pts, uint32_t* codes, uint32_t n, float min_coord,
                           float range, tc_stream stream) {
    if (!ctx || (n && (!pts || !codes))) return TC_ERR_INVALID;
    if (reinterpret_cast<uintptr_t>(pts) & 15u) return TC_ERR_INVALID;
    if (n == 0) return TC_OK;
    ProfScope prof(ctx, as_stream(stream), K_MORTON);
    hipLaunchKernelGGL(k_morton30,dim3(morton_grid(ctx, n)), dim3(MORTON_BLOCK), 0, as_stream(stream),
                       reinterpret_cast<const float4*>(pts), codes, n, min_coord, range);
    TC_LAUNCH_CHECK();
    return TC_OK;
}

extern "C" int tc_morton63(tc_ctx* ctx, const float* pts, uint64_t* codes, uint32_t n, float min_coord,
                           float range, tc_stream stream) {
    if (!ctx || (n && (!pts || !codes))) return TC_ERR_INVALID;
    if (reinterpret_cast<uintptr_t>(pts) & 15u) return TC_ERR_INVALID;
    if (n == 0) return TC_OK;
    ProfScope prof(ctx, as_stream(stream), K_MORTON);
    hipLaunchKernelGGL(k_morton63,dim3(morton_grid(ctx, n)), dim3(MORTON_BLOCK), 0, as_stream(stream),
                       reinterpret_cast<const float4*>(pts), codes, n, min_coord, range);
    TC_LAUNCH_CHECK();
    return TC_OK;
}

extern "C" int tc_morton30_decode(tc_ctx* ctx, const uint32_t* codes, float* corners, uint32_t n,
                                  float min_coord, float range, tc_stream stream) {
    if (!ctx || (n && (!codes || !corners))) return TC_ERR_INVALID;
    if (reinterpret_cast<uintptr_t>(corners) & 15u) return TC_ERR_INVALID;
    if (n == 0) return TC_OK;
    ProfScope prof(ctx, as_stream(stream), K_DECODE);
    hipLaunchKernelGGL(k_morton30_decode, dim3(morton_grid(ctx, n)), dim3(MORTON_BLOCK), 0,
                       as_stream(stream), codes, reinterpret_cast<float4*>(corners), n, min_coord, range);
    TC_LAUNCH_CHECK();
    return TC_OK;
}
