// capi.cpp -- context, scratch management, the plain pipeline and the
// host-array mirrors of the reference's per-stage Rust functions.
//
// The reference creates a Vulkan instance, device and queue inside every stage
// call and round-trips each stage's arrays through host-visible buffers
// (e.g. src/morton.rs:84-168,227-357).  Here one context owns the device
// scratch; device-pointer entry points are stream-ordered; the host-array
// mirrors copy in, run, copy out.
#include <hip/hip_runtime.h>

#include <dlfcn.h>

#include <algorithm>
#include <cstring>
#include <new>

#include "tc_internal.hpp"

namespace {

constexpr int FILL_BLOCK = 256;

__global__ __launch_bounds__(FILL_BLOCK) void k_fill_u32(uint32_t* __restrict__ p, uint32_t v, size_t n) {
    const size_t stride = (size_t)gridDim.x * FILL_BLOCK;
    for (size_t i = (size_t)blockIdx.x * FILL_BLOCK + threadIdx.x; i < n; i += stride) p[i] = v;
}

// Reference kernels of known traffic for calibrating hardware counters and for
// the achievable bandwidth (tc_traffic_reference): stream `nvec` 16-byte
// vectors through plain or non-temporal loads (folded with XOR into the sink
// word of the workgroup), write them with 16-byte stores, or copy them.  Each
// workgroup streams one contiguous chunk of TRAFFIC_CHUNK vectors, eight per
// lane in flight, and there are as many workgroups as chunks: on MI355X a
// 1 GiB copy runs at 5.6 TB/s this way, at 4.3-4.8 TB/s as a grid-stride loop
// over the whole buffer.
constexpr uint32_t TRAFFIC_U = 8;
constexpr uint32_t TRAFFIC_CHUNK = 256 * TRAFFIC_U;

template <bool NT>
__global__ __launch_bounds__(256) void k_traffic_read(const uint4* __restrict__ src, size_t nvec,
                                                      uint32_t* __restrict__ sink, uint32_t sink_words) {
    __shared__ uint32_t s_x[4];
    uint32_t x = 0;
    const size_t c0 = (size_t)blockIdx.x * TRAFFIC_CHUNK + threadIdx.x;
    uint4 v[TRAFFIC_U];
#pragma unroll
    for (uint32_t u = 0; u < TRAFFIC_U; ++u) {
        const size_t i = c0 + u * 256;
        v[u] = i < nvec ? (NT ? ld_stream(&src[i]) : src[i]) : make_uint4(0, 0, 0, 0);
    }
#pragma unroll
    for (uint32_t u = 0; u < TRAFFIC_U; ++u) x ^= v[u].x ^ v[u].y ^ v[u].z ^ v[u].w;
    for (int off = 32; off >= 1; off >>= 1) x ^= __shfl_xor(x, off, 64);
    if ((threadIdx.x & 63u) == 0) s_x[threadIdx.x >> 6] = x;
    __syncthreads();
    if (threadIdx.x == 0) atomicXor(&sink[blockIdx.x % sink_words], s_x[0] ^ s_x[1] ^ s_x[2] ^ s_x[3]);
}

__global__ __launch_bounds__(256) void k_traffic_write(uint4* __restrict__ dst, size_t nvec, uint32_t v) {
    const size_t c0 = (size_t)blockIdx.x * TRAFFIC_CHUNK + threadIdx.x;
#pragma unroll
    for (uint32_t u = 0; u < TRAFFIC_U; ++u)
        if (c0 + u * 256 < nvec) dst[c0 + u * 256] = make_uint4(v, v, v, v);
}

// Copy: eight non-temporal 16-byte loads in flight per lane, then eight 16-byte stores.
__global__ __launch_bounds__(256) void k_traffic_copy(const uint4* __restrict__ src, uint4* __restrict__ dst,
                                                      size_t nvec) {
    const size_t c0 = (size_t)blockIdx.x * TRAFFIC_CHUNK + threadIdx.x;
    uint4 v[TRAFFIC_U];
#pragma unroll
    for (uint32_t u = 0; u < TRAFFIC_U; ++u)
        if (c0 + u * 256 < nvec) v[u] = ld_stream(&src[c0 + u * 256]);
#pragma unroll
    for (uint32_t u = 0; u < TRAFFIC_U; ++u)
        if (c0 + u * 256 < nvec) dst[c0 + u * 256] = v[u];
}

struct DevBuf {
    void* p = nullptr;
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
    int alloc(size_t bytes) {
        if (bytes == 0) bytes = 16;
        return hipMalloc(&p, bytes) == hipSuccess ? TC_OK : TC_ERR_NOMEM;
    }
    template <typename T> T* as() const { return static_cast<T*>(p); }
};

int set_device(tc_ctx* ctx) { return hipSetDevice(ctx->device) == hipSuccess ? TC_OK : TC_ERR_HIP; }

}  // namespace

extern "C" int tc_traffic_reference(tc_ctx* ctx, void* buf, size_t bytes, int mode, uint32_t* sink,
                                    tc_stream stream) {
    if (!ctx || !buf || (mode < 2 && !sink) || mode < 0 || mode > 3) return TC_ERR_INVALID;
    if ((reinterpret_cast<uintptr_t>(buf) & 15u) || (bytes & (mode == 3 ? 31u : 15u))) return TC_ERR_INVALID;
    const size_t nvec = bytes / 16;
    const size_t moved = mode == 3 ? nvec / 2 : nvec;  // vectors each workgroup set streams through
    const size_t chunks = (moved + TRAFFIC_CHUNK - 1) / TRAFFIC_CHUNK;
    if (chunks == 0) return TC_OK;
    if (chunks > 0x7FFFFFFFu) return TC_ERR_TOO_LARGE;
    const uint32_t grid = (uint32_t)chunks;
    hipStream_t s = as_stream(stream);
    if (mode == 0)
        hipLaunchKernelGGL(k_traffic_read<false>, dim3(grid), dim3(256), 0, s, static_cast<const uint4*>(buf), nvec,
                           sink, (uint32_t)TC_TRAFFIC_SINK_WORDS);
    else if (mode == 1)
        hipLaunchKernelGGL(k_traffic_read<true>, dim3(grid), dim3(256), 0, s, static_cast<const uint4*>(buf), nvec,
                           sink, (uint32_t)TC_TRAFFIC_SINK_WORDS);
    else if (mode == 2)
        hipLaunchKernelGGL(k_traffic_write, dim3(grid), dim3(256), 0, s, static_cast<uint4*>(buf), nvec, 0x5A5A5A5Au);
    else
        hipLaunchKernelGGL(k_traffic_copy, dim3(grid), dim3(256), 0, s, static_cast<const uint4*>(buf),
                           static_cast<uint4*>(buf) + nvec / 2, nvec / 2);
    TC_LAUNCH_CHECK();
    return TC_OK;
}

int tc_fill_u32(tc_ctx* ctx, uint32_t* ptr, uint32_t value, size_t count, hipStream_t s) {
    if (count == 0) return TC_OK;
    size_t blocks = (count + FILL_BLOCK - 1) / FILL_BLOCK;
    const size_t cap = (size_t)ctx->num_cus * 8;
    if (blocks > cap) blocks = cap;
    ProfScope prof(ctx, s, K_FILL);
    hipLaunchKernelGGL(k_fill_u32, dim3((uint32_t)blocks), dim3(FILL_BLOCK), 0, s, ptr, value, count);
    TC_LAUNCH_CHECK();
    return TC_OK;
}

int tc_ensure_scratch(tc_ctx* ctx, size_t bytes) {
    if (bytes <= ctx->scratch_bytes) return TC_OK;
    if (ctx->scratch) {
        // Earlier work may still use the old block.
        TC_HIP(hipDeviceSynchronize());
        TC_HIP(hipFree(ctx->scratch));
        ctx->scratch = nullptr;
        ctx->scratch_bytes = 0;
    }
    size_t want = bytes + bytes / 4;
    want = (want + 4095) & ~(size_t)4095;
    if (hipMalloc(&ctx->scratch, want) != hipSuccess) {
        ctx->scratch = nullptr;
        return TC_ERR_NOMEM;
    }
    ctx->scratch_bytes = want;
    return TC_OK;
}

extern "C" const char* tc_status_string(int status) {
    switch (status) {
        case TC_OK: return "ok";
        case TC_ERR_INVALID: return "invalid argument";
        case TC_ERR_NOMEM: return "out of memory";
        case TC_ERR_TOO_LARGE: return "size above the 2^30 limit";
        case TC_ERR_HIP: return "HIP runtime error";
        case TC_ERR_OVER_BUDGET: return "over the memory budget";
        case TC_ERR_CAPACITY: return "output capacity too small";
        case TC_ERR_BUSY: return "busy";
        case TC_ERR_NOT_FOUND: return "not found";
        case TC_ERR_NO_DEVICE: return "no usable GPU";
        case TC_ERR_BAD_TREE: return "tree invariant violated";
        case TC_ERR_IO: return "file I/O error";
        case TC_ERR_CANCELLED: return "cancelled";
        default: return "unknown status";
    }
}

extern "C" const char* tc_version(void) { return "tc 0.1 (gfx950)"; }

// ------------------------------------------------------------------ profiler
hipEvent_t tc_prof_event(tc_ctx* ctx) {
    if (!ctx->prof_pool.empty()) {
        hipEvent_t e = ctx->prof_pool.back();
        ctx->prof_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

const char* tc_kernel_name_internal(int id) {
    static const char* names[K_COUNT] = {"morton", "histogram", "binning", "unique", "scan", "radix_tree",
                                         "edge_count", "edge_count_scan", "octree", "octree_link",
                                         "fill", "decode", "segmented"};
    return (id >= 0 && id < K_COUNT) ? names[id] : "?";
}

extern "C" const char* tc_kernel_name(int id) { return tc_kernel_name_internal(id); }

// ------------------------------------------------------------ roctx markers
TcRoctx g_tc_roctx;

extern "C" int tc_set_markers(tc_ctx* ctx, int on) {
    if (!ctx) return TC_ERR_INVALID;
    if (on && !g_tc_roctx.push) {
        static std::mutex mu;
        std::lock_guard<std::mutex> lock(mu);
        if (!g_tc_roctx.push) {
            void* h = dlopen("librocprofiler-sdk-roctx.so.1", RTLD_NOW | RTLD_LOCAL);
            if (!h) h = dlopen("libroctx64.so.4", RTLD_NOW | RTLD_LOCAL);
            if (!h) return TC_ERR_NOT_FOUND;
            auto push = reinterpret_cast<int (*)(const char*)>(dlsym(h, "roctxRangePushA"));
            auto pop = reinterpret_cast<int (*)()>(dlsym(h, "roctxRangePop"));
            if (!push || !pop) return TC_ERR_NOT_FOUND;
            g_tc_roctx.pop = pop;
            g_tc_roctx.push = push;
        }
    }
    ctx->markers_on = on != 0;
    return TC_OK;
}

extern "C" int tc_profile_enable(tc_ctx* ctx, int on) {
    if (!ctx) return TC_ERR_INVALID;
    ctx->prof_on = on != 0;
    return TC_OK;
}

extern "C" int tc_profile_collect(tc_ctx* ctx) {
    if (!ctx) return TC_ERR_INVALID;
    for (ProfRec& r : ctx->prof_pending) {
        if (hipEventSynchronize(r.b) != hipSuccess) return TC_ERR_HIP;
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, r.a, r.b) == hipSuccess) {
            ctx->prof_ms[r.id] += ms;
            ctx->prof_n[r.id] += 1;
        }
        ctx->prof_pool.push_back(r.a);
        ctx->prof_pool.push_back(r.b);
    }
    ctx->prof_pending.clear();
    return TC_OK;
}

extern "C" int tc_profile_reset(tc_ctx* ctx) {
    if (!ctx) return TC_ERR_INVALID;
    TC_TRY(tc_profile_collect(ctx));
    for (int i = 0; i < K_COUNT; ++i) {
        ctx->prof_ms[i] = 0.0;
        ctx->prof_n[i] = 0;
    }
    return TC_OK;
}

extern "C" int tc_profile_get(tc_ctx* ctx, int id, double* total_ms, uint64_t* launches) {
    if (!ctx || !total_ms || !launches || id < 0 || id >= K_COUNT) return TC_ERR_INVALID;
    *total_ms = ctx->prof_ms[id];
    *launches = ctx->prof_n[id];
    return TC_OK;
}

extern "C" int tc_create(tc_ctx** out, int device) {
    if (!out) return TC_ERR_INVALID;
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) return TC_ERR_NO_DEVICE;
    if (device < 0 || device >= count) return TC_ERR_INVALID;
    tc_ctx* ctx = new (std::nothrow) tc_ctx();
    if (!ctx) return TC_ERR_NOMEM;
    ctx->device = device;
    std::memset(ctx->tune, -1, sizeof(ctx->tune));
    if (hipSetDevice(device) != hipSuccess) {
        delete ctx;
        return TC_ERR_HIP;
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess && prop.multiProcessorCount > 0)
        ctx->num_cus = prop.multiProcessorCount;
    const size_t status_bytes = tc_scan_status_bytes();
    const size_t acc_bytes = tc_hist_acc_bytes();
    if (hipMalloc(&ctx->small, SMALL_WORDS * 4) != hipSuccess ||
        hipMalloc(&ctx->scan_status, status_bytes) != hipSuccess ||
        hipMalloc(&ctx->hist_acc, acc_bytes) != hipSuccess ||
        hipMemset(ctx->small, 0, SMALL_WORDS * 4) != hipSuccess ||
        hipMemset(ctx->hist_acc, 0, acc_bytes) != hipSuccess ||
        hipMemset(ctx->scan_status, 0, status_bytes) != hipSuccess ||
        hipDeviceSynchronize() != hipSuccess) {
        tc_destroy(ctx);
        return TC_ERR_NOMEM;
    }
    // A failed lookup only disables tc_sort_keys_u32_noalloc (it then returns TC_ERR_HIP).
    (void)tc_noalloc_init(ctx);
    *out = ctx;
    return TC_OK;
}

extern "C" int tc_destroy(tc_ctx* ctx) {
    if (!ctx) return TC_ERR_INVALID;
    (void)hipSetDevice(ctx->device);
    (void)hipDeviceSynchronize();
    (void)tc_profile_collect(ctx);
    for (hipEvent_t e : ctx->prof_pool) (void)hipEventDestroy(e);
    if (ctx->scratch) (void)hipFree(ctx->scratch);
    if (ctx->small) (void)hipFree(ctx->small);
    if (ctx->hist_acc) (void)hipFree(ctx->hist_acc);
    if (ctx->scan_status) (void)hipFree(ctx->scan_status);
    delete ctx;
    return TC_OK;
}

extern "C" int tc_reserve(tc_ctx* ctx, uint32_t max_n) {
    if (!ctx) return TC_ERR_INVALID;
    if (max_n > TC_MAX_SORT_N) return TC_ERR_TOO_LARGE;
    TC_TRY(set_device(ctx));
    return tc_ensure_scratch(ctx, tc_sort_scratch_max_bytes(max_n));
}

extern "C" int tc_set_sort_params(tc_ctx* ctx, int block, int kpt) {
    return tc_set_sort_plan(ctx, block, kpt, 8);
}

extern "C" int tc_get_sort_params(tc_ctx* ctx, int* block, int* kpt) {
    if (!ctx || !block || !kpt) return TC_ERR_INVALID;
    *block = ctx->sort_block;
    *kpt = ctx->sort_kpt;
    return TC_OK;
}

extern "C" int tc_get_sort_plan(tc_ctx* ctx, int* block, int* kpt, int* digit_bits) {
    if (!ctx || !block || !kpt || !digit_bits) return TC_ERR_INVALID;
    *block = ctx->sort_block;
    *kpt = ctx->sort_kpt;
    *digit_bits = ctx->sort_bits;
    return TC_OK;
}

// ------------------------------------------------------------ plain pipeline
extern "C" int tc_tree_alloc(tc_ctx* ctx, uint32_t capacity, int code_bits, tc_tree* t) {
    return tc_tree_alloc_ex(ctx, capacity, code_bits, 0, t);
}

extern "C" int tc_tree_alloc_ex(tc_ctx* ctx, uint32_t capacity, int code_bits, int flags, tc_tree* t) {
    if (!ctx || !t || (code_bits != 30 && code_bits != 63)) return TC_ERR_INVALID;
    if (flags & ~TC_TREE_POINT_INDEX) return TC_ERR_INVALID;
    if (capacity > TC_MAX_SORT_N) return TC_ERR_TOO_LARGE;
    TC_TRY(set_device(ctx));
    std::memset(t, 0, sizeof(*t));
    t->capacity = capacity;
    t->code_bits = code_bits;
    t->node_capacity = tc_octree_node_bound(capacity, code_bits);
    if (t->node_capacity > 0xFFFFFFFFull) t->node_capacity = 0xFFFFFFFFull;
    const size_t n = capacity ? capacity : 1;
    const size_t ks = code_bits == 30 ? 4 : 8;
    bool ok = hipMalloc(&t->sorted, n * ks) == hipSuccess && hipMalloc(&t->alt, n * ks) == hipSuccess &&
              hipMalloc(&t->unique, n * ks) == hipSuccess &&
              hipMalloc((void**)&t->prefix_n, n) == hipSuccess &&
              hipMalloc((void**)&t->has_leaf_left, n) == hipSuccess &&
              hipMalloc((void**)&t->has_leaf_right, n) == hipSuccess &&
              hipMalloc((void**)&t->left_child, n * 4) == hipSuccess &&
              hipMalloc((void**)&t->parent, n * 4) == hipSuccess &&
              hipMalloc((void**)&t->edge_count, n * 4) == hipSuccess &&
              hipMalloc((void**)&t->offsets, (n + 1) * 4) == hipSuccess &&
              hipMalloc((void**)&t->nodes, std::max<uint64_t>(t->node_capacity, 1) * sizeof(tc_oct_node)) ==
                  hipSuccess &&
              hipMalloc((void**)&t->d_counts, 16) == hipSuccess;
    if (ok && (flags & TC_TREE_POINT_INDEX))
        ok = hipMalloc((void**)&t->point_index, n * 4) == hipSuccess &&
             hipMalloc((void**)&t->index_alt, n * 4) == hipSuccess &&
             hipMalloc((void**)&t->leaf_first, (n + 1) * 4) == hipSuccess;
    if (!ok) {
        tc_tree_free(ctx, t);
        return TC_ERR_NOMEM;
    }
    return tc_reserve(ctx, capacity);
}

extern "C" int tc_tree_free(tc_ctx* ctx, tc_tree* t) {
    if (!ctx || !t) return TC_ERR_INVALID;
    (void)hipSetDevice(ctx->device);
    void* ptrs[] = {t->sorted, t->alt, t->unique, t->prefix_n, t->has_leaf_left, t->has_leaf_right,
                    t->left_child, t->parent, t->edge_count, t->offsets, t->nodes, t->d_counts,
                    t->point_index, t->index_alt, t->leaf_first};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    std::memset(t, 0, sizeof(*t));
    return TC_OK;
}

extern "C" int tc_build_tree(tc_ctx* ctx, const float* pts, uint32_t n, float min_coord, float range,
                             int link_leaves, tc_tree* t, tc_stream stream) {
    if (!ctx || !t || (n && !pts)) return TC_ERR_INVALID;
    if (n > t->capacity) return TC_ERR_CAPACITY;
    hipStream_t s = as_stream(stream);
    const MarkRange mark(ctx, "tc_build_tree");
    t->n_points = n;
    t->min_coord = min_coord;
    t->range = range;
    t->link_leaves = link_leaves ? 1 : 0;
    // An empty cloud reads back as empty; otherwise unique writes the count.
    if (n == 0) return tc_fill_u32(ctx, t->d_counts, 0u, 4, s);
    const uint32_t node_cap = (uint32_t)std::min<uint64_t>(t->node_capacity, 0xFFFFFFFFull);
    // morton -> sort -> unique (count stays on the device) -> radix tree with
    // fused edge counts -> exclusive scan -> octree.  Stages after unique read
    // the unique count from d_counts; the scan always writes offsets[n_brt].
    // With a point index the sort carries each point's index as its payload and
    // unique also records where each code's run starts (the leaf -> points map).
    // The scans skip their one-lane epoch kernel: the sort's histogram-finish
    // kernel and the radix-tree kernel (launched when n >= 2) advance the tag.
    const bool indexed = t->point_index != nullptr;
    if (t->code_bits == 30) {
        uint32_t* sorted = static_cast<uint32_t*>(t->sorted);
        uint32_t* uniq = static_cast<uint32_t*>(t->unique);
        if (indexed) {
            TC_TRY(tc_morton_sort_pairs_u32(ctx, pts, sorted, t->point_index, static_cast<uint32_t*>(t->alt),
                                            t->index_alt, n, min_coord, range, stream));
            TC_TRY(tc_launch_scan(ctx, MODE_UNIQUE32, sorted, nullptr, uniq, t->leaf_first, t->d_counts, n,
                                  nullptr, nullptr, 0, nullptr, nullptr, s, true));
        } else {
            TC_TRY(tc_morton_sort_u32(ctx, pts, sorted, static_cast<uint32_t*>(t->alt), n, min_coord, range, stream));
            TC_TRY(tc_launch_scan(ctx, MODE_UNIQUE32, sorted, nullptr, uniq, nullptr, t->d_counts, n,
                                  nullptr, nullptr, 0, nullptr, nullptr, s, true));
        }
        TC_TRY(tc_build_radix_tree_ec_u32(ctx, uniq, n, t->d_counts, t->prefix_n, t->has_leaf_left,
                                          t->has_leaf_right, t->left_child, t->parent, t->edge_count, stream));
        TC_TRY(tc_launch_scan(ctx, MODE_SCAN, t->edge_count, nullptr, t->offsets, nullptr, t->d_counts + 1, n - 1,
                              t->d_counts, nullptr, 0, nullptr, nullptr, s, n >= 2));
        if (n >= 2)
            TC_TRY(tc_build_octree_u32(ctx, t->nodes, node_cap, t->offsets, t->edge_count, uniq, t->prefix_n,
                                       t->parent, t->left_child, t->has_leaf_left, t->has_leaf_right, n - 1,
                                       t->d_counts, min_coord, range, link_leaves, stream));
    } else {
        uint64_t* sorted = static_cast<uint64_t*>(t->sorted);
        uint64_t* uniq = static_cast<uint64_t*>(t->unique);
        if (indexed) {
            TC_TRY(tc_morton_sort_pairs_u64(ctx, pts, sorted, t->point_index, static_cast<uint64_t*>(t->alt),
                                            t->index_alt, n, min_coord, range, stream));
            TC_TRY(tc_launch_scan(ctx, MODE_UNIQUE64, sorted, nullptr, uniq, t->leaf_first, t->d_counts, n,
                                  nullptr, nullptr, 0, nullptr, nullptr, s, true));
        } else {
            TC_TRY(tc_morton_sort_u64(ctx, pts, sorted, static_cast<uint64_t*>(t->alt), n, min_coord, range, stream));
            TC_TRY(tc_launch_scan(ctx, MODE_UNIQUE64, sorted, nullptr, uniq, nullptr, t->d_counts, n,
                                  nullptr, nullptr, 0, nullptr, nullptr, s, true));
        }
        TC_TRY(tc_build_radix_tree_ec_u64(ctx, uniq, n, t->d_counts, t->prefix_n, t->has_leaf_left,
                                          t->has_leaf_right, t->left_child, t->parent, t->edge_count, stream));
        TC_TRY(tc_launch_scan(ctx, MODE_SCAN, t->edge_count, nullptr, t->offsets, nullptr, t->d_counts + 1, n - 1,
                              t->d_counts, nullptr, 0, nullptr, nullptr, s, n >= 2));
        if (n >= 2)
            TC_TRY(tc_build_octree_u64(ctx, t->nodes, node_cap, t->offsets, t->edge_count, uniq, t->prefix_n,
                                       t->parent, t->left_child, t->has_leaf_left, t->has_leaf_right, n - 1,
                                       t->d_counts, min_coord, range, link_leaves, stream));
    }
    return TC_OK;
}

extern "C" int tc_tree_read_counts(tc_ctx* ctx, tc_tree* t, tc_stream stream) {
    if (!ctx || !t) return TC_ERR_INVALID;
    hipStream_t s = as_stream(stream);
    TC_HIP(hipStreamSynchronize(s));
    uint32_t u = 0;
    TC_HIP(hipMemcpy(&u, t->d_counts, 4, hipMemcpyDeviceToHost));
    t->n_unique = u;
    t->n_brt = u > 1 ? u - 1 : 0;
    t->n_nodes = 0;
    if (t->n_brt > 0) TC_HIP(hipMemcpy(&t->n_nodes, t->offsets + t->n_brt, 4, hipMemcpyDeviceToHost));
    if (t->n_nodes > t->node_capacity) return TC_ERR_CAPACITY;
    return TC_OK;
}

// ------------------------------------------- tree download, check, export
namespace {
// Host copies of a device tree's arrays, described by a tc_tree_host.
struct HostTree {
    std::vector<uint8_t> sorted, unique, prefix_n, hll, hlr;
    std::vector<int32_t> left, parent;
    std::vector<uint32_t> ec, offsets, point_index, leaf_first;
    std::vector<tc_oct_node> nodes;
    tc_tree_host view{};
};

template <typename T>
int d2h(std::vector<T>& v, const void* d, size_t count) {
    v.resize(count);
    if (count) TC_HIP(hipMemcpy(v.data(), d, count * sizeof(T), hipMemcpyDeviceToHost));
    return TC_OK;
}

int download_tree(tc_ctx* ctx, tc_tree* t, tc_stream stream, HostTree* h) {
    TC_TRY(tc_tree_read_counts(ctx, t, stream));
    const size_t ks = t->code_bits == 30 ? 4 : 8;
    const uint32_t n = t->n_points, u = t->n_unique, nb = t->n_brt;
    TC_TRY(d2h(h->sorted, t->sorted, n * ks));
    TC_TRY(d2h(h->unique, t->unique, u * ks));
    TC_TRY(d2h(h->prefix_n, t->prefix_n, nb));
    TC_TRY(d2h(h->hll, t->has_leaf_left, nb));
    TC_TRY(d2h(h->hlr, t->has_leaf_right, nb));
    TC_TRY(d2h(h->left, t->left_child, nb));
    TC_TRY(d2h(h->parent, t->parent, nb));
    TC_TRY(d2h(h->ec, t->edge_count, nb));
    TC_TRY(d2h(h->offsets, t->offsets, nb ? nb + 1 : 0));
    TC_TRY(d2h(h->nodes, t->nodes, t->n_nodes));
    if (t->point_index) {
        TC_TRY(d2h(h->point_index, t->point_index, n));
        if (n) TC_TRY(d2h(h->leaf_first, t->leaf_first, u + 1));
        else h->leaf_first.assign(1, 0u);
    }
    tc_tree_host& v = h->view;
    v.code_bits = t->code_bits;
    v.min_coord = t->min_coord;
    v.range = t->range;
    v.link_leaves = t->link_leaves;
    v.n_points = n;
    v.n_unique = u;
    v.n_nodes = t->n_nodes;
    v.sorted = h->sorted.data();
    v.unique = h->unique.data();
    v.prefix_n = h->prefix_n.data();
    v.has_leaf_left = h->hll.data();
    v.has_leaf_right = h->hlr.data();
    v.left_child = h->left.data();
    v.parent = h->parent.data();
    v.edge_count = h->ec.data();
    v.offsets = nb ? h->offsets.data() : nullptr;
    v.nodes = h->nodes.data();
    v.point_index = t->point_index ? h->point_index.data() : nullptr;
    v.leaf_first = t->point_index ? h->leaf_first.data() : nullptr;
    return TC_OK;
}
}  // namespace

extern "C" int tc_tree_validate(tc_ctx* ctx, tc_tree* t, char* msg, size_t msg_len, tc_stream stream) {
    if (!ctx || !t) return TC_ERR_INVALID;
    HostTree h;
    TC_TRY(download_tree(ctx, t, stream, &h));
    return tc_tree_validate_host(&h.view, msg, msg_len);
}

extern "C" int tc_tree_export(tc_ctx* ctx, tc_tree* t, const char* path, tc_stream stream) {
    if (!ctx || !t || !path) return TC_ERR_INVALID;
    HostTree h;
    TC_TRY(download_tree(ctx, t, stream, &h));
    return tc_tree_write_host(&h.view, path);
}

// ------------------------------------------------------ host-array mirrors
// Each mirror follows the argument order of the cited Rust function; the
// reference's group_size (its dispatch width) is accepted and ignored because
// the launch shape is chosen per kernel here.
namespace {
int h2d(void* d, const void* h, size_t bytes) {
    if (!bytes) return TC_OK;
    return hipMemcpy(d, h, bytes, hipMemcpyHostToDevice) == hipSuccess ? TC_OK : TC_ERR_HIP;
}
int d2h(void* h, const void* d, size_t bytes) {
    if (!bytes) return TC_OK;
    return hipMemcpy(h, d, bytes, hipMemcpyDeviceToHost) == hipSuccess ? TC_OK : TC_ERR_HIP;
}
int sync0() { return hipDeviceSynchronize() == hipSuccess ? TC_OK : TC_ERR_HIP; }
}  // namespace

// compute_morton, src/morton.rs:76-83.
extern "C" int tc_host_compute_morton(tc_ctx* ctx, const float* input, uint32_t* keys, uint32_t n, float min_coord,
                                      float range, uint32_t group_size) {
    (void)group_size;
    if (!ctx || (n && (!input || !keys))) return TC_ERR_INVALID;
    std::lock_guard<std::mutex> lk(ctx->mu);
    TC_TRY(set_device(ctx));
    DevBuf dp, dk;
    TC_TRY(dp.alloc((size_t)n * 16));
    TC_TRY(dk.alloc((size_t)n * 4));
    TC_TRY(h2d(dp.p, input, (size_t)n * 16));
    TC_TRY(tc_morton30(ctx, dp.as<float>(), dk.as<uint32_t>(), n, min_coord, range, nullptr));
    TC_TRY(sync0());
    return d2h(keys, dk.p, (size_t)n * 4);
}

// histogram, src/histogram.rs:38 (global_hist holds 4 x 256 words).
extern "C" int tc_host_histogram(tc_ctx* ctx, uint32_t input_size, const uint32_t* input, uint32_t* global_hist) {
    if (!ctx || !global_hist || (input_size && !input)) return TC_ERR_INVALID;
    std::lock_guard<std::mutex> lk(ctx->mu);
    TC_TRY(set_device(ctx));
    DevBuf dk, dh;
    TC_TRY(dk.alloc((size_t)input_size * 4));
    TC_TRY(dh.alloc(4 * 256 * 4));
    TC_TRY(h2d(dk.p, input, (size_t)input_size * 4));
    TC_TRY(tc_global_histogram_u32(ctx, dk.as<uint32_t>(), input_size, dh.as<uint32_t>(), nullptr));
    TC_TRY(sync0());
    return d2h(global_hist, dh.p, 4 * 256 * 4);
}

// radix_sort, src/radixsort.rs:36 (the pipeline's contract: ascending u32 keys).
extern "C" int tc_host_radix_sort(tc_ctx* ctx, const uint32_t* in, uint32_t n, uint32_t* out, uint32_t group_size) {
    (void)group_size;
    if (!ctx || (n && (!in || !out))) return TC_ERR_INVALID;
    std::lock_guard<std::mutex> lk(ctx->mu);
    TC_TRY(set_device(ctx));
    DevBuf dk, da;
    TC_TRY(dk.alloc((size_t)n * 4));
    TC_TRY(da.alloc((size_t)n * 4));
    TC_TRY(h2d(dk.p, in, (size_t)n * 4));
    TC_TRY(tc_sort_keys_u32(ctx, dk.as<uint32_t>(), da.as<uint32_t>(), n, nullptr));
    TC_TRY(sync0());
    return d2h(out, dk.p, (size_t)n * 4);
}

// unique, src/unique.rs:37: u_sort[0..U) receives the unique keys, U returned.
extern "C" int tc_host_unique(tc_ctx* ctx, uint32_t* u_sort, uint32_t n, uint32_t* n_unique) {
    if (!ctx || !n_unique || (n && !u_sort)) return TC_ERR_INVALID;
    std::lock_guard<std::mutex> lk(ctx->mu);
    TC_TRY(set_device(ctx));
    DevBuf dk, dout, dc;
    TC_TRY(dk.alloc((size_t)n * 4));
    TC_TRY(dout.alloc((size_t)n * 4));
    TC_TRY(dc.alloc(4));
    TC_TRY(h2d(dk.p, u_sort, (size_t)n * 4));
    TC_TRY(tc_unique_u32(ctx, dk.as<uint32_t>(), dout.as<uint32_t>(), n, dc.as<uint32_t>(), nullptr));
    TC_TRY(sync0());
    TC_TRY(d2h(n_unique, dc.p, 4));
    return d2h(u_sort, dout.p, (size_t)*n_unique * 4);
}

// build_radix_tree, src/build_radix_tree.rs:34-43: n is the node count
// (n_unique - 1, as src/main.rs:106-107 passes); codes holds n + 1 keys.
extern "C" int tc_host_build_radix_tree(tc_ctx* ctx, int32_t n, const uint32_t* codes, uint8_t* prefix_n,
                                        uint8_t* hll, uint8_t* hlr, int32_t* left, int32_t* parent,
                                        uint32_t group_size) {
    (void)group_size;
    if (!ctx || n < 0) return TC_ERR_INVALID;
    if (n == 0) return TC_OK;
    if (!codes || !prefix_n || !hll || !hlr || !left || !parent) return TC_ERR_INVALID;
    std::lock_guard<std::mutex> lk(ctx->mu);
    TC_TRY(set_device(ctx));
    const size_t nb = (size_t)n, nk = nb + 1;
    DevBuf dc, dp, dl, dr, dlc, dpa;
    TC_TRY(dc.alloc(nk * 4));
    TC_TRY(dp.alloc(nb));
    TC_TRY(dl.alloc(nb));
    TC_TRY(dr.alloc(nb));
    TC_TRY(dlc.alloc(nb * 4));
    TC_TRY(dpa.alloc(nb * 4));
    TC_TRY(h2d(dc.p, codes, nk * 4));
    TC_TRY(tc_build_radix_tree_u32(ctx, dc.as<uint32_t>(), (uint32_t)nk, nullptr, dp.as<uint8_t>(), dl.as<uint8_t>(),
                                   dr.as<uint8_t>(), dlc.as<int32_t>(), dpa.as<int32_t>(), nullptr));
    TC_TRY(sync0());
    TC_TRY(d2h(prefix_n, dp.p, nb));
    TC_TRY(d2h(hll, dl.p, nb));
    TC_TRY(d2h(hlr, dr.p, nb));
    TC_TRY(d2h(left, dlc.p, nb * 4));
    return d2h(parent, dpa.p, nb * 4);
}

// edge_count, src/edge_count.rs:41-47: n_brt_nodes radix nodes (n_unique - 1).
extern "C" int tc_host_edge_count(tc_ctx* ctx, const uint8_t* prefix_n, const int32_t* parent, uint32_t* ec,
                                  uint32_t n_brt, uint32_t group_size) {
    (void)group_size;
    if (!ctx || (n_brt && (!prefix_n || !parent || !ec))) return TC_ERR_INVALID;
    if (n_brt == 0) return TC_OK;
    std::lock_guard<std::mutex> lk(ctx->mu);
    TC_TRY(set_device(ctx));
    DevBuf dp, dpa, de;
    TC_TRY(dp.alloc(n_brt));
    TC_TRY(dpa.alloc((size_t)n_brt * 4));
    TC_TRY(de.alloc((size_t)n_brt * 4));
    TC_TRY(h2d(dp.p, prefix_n, n_brt));
    TC_TRY(h2d(dpa.p, parent, (size_t)n_brt * 4));
    TC_TRY(tc_edge_count(ctx, dp.as<uint8_t>(), dpa.as<int32_t>(), de.as<uint32_t>(), n_brt, nullptr, nullptr));
    TC_TRY(sync0());
    return d2h(ec, de.p, (size_t)n_brt * 4);
}

// partial_sum, src/partial_sum.rs:43 with the fix-up of src/main.rs:127-129:
// d_first[0 .. last-first] = exclusive scan of u_edge_count[first .. last).
extern "C" int tc_host_partial_sum(tc_ctx* ctx, const uint32_t* in, uint32_t first, uint32_t last, uint32_t* d_first) {
    if (!ctx || !d_first || last < first || (last > first && !in)) return TC_ERR_INVALID;
    std::lock_guard<std::mutex> lk(ctx->mu);
    TC_TRY(set_device(ctx));
    const size_t n = last - first;
    DevBuf di, dout;
    TC_TRY(di.alloc(n * 4));
    TC_TRY(dout.alloc((n + 1) * 4));
    TC_TRY(h2d(di.p, in + first, n * 4));
    TC_TRY(tc_exclusive_scan_u32(ctx, di.as<uint32_t>(), dout.as<uint32_t>(), (uint32_t)n, nullptr, nullptr));
    TC_TRY(sync0());
    return d2h(d_first, dout.p, (n + 1) * 4);
}

// build_octree, src/octree.rs:73-87: n radix nodes; oct_nodes must hold
// node_offsets[n] entries (the root included, at index 0).
extern "C" int tc_host_build_octree(tc_ctx* ctx, tc_oct_node* oct_nodes, const uint32_t* offsets,
                                    const uint32_t* counts, const uint32_t* codes, const uint8_t* prefix_n,
                                    const int32_t* parents, float min_coord, float range, int32_t n,
                                    const uint8_t* hll, const uint8_t* hlr, const int32_t* left,
                                    uint32_t group_size) {
    (void)group_size;
    if (!ctx || n < 0) return TC_ERR_INVALID;
    if (n == 0) return TC_OK;
    if (!oct_nodes || !offsets || !counts || !codes || !prefix_n || !parents || !hll || !hlr || !left)
        return TC_ERR_INVALID;
    std::lock_guard<std::mutex> lk(ctx->mu);
    TC_TRY(set_device(ctx));
    const size_t nb = (size_t)n;
    const uint32_t total = offsets[nb];
    DevBuf dn, doff, dcnt, dcode, dpn, dpar, dl, dr, dlc;
    TC_TRY(dn.alloc((size_t)total * sizeof(tc_oct_node)));
    TC_TRY(doff.alloc((nb + 1) * 4));
    TC_TRY(dcnt.alloc(nb * 4));
    TC_TRY(dcode.alloc((nb + 1) * 4));
    TC_TRY(dpn.alloc(nb));
    TC_TRY(dpar.alloc(nb * 4));
    TC_TRY(dl.alloc(nb));
    TC_TRY(dr.alloc(nb));
    TC_TRY(dlc.alloc(nb * 4));
    TC_TRY(h2d(doff.p, offsets, (nb + 1) * 4));
    TC_TRY(h2d(dcnt.p, counts, nb * 4));
    TC_TRY(h2d(dcode.p, codes, (nb + 1) * 4));
    TC_TRY(h2d(dpn.p, prefix_n, nb));
    TC_TRY(h2d(dpar.p, parents, nb * 4));
    TC_TRY(h2d(dl.p, hll, nb));
    TC_TRY(h2d(dr.p, hlr, nb));
    TC_TRY(h2d(dlc.p, left, nb * 4));
    TC_TRY(tc_build_octree_u32(ctx, dn.as<tc_oct_node>(), total, doff.as<uint32_t>(), dcnt.as<uint32_t>(),
                               dcode.as<uint32_t>(), dpn.as<uint8_t>(), dpar.as<int32_t>(), dlc.as<int32_t>(),
                               dl.as<uint8_t>(), dr.as<uint8_t>(), (uint32_t)nb, nullptr, min_coord, range, 1,
                               nullptr));
    TC_TRY(sync0());
    return d2h(oct_nodes, dn.p, (size_t)total * sizeof(tc_oct_node));
}
