/*
 * tc.h -- C ABI of the MI355X-native tree-construction library (libtc.so).
 *
 * Stage chain (reference: ucsc-redwood/glsl-tree-construction, paths relative
 * to vulkan/radixsort-vulkan/):
 *
 *   points --morton--> codes --sort--> sorted --unique--> keys
 *          --radix tree--> (prefix_n, has_leaf_left/right, left_child, parent)
 *          --edge count--> counts --exclusive scan--> offsets --octree--> nodes
 *
 * Conventions
 *  - Device-pointer entry points (tc_<stage>) are asynchronous on the given HIP
 *    stream (passed as void*; NULL is the legacy default stream).  The caller
 *    owns every array it passes; the context owns scratch space.
 *  - Host-array entry points (tc_host_<stage>) mirror the reference's per-stage
 *    Rust functions one for one: they copy host arrays in, run the stage on the
 *    GPU and copy results out, blocking.
 *  - Every function returns TC_OK (0) or a negative tc_status; nothing throws.
 *  - A tc_ctx is externally synchronised: use it from one host thread and one
 *    stream at a time (its scratch is reused by every call).
 *  - Point arrays are float[n][4] (x, y, z, w), 16 B per point.
 *  - Counts are limited to 2^30 per sort (the lookback words pack a 30-bit count).
 *  - Optional device counts: the tree stages take a host capacity and an
 *    optional device pointer `d_n_keys` holding the number of unique keys U
 *    (as written by tc_unique_*).  When it is non-NULL the stage reads U on the
 *    device, so a whole build can be queued without a host synchronisation.
 */
#ifndef TC_H
#define TC_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct tc_ctx tc_ctx;
typedef void* tc_stream; /* hipStream_t */

enum tc_status {
    TC_OK = 0,
    TC_ERR_INVALID = -1,     /* bad argument (NULL pointer, bad size, bad flag) */
    TC_ERR_NOMEM = -2,       /* device or host allocation failed */
    TC_ERR_TOO_LARGE = -3,   /* n above the 2^30 packing limit */
    TC_ERR_HIP = -4,         /* a HIP runtime call failed */
    TC_ERR_OVER_BUDGET = -5, /* request exceeds the memory budget */
    TC_ERR_CAPACITY = -6,    /* output capacity too small */
    TC_ERR_BUSY = -7,        /* queue full / try again */
    TC_ERR_NOT_FOUND = -8,   /* unknown request id */
    TC_ERR_NO_DEVICE = -9,   /* no usable GPU */
    TC_ERR_BAD_TREE = -10,   /* tc_tree_validate*: an invariant does not hold */
    TC_ERR_IO = -11,         /* file could not be written */
    TC_ERR_CANCELLED = -12   /* served request cancelled before it started */
};

#define TC_MAX_SORT_N (1u << 30)

/* 64-byte octree node: OctNode of shader/octree.comp:14-20 (std430 stride 64 B)
 * with an explicit pad so host and device strides agree. */
typedef struct {
    int32_t children[8];
    float corner[4];
    float cell_size;
    int32_t child_node_mask;
    int32_t child_leaf_mask;
    int32_t pad_;
} tc_oct_node;

const char* tc_status_string(int status);
const char* tc_version(void);

/* Context: device, scratch (histograms, lookback words, tile counters). */
int tc_create(tc_ctx** out, int device);
int tc_destroy(tc_ctx* ctx);
/* Pre-size scratch for sorts/scans of up to max_n elements so that later
 * calls allocate nothing (needed before stream capture). */
int tc_reserve(tc_ctx* ctx, uint32_t max_n);
/* Launch parameters of the sort's binning pass, one of the compiled shapes
 * (block, keys per thread) with 8-bit digits: 256 x 8, 512 x {8, 12, 16, 20,
 * 24, 28, 32, 36} and 1024 x 16.  0 keeps the current value. */
int tc_set_sort_params(tc_ctx* ctx, int block, int keys_per_thread);
int tc_get_sort_params(tc_ctx* ctx, int* block, int* keys_per_thread);
/* The same with the digit plan (no reference counterpart; the reference
 * hard-codes 8-bit digits, shader/new_radix_sort.comp:12-16): digit_bits 8
 * sorts in 8-bit passes (the last one narrower for 30- and 63-bit Morton
 * codes); 11 sorts in ceil(key_bits / 11) passes of balanced widths of at most
 * 11 bits (30-bit codes: 3 x 10; 32-bit keys: 11, 11, 10; 63-bit codes:
 * 3 x 11 + 3 x 10; 64-bit keys: 4 x 11 + 2 x 10), built for 1024 x 8 and
 * 1024 x 16.  A plan with an odd number of passes leaves the result in the
 * caller's keys all the same (Morton sorts write their codes to the ping-pong
 * buffers first; plain sorts copy back at the end). */
int tc_set_sort_plan(tc_ctx* ctx, int block, int keys_per_thread, int digit_bits);
int tc_get_sort_plan(tc_ctx* ctx, int* block, int* keys_per_thread, int* digit_bits);
/* Every built (block, keys per thread, digit bits) combination, the tuner's
 * choices: writes up to max entries (any pointer may be NULL) and returns
 * how many there are. */
int tc_sort_plan_list(int* blocks, int* keys_per_thread, int* digit_bits, int max);
/* Autotuner (no reference counterpart): times every binning-pass launch shape
 * and digit plan (tc_sort_plan_list) on random keys at sizes 2^12, 2^14, ... up
 * to max_n, for u32 and u64 keys with and without a u32 payload, and from then
 * on each sort uses the fastest measured at the size nearest its own (batches:
 * at their total point capacity, among the 8-bit plans their segmented passes
 * use).  Blocking.  tc_set_sort_params / tc_set_sort_plan switch back to one
 * fixed shape. */
int tc_autotune_sort(tc_ctx* ctx, uint32_t max_n);
/* The shape a sort of n keys (key_bits 32 or 64, pairs 0/1) would use now. */
int tc_tuned_sort_shape(tc_ctx* ctx, int key_bits, int pairs, uint32_t n, int* block, int* keys_per_thread);
/* The same with the digit plan (8 or 11, see tc_set_sort_plan). */
int tc_tuned_sort_plan(tc_ctx* ctx, int key_bits, int pairs, uint32_t n, int* block, int* keys_per_thread,
                       int* digit_bits);
/* Restore one entry of a saved tuning table (size bucket ceil(log2 n)); the
 * first form restores an 8-bit plan. */
int tc_set_tuned_sort_shape(tc_ctx* ctx, int key_bits, int pairs, int log2_n, int block, int keys_per_thread);
int tc_set_tuned_sort_plan(tc_ctx* ctx, int key_bits, int pairs, int log2_n, int block, int keys_per_thread,
                           int digit_bits);
int tc_clear_sort_tuning(tc_ctx* ctx);
/* The binning-pass launch shapes the library is built with (the tuner's
 * choices, without their digit plans; tc_sort_plan_list has those): writes up
 * to max (block, keys per thread) pairs and returns how many there are. */
int tc_sort_shape_list(int* blocks, int* keys_per_thread, int max);
/* Bound on a workgroup's look-back wait (polls of a predecessor's status
 * without progress) in the sort and scan kernels; past it the workgroup
 * computes that predecessor tile's counts from the input itself and looks on
 * below it, so progress never depends on workgroup start order.  0 makes every
 * workgroup compute every predecessor that way (used to test it).  Default
 * TC_LOOKBACK_SPIN_DEFAULT. */
#define TC_LOOKBACK_SPIN_DEFAULT 64u
int tc_set_lookback_spin_limit(tc_ctx* ctx, uint32_t polls);

/* Per-launch event profiler (the reference times each stage with
 * Instant::now() around submit + fence, e.g. src/morton.rs:327-352).  When
 * enabled, every kernel launch is bracketed by two HIP events on its own
 * stream; collect synchronises on them and accumulates per-kernel totals.
 * Kernel ids: 0 morton, 1 histogram, 2 binning, 3 unique, 4 scan,
 * 5 radix_tree, 6 edge_count, 7 edge_count_scan, 8 octree,
 * 9 octree_link, 10 fill, 11 decode, 12 segmented. */
int tc_profile_enable(tc_ctx* ctx, int on);
/* roctx ranges (no reference counterpart; its timers are Instant::now() around
 * each submit, e.g. src/morton.rs:327-352): with markers on, every kernel
 * launch and every tc_build_tree / tc_batch_build / tc_batch_replay call is a
 * named range for `rocprofv3 --marker-trace` (host-side: the time spent
 * enqueuing).  The ROCm marker library is loaded on first use;
 * TC_ERR_NOT_FOUND if it is absent. */
int tc_set_markers(tc_ctx* ctx, int on);
int tc_profile_collect(tc_ctx* ctx);
/* Measurement aid (no reference counterpart): a reference kernel of known HBM
 * traffic for calibrating hardware counters (rocprofv3 --pmc) and for the
 * achievable bandwidth.  mode 0 reads `bytes` (a multiple of 16, 16-byte
 * aligned buf) with 16-byte vector loads, mode 1 the same with non-temporal
 * loads (the form the pipeline uses for data it reads once), mode 2 writes
 * `bytes` with 16-byte stores, mode 3 copies the first half of buf into the
 * second (bytes a multiple of 32; non-temporal loads, 16-byte stores).  The read
 * modes fold every word into sink (TC_TRAFFIC_SINK_WORDS words) with atomic
 * XOR, so the XOR of the sink's words, zeroed first, is the XOR of the data. */
#define TC_TRAFFIC_SINK_WORDS 2048
int tc_traffic_reference(tc_ctx* ctx, void* buf, size_t bytes, int mode, uint32_t* sink, tc_stream stream);
int tc_profile_reset(tc_ctx* ctx);
int tc_profile_get(tc_ctx* ctx, int kernel_id, double* total_ms, uint64_t* launches);
const char* tc_kernel_name(int kernel_id);

/* ---- Morton encoding: k_ComputeMortonCode, shader/morton.comp:60-67;
 *      host wrapper compute_morton, src/morton.rs:76-83. */
int tc_morton30(tc_ctx* ctx, const float* pts_xyzw, uint32_t* codes, uint32_t n, float min_coord,
                float range, tc_stream stream);
/* 63-bit codes, 21 bits per axis (no reference counterpart; same formula with
 * bit scale 2^21). */
int tc_morton63(tc_ctx* ctx, const float* pts_xyzw, uint64_t* codes, uint32_t n, float min_coord,
                float range, tc_stream stream);
/* Decode codes to cell corners: morton32_to_xyz, src/morton.rs:43-59. */
int tc_morton30_decode(tc_ctx* ctx, const uint32_t* codes, float* corners_xyzw, uint32_t n,
                       float min_coord, float range, tc_stream stream);
/* Host-only scalar decode (no GPU needed): src/morton.rs:43. */
void tc_morton32_to_xyz(float ret[4], uint32_t code, float min_coord, float range);
void tc_morton64_to_xyz(float ret[4], uint64_t code, float min_coord, float range);

/* ---- Global histogram: k_GlobalHistogram, shader/histogram.comp:103-195;
 *      host wrapper histogram(), src/histogram.rs:38.
 * hist[256*d + b] = #{keys with digit d < b} (exclusive prefix, digit-major);
 * 4 digits for u32 keys, 8 for u64. */
int tc_global_histogram_u32(tc_ctx* ctx, const uint32_t* keys, uint32_t n, uint32_t* hist_4x256,
                            tc_stream stream);
int tc_global_histogram_u64(tc_ctx* ctx, const uint64_t* keys, uint32_t n, uint32_t* hist_8x256,
                            tc_stream stream);

/* ---- One OneSweep digit-binning pass: k_DigitBinning,
 *      shader/new_radix_sort.comp:116-314, Params{pass_num, radix_shift}
 *      src/main.rs:175-180.  Stable scatter of src by digit
 *      (key >> radix_shift) & 255 using hist (the full digit-major table from
 *      tc_global_histogram_*).  Values may be NULL. */
int tc_binning_pass_u32(tc_ctx* ctx, const uint32_t* src_k, uint32_t* dst_k, const uint32_t* src_v,
                        uint32_t* dst_v, uint32_t n, const uint32_t* hist, uint32_t radix_shift,
                        tc_stream stream);
int tc_binning_pass_u64(tc_ctx* ctx, const uint64_t* src_k, uint64_t* dst_k, const uint32_t* src_v,
                        uint32_t* dst_v, uint32_t n, const uint32_t* hist, uint32_t radix_shift,
                        tc_stream stream);

/* ---- Full stable LSD sorts (histogram + one binning pass per digit of the
 *      digit plan: 4 or 8 with 8-bit digits), ascending.
 * Contract of u_sort.sort(), src/main.rs:85; API shape of radix_sort,
 * src/radixsort.rs:36.  The result ends in keys (and vals); alt buffers are
 * ping-pong space of n elements. */
int tc_sort_keys_u32(tc_ctx* ctx, uint32_t* keys, uint32_t* keys_alt, uint32_t n, tc_stream stream);
int tc_sort_pairs_u32(tc_ctx* ctx, uint32_t* keys, uint32_t* vals, uint32_t* keys_alt,
                      uint32_t* vals_alt, uint32_t n, tc_stream stream);
int tc_sort_keys_u64(tc_ctx* ctx, uint64_t* keys, uint64_t* keys_alt, uint32_t n, tc_stream stream);
/* Morton encode + sort in one call: the codes of pts_xyzw land in keys, sorted;
 * the sort's histogram is accumulated by the Morton kernel itself, so the
 * codes are not re-read for it (same result as tc_morton30/63 followed by
 * tc_sort_keys_u32/64). */
int tc_morton_sort_u32(tc_ctx* ctx, const float* pts_xyzw, uint32_t* keys, uint32_t* keys_alt, uint32_t n,
                       float min_coord, float range, tc_stream stream);
int tc_morton_sort_u64(tc_ctx* ctx, const float* pts_xyzw, uint64_t* keys, uint64_t* keys_alt, uint32_t n,
                       float min_coord, float range, tc_stream stream);
/* Morton codes sorted together with each point's index (values start as
 * 0..n-1 and travel with the stable sort): keys[k] is the k-th smallest code
 * and index[k] the input position of its point.  No reference counterpart
 * (the reference sorts codes only, src/main.rs:85); the point -> leaf map. */
int tc_morton_sort_pairs_u32(tc_ctx* ctx, const float* pts_xyzw, uint32_t* keys, uint32_t* index,
                             uint32_t* keys_alt, uint32_t* index_alt, uint32_t n, float min_coord,
                             float range, tc_stream stream);
int tc_morton_sort_pairs_u64(tc_ctx* ctx, const float* pts_xyzw, uint64_t* keys, uint32_t* index,
                             uint64_t* keys_alt, uint32_t* index_alt, uint32_t n, float min_coord,
                             float range, tc_stream stream);
int tc_sort_pairs_u64(tc_ctx* ctx, uint64_t* keys, uint32_t* vals, uint64_t* keys_alt,
                      uint32_t* vals_alt, uint32_t n, tc_stream stream);
/* tc_sort_keys_u32 that allocates nothing.  No reference counterpart: it is for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * serving loops, harnesses that forbid hipMalloc).  Its scratch is a static
 * 5.03 MiB device array per GPU, held until the process ends, found at
 * tc_create; it calls no hipMalloc/hipFree.  n <= 2^24 (TC_ERR_TOO_LARGE
 * above).  Every call on one GPU shares that array, whatever the context:
 * calls that overlap on one GPU (other streams, other contexts) race and must
 * be serialised by the caller.  Same result as tc_sort_keys_u32. */
int tc_sort_keys_u32_noalloc(tc_ctx* ctx, uint32_t* keys, uint32_t* keys_alt, uint32_t n, tc_stream stream);

/* ---- Unique: k_CountUnique, shader/unique.comp:31-45; host wrapper
 *      unique(), src/unique.rs:37.  out[0..U) = first key of each run of equal
 *      keys; U is written to *d_n_unique (device).  out must not overlap sorted
 *      (TC_ERR_INVALID). */
int tc_unique_u32(tc_ctx* ctx, const uint32_t* sorted, uint32_t* out, uint32_t n,
                  uint32_t* d_n_unique, tc_stream stream);
int tc_unique_u64(tc_ctx* ctx, const uint64_t* sorted, uint64_t* out, uint32_t n,
                  uint32_t* d_n_unique, tc_stream stream);
/* Unique with run starts (the same std::unique semantics as above, plus
 * run_first[j] = sorted position of the first key of run j and
 * run_first[U] = n; run_first has n + 1 entries).  Run j of sorted codes is
 * leaf j of the radix tree, so with the sort permutation this maps each leaf
 * to its points.  No reference counterpart (the reference keeps only the
 * codes, src/unique.rs:278-283). */
int tc_unique_runs_u32(tc_ctx* ctx, const uint32_t* sorted, uint32_t* out, uint32_t* run_first,
                       uint32_t n, uint32_t* d_n_unique, tc_stream stream);
int tc_unique_runs_u64(tc_ctx* ctx, const uint64_t* sorted, uint64_t* out, uint32_t* run_first,
                       uint32_t n, uint32_t* d_n_unique, tc_stream stream);

/* ---- Karras binary radix tree: k_BuildRadixTree,
 *      shader/build_radix_tree.comp:71-153; host wrapper build_radix_tree,
 *      src/build_radix_tree.rs:34-43.  n_keys strictly increasing codes
 *      (30-bit in u32, 63-bit in u64) give n_keys-1 internal nodes; each output
 *      array holds n_keys-1 entries. */
int tc_build_radix_tree_u32(tc_ctx* ctx, const uint32_t* codes, uint32_t n_keys,
                            const uint32_t* d_n_keys, uint8_t* prefix_n, uint8_t* has_leaf_left,
                            uint8_t* has_leaf_right, int32_t* left_child, int32_t* parent,
                            tc_stream stream);
int tc_build_radix_tree_u64(tc_ctx* ctx, const uint64_t* codes, uint32_t n_keys,
                            const uint32_t* d_n_keys, uint8_t* prefix_n, uint8_t* has_leaf_left,
                            uint8_t* has_leaf_right, int32_t* left_child, int32_t* parent,
                            tc_stream stream);

/* Radix tree and edge counts in one pass (countEdges, shader/edge_count.comp:30-41,
 * fused): each node writes the counts of its internal children, node 0 writes
 * the root's 1; edge_count holds n_keys-1 entries.  Same arrays as
 * tc_build_radix_tree_* followed by tc_edge_count. */
int tc_build_radix_tree_ec_u32(tc_ctx* ctx, const uint32_t* codes, uint32_t n_keys,
                               const uint32_t* d_n_keys, uint8_t* prefix_n, uint8_t* has_leaf_left,
                               uint8_t* has_leaf_right, int32_t* left_child, int32_t* parent,
                               uint32_t* edge_count, tc_stream stream);
int tc_build_radix_tree_ec_u64(tc_ctx* ctx, const uint64_t* codes, uint32_t n_keys,
                               const uint32_t* d_n_keys, uint8_t* prefix_n, uint8_t* has_leaf_left,
                               uint8_t* has_leaf_right, int32_t* left_child, int32_t* parent,
                               uint32_t* edge_count, tc_stream stream);

/* ---- Edge count: countEdges, shader/edge_count.comp:30-41; host wrapper
 *      edge_count(), src/edge_count.rs:41-47.  n_brt = n_keys-1 entries;
 *      entry 0 (the root) is 1. */
int tc_edge_count(tc_ctx* ctx, const uint8_t* prefix_n, const int32_t* parent, uint32_t* edge_count,
                  uint32_t n_brt, const uint32_t* d_n_keys, tc_stream stream);

/* ---- Exclusive scan: k_PartialSum, shader/partial_sum.comp:23-39 with the
 *      fix-up of src/main.rs:127-129; host wrapper partial_sum(),
 *      src/partial_sum.rs:43.  out has n+1 entries; out[n] = total.
 *      With d_n_keys the element count is (*d_n_keys - 1).  out must not
 *      overlap in (TC_ERR_INVALID). */
int tc_exclusive_scan_u32(tc_ctx* ctx, const uint32_t* in, uint32_t* out, uint32_t n,
                          const uint32_t* d_n_keys, tc_stream stream);
/* Edge count and exclusive scan in one pass (writes both arrays). */
int tc_edge_count_scan(tc_ctx* ctx, const uint8_t* prefix_n, const int32_t* parent,
                       uint32_t* edge_count, uint32_t* offsets, uint32_t n_brt,
                       const uint32_t* d_n_keys, tc_stream stream);

/* ---- Octree: k_MakeOctNodes + k_LinkLeafNodes, shader/octree.comp:126-236,
 *      root set-up src/main.rs:140-153; host wrapper build_octree(),
 *      src/octree.rs:73-87.  nodes must hold offsets[n_brt] entries
 *      (nodes_capacity is checked on the device; nodes beyond it are not
 *      written).  Leaves are linked when link_leaves != 0. */
int tc_build_octree_u32(tc_ctx* ctx, tc_oct_node* nodes, uint32_t nodes_capacity,
                        const uint32_t* offsets, const uint32_t* edge_count, const uint32_t* codes,
                        const uint8_t* prefix_n, const int32_t* parent, const int32_t* left_child,
                        const uint8_t* has_leaf_left, const uint8_t* has_leaf_right, uint32_t n_brt,
                        const uint32_t* d_n_keys, float min_coord, float range, int link_leaves,
                        tc_stream stream);
int tc_build_octree_u64(tc_ctx* ctx, tc_oct_node* nodes, uint32_t nodes_capacity,
                        const uint32_t* offsets, const uint32_t* edge_count, const uint64_t* codes,
                        const uint8_t* prefix_n, const int32_t* parent, const int32_t* left_child,
                        const uint8_t* has_leaf_left, const uint8_t* has_leaf_right, uint32_t n_brt,
                        const uint32_t* d_n_keys, float min_coord, float range, int link_leaves,
                        tc_stream stream);
/* Upper bound on the node count for n_keys unique codes (1 + sum over levels of
 * min(8^L, n_keys/2)); the pipeline sizes its node array with it. */
uint64_t tc_octree_node_bound(uint64_t n_keys, int code_bits);

/* ---- Whole plain pipeline (the stage order of src/main.rs:60-170) on
 *      device-resident buffers, queued without any host synchronisation. */
typedef struct {
    uint32_t capacity;        /* max points */
    int code_bits;            /* 30 (u32 codes) or 63 (u64 codes) */
    uint64_t node_capacity;   /* octree node capacity */
    void* sorted;             /* [capacity] Morton codes, sorted in place */
    void* alt;                /* [capacity] sort ping-pong space */
    void* unique;             /* [capacity] unique codes */
    uint8_t* prefix_n;        /* [capacity] radix tree */
    uint8_t* has_leaf_left;
    uint8_t* has_leaf_right;
    int32_t* left_child;
    int32_t* parent;
    uint32_t* edge_count;     /* [capacity] */
    uint32_t* offsets;        /* [capacity+1] */
    tc_oct_node* nodes;       /* [node_capacity] */
    uint32_t* d_counts;       /* device [4], written by every tc_build_tree on the
                                 device (no host sync): n_unique, n_brt, n_nodes;
                                 word 3 is reserved (not written) */
    /* filled by tc_tree_read_counts */
    uint32_t n_points, n_unique, n_brt, n_nodes;
    /* Point -> leaf map (NULL unless allocated with TC_TREE_POINT_INDEX).  No
     * reference counterpart: the reference sorts codes only (src/main.rs:85).
     * Leaf j (unique code j) holds the points point_index[leaf_first[j] ..
     * leaf_first[j+1]); point_index is the stable sort permutation (sorted
     * position -> input index) and leaf_first[n_unique] = n_points. */
    uint32_t* point_index;    /* [capacity] */
    uint32_t* index_alt;      /* [capacity] sort ping-pong space for the index */
    uint32_t* leaf_first;     /* [capacity+1] */
    /* the last build's parameters (set by tc_build_tree) */
    float min_coord, range;
    int link_leaves;
} tc_tree;

#define TC_TREE_POINT_INDEX 1 /* tc_tree_alloc_ex flag: also build the point -> leaf map */

int tc_tree_alloc(tc_ctx* ctx, uint32_t capacity, int code_bits, tc_tree* tree);
int tc_tree_alloc_ex(tc_ctx* ctx, uint32_t capacity, int code_bits, int flags, tc_tree* tree);
int tc_tree_free(tc_ctx* ctx, tc_tree* tree);
int tc_build_tree(tc_ctx* ctx, const float* pts_xyzw, uint32_t n, float min_coord, float range,
                  int link_leaves, tc_tree* tree, tc_stream stream);
/* Synchronise the stream and copy the device counts into the tree struct. */
int tc_tree_read_counts(tc_ctx* ctx, tc_tree* tree, tc_stream stream);

/* ---- Tree export and validation (host side).  No reference counterpart
 *      beyond the commented-out little-endian dump of the sorted keys at
 *      src/main.rs:761-768.
 *
 * tc_tree_host describes one tree in host memory (the arrays of tc_tree after
 * a download; optional arrays may be NULL).  Node and radix arrays have the
 * layouts documented above. */
typedef struct {
    int code_bits;               /* 30 or 63 */
    float min_coord, range;
    int link_leaves;             /* leaves were linked into the octree */
    uint32_t n_points, n_unique, n_nodes;
    const void* sorted;          /* [n_points] sorted codes (optional) */
    const void* unique;          /* [n_unique] */
    const uint8_t* prefix_n;     /* [n_unique-1] each */
    const uint8_t* has_leaf_left;
    const uint8_t* has_leaf_right;
    const int32_t* left_child;
    const int32_t* parent;
    const uint32_t* edge_count;  /* [n_unique-1] */
    const uint32_t* offsets;     /* [n_unique] */
    const tc_oct_node* nodes;    /* [n_nodes] */
    const uint32_t* point_index; /* [n_points] (optional) */
    const uint32_t* leaf_first;  /* [n_unique+1] (optional) */
} tc_tree_host;

/* Checks every invariant that defines the tree, independently of how it was
 * built: unique codes strictly increasing; radix tree links (children, parents,
 * each leaf and non-root node referenced once), each node's key range with the
 * node index at one end, prefix_n = common prefix of the range and the split
 * at its first differing bit; edge counts and offsets; octree child links,
 * halving cell sizes, corners equal to the decoded cell prefixes, every node
 * reachable once from the root, and (link_leaves) every unique code found by
 * descending from the root into its leaf slot; the point -> leaf map when
 * present.  Returns TC_OK or TC_ERR_BAD_TREE with the first violation
 * written to msg (NUL-terminated, at most msg_len bytes; msg may be NULL). */
int tc_tree_validate_host(const tc_tree_host* tree, char* msg, size_t msg_len);
/* Writes the tree to a file: 8-byte magic "TCTREE\0\1", a u64 header
 * length, a JSON header (sizes, parameters and for each array its name,
 * little-endian dtype, count and byte offset), then the arrays, each 64-byte
 * aligned.  Returns TC_OK or TC_ERR_IO. */
int tc_tree_write_host(const tc_tree_host* tree, const char* path);
/* The same on a device tree of the last tc_build_tree (synchronises the
 * stream and copies the arrays to host memory). */
int tc_tree_validate(tc_ctx* ctx, tc_tree* tree, char* msg, size_t msg_len, tc_stream stream);
int tc_tree_export(tc_ctx* ctx, tc_tree* tree, const char* path, tc_stream stream);

/* ---- Batched builds (no reference counterpart: the reference builds one cloud
 *      per call, src/main.rs:60-170).  Many clouds of uneven sizes, concatenated
 *      in one point array, are built with one persistent segmented launch per
 *      stage; each cloud's outputs equal a plain tc_build_tree of its points.
 *      Every kernel reads the batch shape from device memory, so a captured
 *      graph (tc_batch_capture) replays for any batch within the capacity after
 *      tc_batch_set uploads the new plan. */
typedef struct tc_batch tc_batch;

typedef struct {
    uint32_t point_offset, n_points;   /* cloud's points in the concatenated input / sorted codes */
    uint32_t unique_offset, n_unique;  /* cloud's unique codes and radix-node slots (n_unique - 1 nodes) */
    uint32_t node_offset, n_nodes;     /* cloud's octree nodes; stored links are cloud-local */
} tc_cloud_info;

typedef struct {
    int code_bits;
    uint32_t n_clouds, n_points;
    void* sorted;            /* [n_points] codes, sorted within each cloud */
    void* unique;            /* unique codes, cloud c at unique_offset */
    uint8_t* prefix_n;       /* radix arrays, cloud c at unique_offset (cloud-local indices) */
    uint8_t* has_leaf_left;
    uint8_t* has_leaf_right;
    int32_t* left_child;
    int32_t* parent;
    uint32_t* edge_count;    /* cloud c at unique_offset */
    uint32_t* offsets;       /* global node offsets; cloud-local = offsets - node_offset */
    tc_oct_node* nodes;      /* cloud c at node_offset */
    uint64_t node_capacity;
} tc_batch_view;

/* Per-cloud output buffers (device), sized for a cloud of n points: unique
 * [n] codes, radix arrays [n-1], edge_count [n-1], offsets [n],
 * nodes [tc_octree_node_bound(n)].  Filled with the same values as a plain
 * build (cloud-local indices and offsets). */
typedef struct {
    void* unique;
    uint8_t* prefix_n;
    uint8_t* has_leaf_left;
    uint8_t* has_leaf_right;
    int32_t* left_child;
    int32_t* parent;
    uint32_t* edge_count;
    uint32_t* offsets;
    tc_oct_node* nodes;
} tc_tree_ptrs;

int tc_batch_create(tc_ctx* ctx, uint32_t max_clouds, uint32_t max_points, int code_bits, tc_batch** out);
int tc_batch_destroy(tc_batch* batch);
/* Describe the next batch: per-cloud point counts (and min/range, NULL for
 * 0 / 1024); uploads the plan asynchronously on the stream. */
int tc_batch_set(tc_batch* batch, uint32_t n_clouds, const uint32_t* counts, const float* mins,
                 const float* ranges, tc_stream stream);
/* As tc_batch_set, for clouds held in separate device buffers: cloud c is read
 * from cloud_pts[c] (16-byte aligned xyzw), and when outs is non-NULL the
 * kernels that produce each output also store cloud c's copy in outs[c] (its
 * octree nodes only there; the other concatenated arrays stay valid).  Build
 * or capture with pts_xyzw = NULL after this call. */
int tc_batch_set_io(tc_batch* batch, uint32_t n_clouds, const uint32_t* counts, const float* mins,
                    const float* ranges, const float* const* cloud_pts, const tc_tree_ptrs* outs,
                    tc_stream stream);
int tc_batch_build(tc_batch* batch, const float* pts_xyzw, int link_leaves, tc_stream stream);
/* Capture the whole batched pipeline (for pts_xyzw) as a HIP graph; replay it
 * after each tc_batch_set.  The stream must not be the legacy NULL stream. */
int tc_batch_capture(tc_batch* batch, const float* pts_xyzw, int link_leaves, tc_stream stream);
int tc_batch_replay(tc_batch* batch, tc_stream stream);
/* Synchronise the stream and read per-cloud offsets and counts. */
int tc_batch_read_info(tc_batch* batch, tc_cloud_info* info, tc_stream stream);
int tc_batch_arrays(tc_batch* batch, tc_batch_view* view);
/* Queue copies of the per-cloud bases (n_clouds + 1 entries each: cloud c has
 * h_unique_base[c+1] - h_unique_base[c] unique codes and
 * h_node_base[c+1] - h_node_base[c] octree nodes) into pinned host memory;
 * they are valid once the stream has reached this point. */
int tc_batch_read_info_async(tc_batch* batch, uint32_t* h_unique_base, uint32_t* h_node_base, tc_stream stream);
/* Device bytes tc_batch_create allocates for these capacities. */
uint64_t tc_batch_memory_bytes(uint32_t max_clouds, uint32_t max_points, int code_bits);
/* Bytes of the per-cloud output buffers (tc_tree_ptrs) for a cloud of n points,
 * each array 256-byte aligned, nodes last, at the node bound. */
uint64_t tc_tree_result_bytes(uint32_t n, int code_bits);

/* ---- Serving (no reference counterpart).  Requests (device point buffers)
 *      are admitted as they arrive while a device-memory budget allows, and
 *      built in batches on several streams: whenever a stream is free, the
 *      requests waiting at that moment (in arrival order, within the batch
 *      limits and the budget) go out as one segmented build, so work that
 *      arrives while other batches run overlaps them on the device.  Each
 *      result stays in device memory, equal to a plain build of the same
 *      points, until tc_server_release.  Submission and polling are
 *      thread-safe; a scheduler thread owned by the server does the rest. */
typedef struct tc_server tc_server;

typedef struct {
    uint64_t mem_budget_bytes;  /* device bytes for workspaces + held results (default 64 GiB) */
    uint32_t n_streams;         /* concurrent batches (default 2) */
    uint32_t max_batch_clouds;  /* per batch (default 256) */
    uint32_t max_batch_points;  /* per batch, also the largest request (default 2^24) */
    int code_bits;              /* 30 or 63 (default 30) */
    int link_leaves;            /* default 1 */
    uint32_t max_queued;        /* requests waiting for launch before submit returns TC_ERR_BUSY
                                   (default 65536; 0: unbounded) */
} tc_server_config;

typedef struct {
    uint64_t id;
    int32_t status;                      /* TC_OK */
    uint32_t n_points, n_unique, n_nodes;
    tc_tree_ptrs tree;                   /* device arrays, valid until tc_server_release(id) */
    double t_submit, t_start, t_done;    /* seconds on the server's steady clock */
} tc_server_result;

typedef struct {
    uint64_t submitted, completed, rejected, batches;
    uint64_t points_completed;
    double mean_clouds_per_batch;
    double latency_p50_ms, latency_p99_ms, latency_mean_ms, latency_max_ms;
    double first_submit_s, last_done_s;  /* throughput = points_completed / (last_done - first_submit) */
    uint64_t reserved_bytes, peak_reserved_bytes, result_budget_bytes;
    uint64_t cancelled;  /* requests cancelled before they started (counted as completed) */
    uint64_t queued;     /* requests waiting for launch now */
} tc_server_stats;

/* Fills cfg with the defaults. */
void tc_server_config_default(tc_server_config* cfg);
int tc_server_create(tc_ctx* ctx, const tc_server_config* cfg, tc_server** out);
/* Waits for the batches on the device, drops requests not yet launched, then
 * frees everything (held results included). */
int tc_server_destroy(tc_server* srv);
/* Queue a build of n points (device, 16-byte aligned xyzw; must stay valid
 * until the result is polled).  TC_ERR_CAPACITY: n above max_batch_points;
 * TC_ERR_OVER_BUDGET: its result could never fit the budget; TC_ERR_BUSY:
 * max_queued requests are already waiting (back-pressure: retry later). */
int tc_server_submit(tc_server* srv, const float* pts_xyzw, uint32_t n, float min_coord, float range,
                     uint64_t* id);
/* Cancel a request that has not started: it is dropped and comes back from
 * tc_server_poll with status TC_ERR_CANCELLED and no arrays (its points need
 * not stay valid after this returns TC_OK).  TC_ERR_BUSY: it is being built
 * (poll for its result); TC_ERR_NOT_FOUND: unknown or already finished. */
int tc_server_cancel(tc_server* srv, uint64_t id);
/* Move up to max finished results to out (wait up to wait_ms for the first). */
int tc_server_poll(tc_server* srv, tc_server_result* out, uint32_t max, uint32_t* n_out, int wait_ms);
/* Return a polled result's device memory to the budget. */
int tc_server_release(tc_server* srv, uint64_t id);
/* Block until every submitted request has finished (requests waiting for
 * memory need earlier results to be released meanwhile). */
int tc_server_drain(tc_server* srv);
int tc_server_stats_get(tc_server* srv, tc_server_stats* st);
int tc_server_stats_reset(tc_server* srv);

/* Replay an arrival trace against the server from a submitting thread:
 * request i (points pts[i], count n[i]) is submitted at t_arrive[i] seconds
 * after the call starts; results are released as they are polled.  latency_s
 * receives done - arrival per request. */
int tc_server_run_trace(tc_server* srv, uint32_t n_req, const double* t_arrive, const float* const* pts,
                        const uint32_t* n, float min_coord, float range, double* latency_s);
/* The same trace through a one-request-at-a-time plain loop (tc_build_tree,
 * each request waiting for the previous one): the baseline for the server. */
int tc_plain_run_trace(tc_ctx* ctx, uint32_t n_req, const double* t_arrive, const float* const* pts,
                       const uint32_t* n, float min_coord, float range, int code_bits, double* latency_s);

/* Discrete-event simulation of the server's admission and batching policy
 * (the same scheduler code, a modelled device: a batch of P points takes
 * launch_s + per_point_s * P, results are released hold_s after they finish).
 * For testing the policy without a GPU. */
typedef struct {
    uint32_t n_streams, max_batch_clouds, max_batch_points;
    uint64_t result_budget_bytes;
    int code_bits;
    double launch_s, per_point_s, hold_s;
    /* admission back-pressure, as tc_server_config.max_queued (0: unbounded) */
    uint32_t max_queued;
    /* failure injection into the result allocator (an arena of
     * result_budget_bytes, as in the server): after alloc_fail_after successful
     * allocations the next alloc_fail_count attempts fail (0xFFFFFFFF: all of
     * them); a failed placement ends the batch and the request waits. */
    uint32_t alloc_fail_after, alloc_fail_count;
} tc_sim_config;
/* Outputs per request: t_start / t_done (-1 if it never ran), batch_of
 * (0xFFFFFFFF if none); status (optional): TC_OK, TC_ERR_BUSY (rejected at
 * admission), TC_ERR_CAPACITY / TC_ERR_OVER_BUDGET (could never be served),
 * TC_ERR_CANCELLED (t_cancel reached while it waited) or TC_ERR_NOMEM (still
 * waiting when no event could free memory: the call then returns TC_ERR_NOMEM).
 * t_cancel (optional): per request, the time its client cancels it (+inf:
 * never). */
int tc_server_simulate(const tc_sim_config* cfg, uint32_t n_req, const double* t_arrive, const uint32_t* n,
                       double* t_start, double* t_done, uint32_t* batch_of, uint64_t* peak_reserved);
int tc_server_simulate_ex(const tc_sim_config* cfg, uint32_t n_req, const double* t_arrive, const uint32_t* n,
                          const double* t_cancel, double* t_start, double* t_done, uint32_t* batch_of,
                          int32_t* status, uint64_t* peak_reserved);

/* ---- Host-array mirrors of the reference's per-stage functions (blocking).
 * Argument order and meaning follow the Rust functions cited; sizes are
 * explicit and results are written to caller arrays. */
int tc_host_compute_morton(tc_ctx* ctx, const float* input_xyzw, uint32_t* morton_keys, uint32_t n,
                           float min_coord, float range, uint32_t group_size);
int tc_host_histogram(tc_ctx* ctx, uint32_t input_size, const uint32_t* input_data,
                      uint32_t* global_hist);
int tc_host_radix_sort(tc_ctx* ctx, const uint32_t* in_data, uint32_t n, uint32_t* out_data,
                       uint32_t group_size);
int tc_host_unique(tc_ctx* ctx, uint32_t* u_sort, uint32_t n, uint32_t* n_unique);
int tc_host_build_radix_tree(tc_ctx* ctx, int32_t n, const uint32_t* codes, uint8_t* prefix_n,
                             uint8_t* has_leaf_left, uint8_t* has_leaf_right, int32_t* left_child,
                             int32_t* parent, uint32_t group_size);
int tc_host_edge_count(tc_ctx* ctx, const uint8_t* prefix_n, const int32_t* parent,
                       uint32_t* u_edge_count, uint32_t n_brt_nodes, uint32_t group_size);
int tc_host_partial_sum(tc_ctx* ctx, const uint32_t* u_edge_count, uint32_t first, uint32_t last,
                        uint32_t* d_first);
int tc_host_build_octree(tc_ctx* ctx, tc_oct_node* oct_nodes, const uint32_t* node_offsets,
                         const uint32_t* rt_node_counts, const uint32_t* codes,
                         const uint8_t* rt_prefix_n, const int32_t* rt_parents, float min_coord,
                         float range, int32_t n, const uint8_t* rt_has_leaf_left,
                         const uint8_t* rt_has_leaf_right, const int32_t* rt_left_child,
                         uint32_t group_size);

/* ---- Seeded input generators (host only, multi-threaded, no GPU needed).
 * Counter-based SplitMix64 streams: the value for point i depends only on
 * (seed, i), so any thread count gives identical arrays. */
void tc_gen_uniform(float* pts_xyzw, uint64_t n, uint64_t seed, float extent);
void tc_gen_gaussian_mixture(float* pts_xyzw, uint64_t n, uint64_t seed, uint32_t n_centres,
                             float sigma, float extent);

#ifdef __cplusplus
}
#endif
#endif
