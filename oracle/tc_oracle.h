/*
 * tc_oracle.h -- CPU oracle for the tree-construction path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP
 * implementation is compared against; only tests, the smoke check and the
 * CPU-baseline leg of the benchmark may load it.  Product code never links,
 * loads or calls it.
 *
 * It restates, in plain sequential C, the semantics of the reference's
 * GLSL compute shaders (ucsc-redwood/glsl-tree-construction, paths relative
 * to vulkan/radixsort-vulkan/), with the normalisations documented beside
 * each function (undefined or racy reference behaviour replaced by one
 * well-defined rule that the GPU path implements identically).
 */
#ifndef TC_ORACLE_H
#define TC_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* 64-byte octree node, same field order as OctNode in shader/octree.comp:14-20
 * (std430 stride 64 B; the explicit pad makes the host stride match). */
typedef struct {
    int32_t children[8];
    float corner[4];
    float cell_size;
    int32_t child_node_mask;
    int32_t child_leaf_mask;
    int32_t pad_;
} tco_oct_node;

/* ---- Morton encode / decode (shader/morton.comp:36-56, src/morton.rs:43-74) */
uint32_t tco_quantize(float v, float min_coord, float range, float bit_scale);
uint32_t tco_morton30_point(float x, float y, float z, float min_coord, float range);
uint64_t tco_morton63_point(float x, float y, float z, float min_coord, float range);
void tco_morton30(const float* pts_xyzw, uint32_t* codes, size_t n, float min_coord, float range);
void tco_morton63(const float* pts_xyzw, uint64_t* codes, size_t n, float min_coord, float range);
void tco_morton30_decode(uint32_t code, float min_coord, float range, float out[4]);
void tco_morton63_decode(uint64_t code, float min_coord, float range, float out[4]);

/* ---- Global histogram (shader/histogram.comp:103-195) */
void tco_global_histogram_u32(const uint32_t* keys, size_t n, uint32_t* hist_4x256);
void tco_global_histogram_u64(const uint64_t* keys, size_t n, uint32_t* hist_8x256);

/* ---- One stable digit-binning pass (shader/new_radix_sort.comp:116-314) */
void tco_binning_pass_u32(const uint32_t* src_k, uint32_t* dst_k, const uint32_t* src_v, uint32_t* dst_v,
                          size_t n, const uint32_t* digit_prefix_256, int radix_shift);
void tco_binning_pass_u64(const uint64_t* src_k, uint64_t* dst_k, const uint32_t* src_v, uint32_t* dst_v,
                          size_t n, const uint32_t* digit_prefix_256, int radix_shift);

/* ---- Full LSD sorts (8-bit digits; histogram + ping-pong binning passes).
 * values may be NULL (keys only).  tmp buffers must hold n elements. */
void tco_sort_u32(uint32_t* keys, uint32_t* values, size_t n, uint32_t* tmp_k, uint32_t* tmp_v);
void tco_sort_u64(uint64_t* keys, uint32_t* values, size_t n, uint64_t* tmp_k, uint32_t* tmp_v);

/* ---- Unique (shader/unique.comp:31-45): out[0..U) = first of each run. */
uint32_t tco_unique_u32(const uint32_t* sorted, uint32_t* out, size_t n);
uint32_t tco_unique_u64(const uint64_t* sorted, uint64_t* out, size_t n);
uint32_t tco_unique_runs_u32(const uint32_t* sorted, uint32_t* out, uint32_t* first, size_t n);
uint32_t tco_unique_runs_u64(const uint64_t* sorted, uint64_t* out, uint32_t* first, size_t n);

/* ---- Karras binary radix tree (shader/build_radix_tree.comp:44-153).
 * n_keys unique sorted codes -> n_keys-1 internal nodes. code_bits = 30 or 63. */
void tco_build_radix_tree_u32(const uint32_t* codes, uint32_t n_keys, uint8_t* prefix_n,
                              uint8_t* has_leaf_left, uint8_t* has_leaf_right,
                              int32_t* left_child, int32_t* parent);
void tco_build_radix_tree_u64(const uint64_t* codes, uint32_t n_keys, uint8_t* prefix_n,
                              uint8_t* has_leaf_left, uint8_t* has_leaf_right,
                              int32_t* left_child, int32_t* parent);

/* Nodes [i0, i1) only (parent[0] is not written); ranges may run on separate threads. */
void tco_build_radix_tree_range_u32(const uint32_t* codes, uint32_t n_keys, uint32_t i0, uint32_t i1,
                                    uint8_t* prefix_n, uint8_t* has_leaf_left, uint8_t* has_leaf_right,
                                    int32_t* left_child, int32_t* parent);
void tco_build_radix_tree_range_u64(const uint64_t* codes, uint32_t n_keys, uint32_t i0, uint32_t i1,
                                    uint8_t* prefix_n, uint8_t* has_leaf_left, uint8_t* has_leaf_right,
                                    int32_t* left_child, int32_t* parent);

/* ---- Edge count (shader/edge_count.comp:30-41) with the root slot = 1. */
void tco_edge_count(const uint8_t* prefix_n, const int32_t* parent, uint32_t* edge_count,
                    uint32_t n_brt, int code_bits);

/* ---- Exclusive scan: out[0..n] with out[n] = total (shader/partial_sum.comp:23-39). */
uint32_t tco_exclusive_scan_u32(const uint32_t* in, uint32_t* out, size_t n);

/* ---- Octree (shader/octree.comp:126-222, root set-up src/main.rs:140-153).
 * nodes must hold offsets[n_brt] entries.  Returns the node count. */
uint32_t tco_build_octree_u32(tco_oct_node* nodes, const uint32_t* offsets, const uint32_t* edge_count,
                              const uint32_t* codes, const uint8_t* prefix_n, const int32_t* parent,
                              const int32_t* left_child, const uint8_t* has_leaf_left,
                              const uint8_t* has_leaf_right, uint32_t n_brt, float min_coord,
                              float range, int link_leaves);
uint32_t tco_build_octree_u64(tco_oct_node* nodes, const uint32_t* offsets, const uint32_t* edge_count,
                              const uint64_t* codes, const uint8_t* prefix_n, const int32_t* parent,
                              const int32_t* left_child, const uint8_t* has_leaf_left,
                              const uint8_t* has_leaf_right, uint32_t n_brt, float min_coord,
                              float range, int link_leaves);

/* Radix nodes [i0, i1) only, into nodes zeroed beforehand (offsets[n_brt] entries). */
void tco_build_octree_range_u32(tco_oct_node* nodes, const uint32_t* offsets, const uint32_t* edge_count,
                                const uint32_t* codes, const uint8_t* prefix_n, const int32_t* parent,
                                const int32_t* left_child, const uint8_t* has_leaf_left,
                                const uint8_t* has_leaf_right, uint32_t i0, uint32_t i1, float min_coord,
                                float range, int link_leaves);
void tco_build_octree_range_u64(tco_oct_node* nodes, const uint32_t* offsets, const uint32_t* edge_count,
                                const uint64_t* codes, const uint8_t* prefix_n, const int32_t* parent,
                                const int32_t* left_child, const uint8_t* has_leaf_left,
                                const uint8_t* has_leaf_right, uint32_t i0, uint32_t i1, float min_coord,
                                float range, int link_leaves);

/* ---- Multi-threaded CPU build of the whole pipeline (cpu_mt.cpp): the CPU
 * baseline.  Same outputs as the functions above; nodes holds node_cap
 * entries; returns the node count (0xFFFFFFFF when above node_cap) and, in
 * stage_ms[7] (optional), the time of morton, sort, unique, radix tree, edge
 * count, scan, octree. */
uint32_t tco_mt_build_tree_u32(const float* pts_xyzw, uint32_t n, float min_coord, float range, int threads,
                               uint32_t* codes, uint32_t* sorted, uint32_t* tmp, uint32_t* unique,
                               uint32_t* n_unique, uint8_t* prefix_n, uint8_t* has_leaf_left,
                               uint8_t* has_leaf_right, int32_t* left_child, int32_t* parent,
                               uint32_t* edge_count, uint32_t* offsets, tco_oct_node* nodes, uint64_t node_cap,
                               int link_leaves, double* stage_ms);
uint32_t tco_mt_build_tree_u64(const float* pts_xyzw, uint32_t n, float min_coord, float range, int threads,
                               uint64_t* codes, uint64_t* sorted, uint64_t* tmp, uint64_t* unique,
                               uint32_t* n_unique, uint8_t* prefix_n, uint8_t* has_leaf_left,
                               uint8_t* has_leaf_right, int32_t* left_child, int32_t* parent,
                               uint32_t* edge_count, uint32_t* offsets, tco_oct_node* nodes, uint64_t node_cap,
                               int link_leaves, double* stage_ms);
/* std::sort of a copy of n codes into work; returns its milliseconds. */
double tco_std_sort_ms(const uint32_t* codes, uint32_t n, uint32_t* work);
/* Key-value sorts of n u64 keys with u32 values, in place, returning their
 * milliseconds: the stable LSD radix sort on `threads` threads, and
 * std::stable_sort of the pairs. */
double tco_mt_sort_pairs_u64_ms(uint64_t* keys, uint32_t* vals, uint32_t n, int threads);
double tco_std_stable_sort_pairs_u64_ms(uint64_t* keys, uint32_t* vals, uint32_t n);

/* Upper bound on the octree node count for n_keys unique codes of code_bits
 * bits (1 + sum over levels of min(8^L, n_keys/2)). */
uint64_t tco_octree_node_bound(uint64_t n_keys, int code_bits);

#ifdef __cplusplus
}
#endif
#endif
