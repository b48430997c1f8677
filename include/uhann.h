/*
 * uhann.h -- C ABI of the MI355X query engine for the exact, flat and
 * hierarchical nearest-neighbour indexes of martingherold/ultrahigh_ann.
 *
 * Plain C: opaque handles, plain pointers and sizes, an int status per call
 * and a thread-local message (uhann_last_error). Status codes map one to one
 * onto the exception types the reference throws, so a C++ facade can rethrow
 * the same type. No call allocates host memory the caller must free.
 *
 * The reference has no C ABI: its swap point is the set of seven CUDA
 * translation units that CMake substitutes for "CUDA not enabled" stubs
 * (CMakeLists.txt:20-51). Each entry point below names the reference
 * interface it replaces (paths relative to the reference checkout).
 *
 * Threading: handles are immutable after creation and may be shared between
 * threads; a workspace must be used by one thread at a time
 * (hierarchical_l1_ann_index.hpp:18-19). Synchronous calls return with their
 * results in host memory. Indices are uint64_t (std::size_t on LP64).
 */
#ifndef UHANN_H
#define UHANN_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes (reference exception type in brackets) ---------------- */
#define UHANN_OK 0
#define UHANN_ERR_INVALID_ARGUMENT 1 /* std::invalid_argument */
#define UHANN_ERR_LENGTH 2           /* std::length_error */
#define UHANN_ERR_RUNTIME 3          /* std::runtime_error: device errors, no device */
#define UHANN_ERR_OVERFLOW 4         /* std::overflow_error: non-finite results */
#define UHANN_ERR_DOMAIN 5           /* std::domain_error: invalid probabilities */
#define UHANN_ERR_OUT_OF_RANGE 6     /* std::out_of_range */
#define UHANN_ERR_LOGIC 7            /* std::logic_error */

#define UHANN_METRIC_L2 0
#define UHANN_METRIC_L1 1

/* Query strategies. DIRECT is the plain binary32 path, bit-identical to the
 * reference's direct CUDA kernels (CudaDenseL2QueryStrategy::direct,
 * cuda_dense_l2_scan_index.hpp:17-23). */
#define UHANN_STRATEGY_DIRECT 0
/* L2 only. ||q||^2 + ||v||^2 - 2 q.v with the inner product in BF16 on MFMA
 * (FP32 accumulation), a per-query candidate set bounded by the BF16 error,
 * an exact re-rank of the candidates in the direct path's binary32 order and
 * a binary64 re-rank of binary32 near ties. New (the reference's GEMM
 * strategy, cuda_dense_l2_scan_index.hpp:21-22, is unrefined FP32 SGEMM). */
#define UHANN_STRATEGY_GRAM_BF16 1

/* Index kinds reported by uhann_index_info. */
#define UHANN_KIND_EXACT 0
#define UHANN_KIND_FLAT 1
#define UHANN_KIND_HIERARCHICAL 2

/* Message of the last failed call on this thread ("" after a success). */
const char* uhann_last_error(void);

/* 1 when a HIP device is visible, else 0. Replaces cuda_dense_l2_scan_available()
 * (cuda_dense_l2_scan_index.hpp:15) and the per-class cuda_*_available(). */
int uhann_device_available(void);

/* Device name into buf (NUL-terminated, truncated to cap). Replaces
 * CudaDenseL2ScanIndex::device_name() (cuda_dense_l2_scan_index.hpp:87). */
int uhann_device_name(int device, char* buf, size_t cap);

/* ==== dense scan engine ==================================================
 * Device-resident reference rows, scanned in binary32 with a fused
 * lowest-index argmin. */
typedef struct uhann_scan uhann_scan;
typedef struct uhann_workspace uhann_workspace;

/* Copies rows x dim row-major references to `device`. metric selects L2 or
 * L1; weights (L1 only, dim positive finite floats, or NULL) scale |r-q|.
 * Replaces CudaDenseL2ScanIndex(const DenseMatrix&, int device)
 * (cuda_dense_l2_scan_index.hpp:47-48) and CudaDenseL1ScanIndex(reps[,
 * weights], device) (cuda_dense_l1_scan_index.hpp:40-45). Errors: rows == 0
 * or non-finite values or bad weights -> INVALID_ARGUMENT; no device ->
 * RUNTIME. */
int uhann_scan_create(int metric, const float* host_refs, uint64_t rows, uint64_t dim,
                      const float* weights, int device, uhann_scan** out);
void uhann_scan_destroy(uhann_scan* scan);

/* Shape and device-payload bytes. Replaces space_usage()/device()
 * (cuda_dense_l2_scan_index.hpp:85-86). */
int uhann_scan_info(const uhann_scan* scan, uint64_t* rows, uint64_t* dim,
                    uint64_t* payload_bytes, int* device);

/* Per-caller buffers for up to max_batch queries. Replaces
 * make_query_workspace(max_batch) (cuda_dense_l2_scan_index.hpp:58-59). */
int uhann_scan_workspace_create(const uhann_scan* scan, uint64_t max_batch,
                                uhann_workspace** out);
void uhann_workspace_destroy(uhann_workspace* ws);
int uhann_workspace_info(const uhann_workspace* ws, uint64_t* max_batch,
                         uint64_t* payload_bytes);

/* count x dim host queries; out_idx gets count row indices and out_dist (or
 * NULL) the binary32 distances. Synchronous. Validation order as the
 * reference: workspace owner, capacity, finite queries, count == 0 returns.
 * Replaces query_batch (cuda_dense_l2_scan_index.hpp:66-70). */
int uhann_scan_query_batch(const uhann_scan* scan, uhann_workspace* ws, const float* host_queries,
                           uint64_t count, uint64_t* out_idx, float* out_dist, int strategy);

/* Same with count x dim queries already on the scan's device (row stride
 * dim). Results to host memory; synchronous. Replaces query_device_batch
 * (cuda_dense_l2_scan_index.hpp:76-82). */
int uhann_scan_query_device_batch(const uhann_scan* scan, uhann_workspace* ws,
                                  const float* device_queries, uint64_t count, uint64_t* out_idx,
                                  float* out_dist, int strategy);

/* Asynchronous device-to-device form on a caller HIP stream (hipStream_t
 * passed as void*; NULL = default stream). Queries have row stride
 * ld_queries >= dim; device_out_idx / device_out_dist (either may be NULL)
 * receive the results when the stream reaches them. No host
 * synchronisation. No counterpart in the reference (its calls are
 * synchronous on the legacy stream): for callers that batch or capture. */
int uhann_scan_query_device_async(const uhann_scan* scan, uhann_workspace* ws,
                                  const float* device_queries, uint64_t ld_queries,
                                  uint64_t count, uint64_t* device_out_idx,
                                  float* device_out_dist, int strategy, void* stream);

/* As uhann_scan_query_device_async with an explicit register tile: rpt
 * chains per thread (1, 2 or 4), tq x tr (query, reference) pairs per block,
 * group_q query tiles per reference sweep. Returns INVALID_ARGUMENT for a tile
 * that is not compiled in. No counterpart in the reference: used by the
 * autotuner and by the tests that cover every tile. */
int uhann_scan_query_device_tile(const uhann_scan* scan, uhann_workspace* ws,
                                 const float* device_queries, uint64_t ld_queries, uint64_t count,
                                 uint64_t* device_out_idx, float* device_out_dist, int rpt, int tq,
                                 int tr, int group_q, void* stream);

/* Builds the BF16 reference copy and row statistics the GRAM_BF16 strategy
 * needs (rows x dim bf16 + 16 bytes per row of device memory). Call before
 * sharing the handle between threads. L2 only. No counterpart in the
 * reference (its GEMM strategy needs no preparation beyond K1 norms,
 * cuda_dense_l2_scan_index.cu:449-462). */
int uhann_scan_enable_mixed(uhann_scan* scan, int strategy);

/* Counters of the last GRAM_BF16 batch on this workspace: total candidate
 * set size, queries decided in binary64, queries escalated to the direct
 * scan. Synchronises the workspace stream. */
int uhann_workspace_mixed_stats(uhann_workspace* ws, uint64_t* candidates, uint64_t* fp64_reranked,
                                uint64_t* escalated);

/* Stage timing of GRAM_BF16 batches on this workspace with HIP events on the
 * launch stream (enable != 0 turns it on; each timed batch then waits for its
 * last event). uhann_workspace_stage_times copies up to n stage durations in
 * milliseconds: bf16 conversion, Gram GEMM, select + re-rank, escalation. */
int uhann_workspace_set_timing(uhann_workspace* ws, int enable);
int uhann_workspace_stage_times(const uhann_workspace* ws, float* ms, int n);

#ifdef __cplusplus
}
#endif

#endif /* UHANN_H */
