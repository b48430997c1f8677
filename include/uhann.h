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
 * (FP32 accumulation), a per-query candidate set, an exact re-rank of the
 * candidates in the direct path's binary32 order and a binary64 re-rank of
 * binary32 near ties. The candidate bound is statistical (it models the
 * bf16 rounding errors as independent): not a certificate; inputs with many
 * repeated values can exceed it. GRAM_INT8 is the certified form. New (the
 * reference's GEMM strategy, cuda_dense_l2_scan_index.hpp:21-22, is
 * unrefined FP32 SGEMM). */
#define UHANN_STRATEGY_GRAM_BF16 1
/* Hierarchical L2 only. The |I| x p query projection runs as a GEMM in BF16
 * on MFMA (FP32 accumulation) instead of the reference's binary32 K7 order
 * (cuda_float_projection.cu:234-263); the scan over the projected vectors is
 * the direct binary32 scan. New: the reference projects in binary32 only. */
#define UHANN_STRATEGY_PROJECTION_BF16 2
/* L2 only. ||q||^2 + ||v||^2 - 2 s_q s_v q8.v8 with rows quantised to e4m3
 * under a power-of-two per-row scale, inner products on FP8 MFMA (FP32
 * accumulation); the 16 smallest estimates per query are re-ranked exactly in
 * the direct path's binary32 order and binary32 near ties in binary64. The
 * shortlist is not certified: uhann_workspace_shortlist_stats counts the
 * queries whose answer the FP8 error bound proves. New (the reference has no
 * FP8 path). */
#define UHANN_STRATEGY_GRAM_FP8 3
/* L2 only, dim <= 65536. Rows as 16-bit fixed point x^ = s (a + b / 256)
 * (per-row binary32 scale s, int8 planes a and b); <q^, v^> from three int8
 * MFMA products summed exactly in int32, combined in binary64. The candidate
 * set comes from a deterministic bound (Cauchy-Schwarz on each row's actual
 * quantisation error, rounded up), so it holds the direct path's answer for
 * every input; candidates are re-ranked in the direct path's binary32 order
 * and binary32 near ties in binary64; more than 64 candidates send the query
 * to the direct scan on the device. New (the reference has no integer
 * path). */
#define UHANN_STRATEGY_GRAM_INT8 4

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
 * synchronisation in any strategy (the mixed strategy's escalations are
 * handled on the device), so a call can be captured into a HIP graph once
 * the workspace has run one batch of that strategy (its buffers are
 * allocated on first use). No counterpart in the reference (its calls are
 * synchronous on the legacy stream): for callers that batch or capture. */
int uhann_scan_query_device_async(const uhann_scan* scan, uhann_workspace* ws,
                                  const float* device_queries, uint64_t ld_queries,
                                  uint64_t count, uint64_t* device_out_idx,
                                  float* device_out_dist, int strategy, void* stream);

/* uhann_scan_query_device_async with the DIRECT strategy for callers that
 * cannot allocate (code under HIP graph capture, tight serving loops): no
 * workspace, and no device memory is allocated or freed by the call. The
 * argmin keys live in a fixed device array of the library (65536 queries,
 * 512 KiB per GPU, held until the process ends); count > 65536 -> LENGTH.
 * Every call on one GPU shares that array: calls that overlap on one GPU
 * (other streams or threads) race, so order them on one stream. Same
 * results, bit for bit, as uhann_scan_query_device_async(DIRECT). No
 * counterpart in the reference. */
int uhann_scan_query_device_noalloc(const uhann_scan* scan, const float* device_queries, uint64_t ld_queries,
                                    uint64_t count, uint64_t* device_out_idx, float* device_out_dist,
                                    void* stream);

/* As uhann_scan_query_device_async with an explicit register tile: rpt
 * chains per thread (1 or 4), tq x tr (query, reference) pairs per block,
 * group_q query tiles per reference sweep; rpt = 2 selects the
 * shared-query tile (tq queries and group_q groups of tr reference rows per
 * block, one chain per thread, the queries staged in LDS); rpt = 0 selects the
 * pair-per-thread scan for rows of at most 4096 coordinates (and fewer than
 * 2^25 reference rows) with tq (1, 2, 4) queries per thread, or with tr = 8
 * its chain-class form with tq (4, 8) queries per thread (the default for
 * such rows is the chain-class form with tq = 4 up to 320 coordinates, tq = 2
 * above), and rpt = 0, tq = 0 the thread-per-query scan for rows of 1 to 128
 * coordinates. Returns INVALID_ARGUMENT for a tile that is not
 * compiled in or does not fit the rows. No counterpart in the reference:
 * used by the autotuner and by the tests that cover every tile. */
int uhann_scan_query_device_tile(const uhann_scan* scan, uhann_workspace* ws,
                                 const float* device_queries, uint64_t ld_queries, uint64_t count,
                                 uint64_t* device_out_idx, float* device_out_dist, int rpt, int tq,
                                 int tr, int group_q, void* stream);

/* Builds the reference copy and row statistics the GRAM_BF16 strategy (rows x
 * dim bf16 + 32 bytes per row of device memory), the GRAM_FP8 strategy
 * (rows x dim e4m3 bytes + 32 bytes per row) or the GRAM_INT8 strategy (rows
 * x 2 dim bytes + 32 bytes per row; dim > 65536 -> LENGTH) needs. Call before sharing the
 * handle between threads. L2 only. No counterpart in the
 * reference (its GEMM strategy needs no preparation beyond K1 norms,
 * cuda_dense_l2_scan_index.cu:449-462). */
int uhann_scan_enable_mixed(uhann_scan* scan, int strategy);

/* Rows the GRAM_FP8 strategy re-ranks exactly per query: the `rows` smallest
 * FP8 estimates (1 .. 16; 16 when never set). Fewer rows read fewer
 * reference rows in the re-rank; the shortlist is not certified either way
 * (uhann_workspace_shortlist_stats counts the queries it provably holds).
 * Call before sharing the handle between threads. No counterpart in the
 * reference (its unrefined SGEMM keeps no shortlist). */
int uhann_scan_set_fp8_shortlist(uhann_scan* scan, uint32_t rows);

/* Of the last GRAM_FP8 or GRAM_INT8 batch on this workspace: queries whose
 * answer the strategy's error bound proves equal to the direct path's (no row
 * outside the candidates can be closer; every query under GRAM_INT8 unless it
 * escalated). Synchronises the workspace stream. No counterpart in the
 * reference. */
int uhann_workspace_shortlist_stats(uhann_workspace* ws, uint64_t* certified);

/* Counters of the last GRAM_BF16 or GRAM_INT8 batch on this workspace: total
 * candidate set size, queries decided in binary64, queries escalated to the
 * direct scan. Synchronises the workspace stream. */
int uhann_workspace_mixed_stats(uhann_workspace* ws, uint64_t* candidates, uint64_t* fp64_reranked,
                                uint64_t* escalated);

/* Stage timing of GRAM_BF16 batches on this workspace with HIP events on the
 * launch stream (enable != 0 turns it on; each timed batch then waits for its
 * last event). uhann_workspace_stage_times copies up to n stage durations in
 * milliseconds: bf16 conversion, Gram GEMM, select + re-rank, escalation. */
int uhann_workspace_set_timing(uhann_workspace* ws, int enable);
int uhann_workspace_stage_times(const uhann_workspace* ws, float* ms, int n);

/* ==== indexes ============================================================
 * Exact, flat (importance-sampled coordinate multiset I) and hierarchical
 * (I projected to p dimensions) indexes, L2 or L1. Queries are full-width
 * (dim) rows; the gather of I runs on the device. */
typedef struct uhann_index uhann_index;

/* One sampled coordinate, the reference's SampledColumn
 * (include/ultrahigh_ann/coordinate_sampling/coordinate_sample.hpp:10-14). */
typedef struct uhann_sampled_column {
    uint64_t source_column;
    uint64_t multiplicity;
    double inverse_probability;
} uhann_sampled_column;

/* Exact index. Replaces CudaExactL2Index(const DenseMatrix&, int device)
 * (include/ultrahigh_ann/exact/l2/cuda_exact_l2_index.hpp) and
 * CudaExactL1Index (exact/l1/cuda_exact_l1_index.hpp). */
int uhann_exact_create(int metric, const float* host_refs, uint64_t rows, uint64_t dim, int device,
                       uhann_index** out);

/* Flat index from an explicit sample (ascending source columns). L2 scales
 * the sampled references by (float)sqrt(m/p); L1 keeps them raw and weighs
 * |r - q| by (float)(m/p). An empty sample scans zero coordinates (row 0).
 * Replaces CudaSampledCoordinateL2AnnIndex(const DenseMatrix&,
 * CoordinateSample, int device) (coordinate_sampling/
 * cuda_sampled_coordinate_l2_ann_index.hpp:45-48) and its L1 twin
 * (cuda_sampled_coordinate_l1_ann_index.hpp); the probability-based
 * CudaFlatL{1,2}AnnIndex constructors (hnsvw25/l{1,2}/flat/cuda_l*_ann_index.hpp)
 * are uhann_build_coordinate_sample + this call. Errors: column out of range
 * -> OUT_OF_RANGE; zero multiplicity or bad weights -> INVALID_ARGUMENT. */
int uhann_flat_create(int metric, const float* host_refs, uint64_t rows, uint64_t dim,
                      const uhann_sampled_column* sample, uint64_t sample_count, int device,
                      uhann_index** out);

/* Hierarchical index from an explicit sample and its binary64 projection
 * matrix (projection_dim x sample_count, row-major, as drawn by
 * uhann_build_projection). The matrix is narrowed to binary32 and the sampled
 * references projected on the host in binary32, as the reference's FP32
 * build does; queries are projected on the device in the reference's K7
 * order and scanned with squared L2 (L2) or the median of |differences|
 * (L1, projection_dim <= 128). Replaces CudaHierarchicalL2AnnIndex(const
 * DenseMatrix&, CoordinateSample, p, engine, device)
 * (hnsvw25/l2/hierarchical/cuda_hierarchical_l2_ann_index.hpp) and
 * CudaHierarchicalL1AnnIndex. Empty sample with rows > 1 -> RUNTIME. */
int uhann_hierarchical_create(int metric, const float* host_refs, uint64_t rows, uint64_t dim,
                              const uhann_sampled_column* sample, uint64_t sample_count,
                              const double* projection, uint64_t projection_dim, int device,
                              uhann_index** out);

void uhann_index_destroy(uhann_index* index);

/* Replaces space_usage()/IndexSpaceUsage (include/ultrahigh_ann/core/
 * index_space_usage.hpp): kind, metric, shape, unique sampled coordinates,
 * summed multiplicity and payload bytes counted the reference's way. */
int uhann_index_info(const uhann_index* index, int* kind, int* metric, uint64_t* rows, uint64_t* dim,
                     uint64_t* unique_coordinates, uint64_t* multiplicity, uint64_t* payload_bytes);

/* Prepares GRAM_BF16, GRAM_FP8 or GRAM_INT8 for an exact or flat L2 index (see
 * uhann_scan_enable_mixed). */
int uhann_index_enable_mixed(uhann_index* index, int strategy);

/* Replaces make_query_workspace(max_batch) of every facade. */
int uhann_index_workspace_create(const uhann_index* index, uint64_t max_batch,
                                 uhann_workspace** out);

/* Synchronous batch from host queries (count x dim). Replaces the facades'
 * query_batch (e.g. cuda_sampled_coordinate_l2_ann_index.hpp:63-68). */
int uhann_index_query_batch(const uhann_index* index, uhann_workspace* ws, const float* host_queries,
                            uint64_t count, uint64_t* out_idx, float* out_dist, int strategy);

/* Synchronous batch from device queries (row stride dim), host results. */
int uhann_index_query_device_batch(const uhann_index* index, uhann_workspace* ws,
                                   const float* device_queries, uint64_t count, uint64_t* out_idx,
                                   float* out_dist, int strategy);

/* Asynchronous device-to-device batch on a caller stream. Data errors (a
 * non-finite sampled coordinate, projection or median) are kept on the
 * workspace and reported by uhann_workspace_status. No counterpart in the
 * reference. */
int uhann_index_query_device_async(const uhann_index* index, uhann_workspace* ws,
                                   const float* device_queries, uint64_t ld_queries, uint64_t count,
                                   uint64_t* device_out_idx, float* device_out_dist, int strategy,
                                   void* stream);

/* Waits for the device and reports (then clears) a data error left by
 * asynchronous batches: INVALID_ARGUMENT for a non-finite sampled query
 * coordinate, OVERFLOW for a non-finite projection or median. */
int uhann_workspace_status(uhann_workspace* ws);

/* ==== stand-alone hierarchical components ================================
 * The query projection and the median-L1 scan as separate handles, for
 * callers that compose them as the reference's hierarchical facades compose
 * its detail classes. Workspaces are uhann_workspace handles created by the
 * component's own call (a workspace from another handle is rejected). */
typedef struct uhann_projection uhann_projection;
typedef struct uhann_median uhann_median;

/* p x k binary32 projection matrix (row-major) copied to `device`. Replaces
 * detail::CudaFloatProjection(projection_matrix, device)
 * (src/core/cuda_float_projection.hpp). p == 0 -> INVALID_ARGUMENT; a
 * non-finite entry -> OVERFLOW (validate_finite_result). */
int uhann_projection_create(const float* host_matrix, uint64_t p, uint64_t k, int device, uhann_projection** out);
void uhann_projection_destroy(uhann_projection* projection);
int uhann_projection_workspace_create(const uhann_projection* projection, uint64_t max_batch,
                                      uhann_workspace** out);
/* count x k host sampled queries -> *device_projected, a row-major
 * count x p binary32 matrix on the device in the reference's K7 order
 * (cuda_float_projection.cu:234-263), valid until the workspace is used
 * again. Synchronous. Workspace owner, capacity and finite queries are
 * checked first (INVALID_ARGUMENT); a non-finite result -> OVERFLOW.
 * Replaces project_batch_float (both reference strategies: the binary32 K7
 * order is the one this engine reproduces). */
int uhann_projection_project(const uhann_projection* projection, uhann_workspace* ws, const float* host_queries,
                             uint64_t count, const float** device_projected);
int uhann_projection_info(const uhann_projection* projection, uint64_t* projection_dim, uint64_t* sampled_dim,
                          uint64_t* payload_bytes, int* device);

/* rows x p binary32 projected references (row-major) on `device`. Replaces
 * detail::CudaMedianL1ScanIndex(projected_representatives, device)
 * (src/hnsvw25/l1/hierarchical/cuda_median_l1_scan_index.hpp). rows == 0 or
 * p == 0 or p above the supported maximum -> INVALID_ARGUMENT; a non-finite
 * value -> OVERFLOW. */
int uhann_median_create(const float* host_projected_refs, uint64_t rows, uint64_t p, int device, uhann_median** out);
void uhann_median_destroy(uhann_median* median);
int uhann_median_workspace_create(const uhann_median* median, uint64_t max_batch, uhann_workspace** out);
/* Per query (count x p device rows, stride p) the row with the smallest
 * median of |differences| (lowest row on ties) into host out_idx.
 * Synchronous; a non-finite median -> OVERFLOW. Replaces
 * query_device_batch (cuda_median_l1_scan_index.cu:429-505). */
int uhann_median_query_device_batch(const uhann_median* median, uhann_workspace* ws, const float* device_projected,
                                    uint64_t count, uint64_t* out_idx);
int uhann_median_info(const uhann_median* median, uint64_t* rows, uint64_t* projection_dim, uint64_t* payload_bytes,
                      int* device);

/* Device metadata for provenance. Replaces cuda_runtime_info(device)
 * (include/ultrahigh_ann/core/cuda_runtime_info.hpp): name, architecture
 * major/minor (gfx950: 9, 5), total memory, the HIP version this library
 * was built with, the runtime and driver versions. No device -> RUNTIME. */
int uhann_device_properties(int device, char* name, size_t name_cap, int* arch_major, int* arch_minor,
                            uint64_t* total_memory_bytes, int* compiled_runtime_version, int* runtime_version,
                            int* driver_version);

/* ==== host index construction (reference <random> order) =================
 * A std::mt19937_64 held behind a handle; the draws match the reference's
 * libstdc++ distributions call for call. */
typedef struct uhann_rng uhann_rng;

/* std::mt19937_64(seed), the engine every reference constructor takes. */
int uhann_rng_create(uint64_t seed, uhann_rng** out);
void uhann_rng_destroy(uhann_rng* rng);
/* *equal = 1 when both engines are in the same state (the reference's
 * "consume identical randomness" check, cuda_hierarchical_l2_ann_index_test.cpp:88-90). */
int uhann_rng_equal(const uhann_rng* a, const uhann_rng* b, int* equal);
/* Next raw 64-bit draw. */
int uhann_rng_next(uhann_rng* rng, uint64_t* value);

/* build_coordinate_sample (coordinate_sampling.hpp, coordinate_sampling.cpp:
 * 14-45). Writes the kept columns to out (capacity entries; dim always
 * suffices) and their number to *count. Invalid probability -> DOMAIN. */
int uhann_build_coordinate_sample(uhann_rng* rng, const double* probabilities, uint64_t dim,
                                  uint64_t repetitions, uhann_sampled_column* out, uint64_t capacity,
                                  uint64_t* count);

/* compute_sampling_mass (coordinate_sampling/uniform_probabilities.cpp:
 * 12-26): the dim probabilities summed in order. Invalid probability ->
 * DOMAIN. */
int uhann_sampling_mass(const double* probabilities, uint64_t dim, double* mass);

/* make_mass_matched_uniform_probabilities (uniform_probabilities.cpp:28-45):
 * dim copies of mass / dim into out, the probabilities of the uniform index
 * kind (UniformL{1,2}AnnIndex, uniform_l{1,2}_ann_index.cpp; on the GPU the
 * reference builds a CudaSampledCoordinateL{1,2}AnnIndex from their sample,
 * nearest_neighbor_benchmark.cpp:755-768). Mass outside [0, dim] or not
 * finite -> DOMAIN. */
int uhann_uniform_probabilities(uint64_t dim, double mass, double* out);

/* Projection matrix drawn after the sample: L2 JL (Rademacher sums,
 * l2_hierarchical_projection.cpp:53-77), L1 Cauchy
 * (l1_hierarchical_projection.cpp:54-77); out is projection_dim x count
 * binary64, row-major. */
int uhann_build_projection(uhann_rng* rng, int metric, const uhann_sampled_column* sample,
                           uint64_t count, uint64_t projection_dim, double* out);

/* Importance probabilities on the host in binary64 (L2:
 * compute_l2_importance_probabilities, hnsvw25/l2/importance_sampling.cpp:
 * 18-107; L1: hnsvw25/l1/importance_sampling.cpp:19-104), dim doubles. */
int uhann_importance_probabilities(int metric, const float* refs, uint64_t rows, uint64_t dim,
                                   double* out);

/* The same probabilities computed on `device`, bit-identical to
 * uhann_importance_probabilities (binary64 pair distances summed in the CPU
 * order, IEEE division, order-free maximum). Device memory: rows x dim floats
 * + rows^2 doubles. Replaces compute_l{1,2}_importance_probabilities_gpu
 * (hnsvw25/l{1,2}/cuda_importance_sampling.cu), whose FP32 result differs
 * from the CPU policy. */
int uhann_importance_probabilities_device(int metric, const float* host_refs, uint64_t rows,
                                          uint64_t dim, int device, double* out);

/* ==== importance-probability files (the reference's UAP v2 format) ======
 * io/importance_probability_io.cpp and io/sha256.hpp; layout and checks in
 * include/ultrahigh_ann_amd/probability_file.hpp. File and format errors ->
 * RUNTIME with the reference's message ("invalid importance-probability
 * file <path>: <reason>"). */

/* sha256_file: SHA-256 of the file's exact bytes into digest[32]. */
int uhann_sha256_file(const char* path, uint8_t* digest);

/* save_importance_probabilities (importance_probability_io.cpp:124-162):
 * metric UHANN_METRIC_L1 or UHANN_METRIC_L2 (stored as kind 1 / 2),
 * representative count, the representative matrix file's digest[32] and dim
 * probabilities (finite, in [0, 1]). Creates parent directories. */
int uhann_save_probabilities(const char* path, int metric, uint64_t representative_count,
                             const uint8_t* digest, const double* probabilities, uint64_t dim);

/* load_importance_probabilities (importance_probability_io.cpp:164-227):
 * header into *metric, *representative_count and digest[32], the coordinate
 * count into *dim and the values into out. With out == NULL only the header
 * and *dim are returned; capacity < *dim -> LENGTH. */
int uhann_load_probabilities(const char* path, int* metric, uint64_t* representative_count, uint8_t* digest,
                             double* out, uint64_t capacity, uint64_t* dim);

/* The mixed strategy's Gram estimates ||q||^2 + ||v||^2 - 2 <bf16(q), bf16(v)>
 * (binary32 MFMA accumulation) for count device queries against every
 * reference, written to device_out (count rows of ld_out >= rows floats) on
 * `stream`. tile -1 runs what the queries run: the default block shape,
 * with the tiles of a partial last round split along K over the idle
 * compute units (their parts summed in order afterwards); 0..7 force a
 * block shape and keep every tile whole (one sequential K accumulation per
 * output, the same bits for every shape). Needs uhann_scan_enable_mixed. No counterpart in the reference (its GEMM
 * strategy keeps the unrefined SGEMM distances internal): for diagnostics
 * such as full distance tables and for the tests of each block shape. */
int uhann_scan_gram_estimates(const uhann_scan* scan, uhann_workspace* ws, const float* device_queries,
                              uint64_t ld_queries, uint64_t count, float* device_out, uint64_t ld_out,
                              int tile, void* stream);

/* Exact diagnostic distances for count device queries against every
 * reference of an unweighted handle, binary64 in the reference's CPU order
 * (L2: sqrt of the sequential sum of squared differences, L1: the
 * transform_reduce sum), written to device_out (count rows of ld_out >= rows
 * doubles) on `stream`. Replaces the O(B r d) loop of the reference's
 * ExactDistanceTable constructor (benchmarks/benchmark_metrics.cpp:112-192,
 * exact_distance :32-41); the statistics over the table stay on the host. */
int uhann_scan_distance_table(const uhann_scan* scan, const float* device_queries, uint64_t ld_queries,
                              uint64_t count, double* device_out, uint64_t ld_out, void* stream);

/* The GRAM_FP8 strategy's estimates ||q||^2 + ||v||^2 - 2 s_q s_v <q8, v8>
 * for count device queries against every reference, as
 * uhann_scan_gram_estimates. Needs uhann_scan_enable_mixed(GRAM_FP8). No
 * counterpart in the reference: for diagnostics and tests. */
int uhann_scan_fp8_estimates(const uhann_scan* scan, uhann_workspace* ws, const float* device_queries,
                             uint64_t ld_queries, uint64_t count, float* device_out, uint64_t ld_out,
                             void* stream);

/* The GRAM_INT8 strategy's estimates ||q||^2 + ||v||^2 - 2 s_q s_v (H + X /
 * 256) (exact int32 sums, binary64 combination, rounded to binary32) for
 * count device queries against every reference, as uhann_scan_gram_estimates
 * (tile -1: the default shape for the batch with its K split; 0 .. 7 force a
 * block shape, every tile whole; all give the same bits, since the sums are
 * exact). Needs
 * uhann_scan_enable_mixed(GRAM_INT8). No counterpart in the reference: for
 * diagnostics and tests. */
int uhann_scan_int8_estimates(const uhann_scan* scan, uhann_workspace* ws, const float* device_queries,
                              uint64_t ld_queries, uint64_t count, float* device_out, uint64_t ld_out, int tile,
                              void* stream);

/* The GRAM_INT8 row conversion alone: count device rows (stride ld_rows,
 * dim <= 65536 columns) into device_bytes (count rows of 128 ceil(dim / 64)
 * bytes: per 64 columns, 64 bytes of the high plane a, then 64 of the low
 * plane b) and device_stats (count records of 32 bytes: binary64 ||x||^2,
 * then binary32 s, ||x - x^||, ||x^||, s ||b|| / 256, ||x||, padding). No
 * counterpart in the reference: lets the tests check the codes and the
 * rounded-up norms the certificate is built from. */
int uhann_int8_rows(const float* device_rows, uint64_t ld_rows, uint64_t count, uint64_t dim, uint8_t* device_bytes,
                    void* device_stats, void* stream);

/* Autotuner of the direct scan: times every compiled kernel shape (the
 * tiled register tiles and the pair-per-thread scan) with HIP events for
 * batches of 1, 2, 4, ... up to max_batch device queries (row stride
 * ld_queries) and keeps the fastest per batch bucket in the handle; later
 * DIRECT calls use it. Writes one row of 6 ints per tuned bucket to table
 * (bucket b covers batches 2^(b-1)+1 .. 2^b; rpt, tq, tr, group_q;
 * microseconds) up to capacity rows, and the row count to *count. Call before
 * sharing the handle between threads. No counterpart in the reference. */
int uhann_scan_autotune(uhann_scan* scan, uhann_workspace* ws, const float* device_queries,
                        uint64_t ld_queries, uint64_t max_batch, int* table, int capacity, int* count);

/* uhann_scan_autotune for the direct scan inside an exact, flat or
 * hierarchical-L2 index, fed with max_batch full-width device queries (the
 * gather and projection run first). INVALID_ARGUMENT for hierarchical L1. */
int uhann_index_autotune(uhann_index* index, uhann_workspace* ws, const float* device_queries,
                         uint64_t ld_queries, uint64_t max_batch, int* table, int capacity, int* count);

/* ------------------------------------------------------------------------
 * Serving loop. No counterpart in the reference, which runs one facade
 * query_batch at a time (nearest_neighbor_benchmark.cpp:386-413 times them);
 * each request here executes what that facade's query_batch computes
 * (cuda_sampled_coordinate_l{1,2}_ann_index.cpp, cuda_hierarchical_l{1,2}_
 * ann_index.cpp, cuda_exact_l{1,2}_index.cpp), bit for bit.
 *
 * A server owns a set of indexes on one device. Registered indexes are built
 * in HBM when their device bytes fit the server's budget (first come, first
 * served); the others wait, and removing an index admits them. Requests (one
 * batch of queries for one index) are queued and a worker thread packs the
 * queued requests of resident indexes into ticks: up to max_tick_queries
 * query rows and max_tick_requests requests, whose work runs together in one
 * set of grouped persistent kernels (one hipGraph per tick shape, captured on
 * first use and replayed). Two ticks are in flight: requests that arrive
 * while one runs are packed into the next as rows free up. A request's
 * results do not depend on what else shares its tick.
 * ------------------------------------------------------------------------ */
typedef struct uhann_server uhann_server;

typedef struct uhann_server_config {
    int device;
    uint64_t hbm_budget_bytes;   /* device bytes of resident indexes */
    uint32_t max_tick_queries;   /* query rows per tick (>= the largest request) */
    uint32_t max_tick_requests;  /* requests per tick */
    uint64_t scratch_floats;     /* per tick: gathered + projected queries */
    uint64_t staging_floats;     /* per tick: host queries copied to the device */
    int use_graph;               /* 1: replay captured graphs, 0: launch the kernels */
    int pairs_tq;                /* queries per block of the pair-per-thread scans (1, 2: whole
                                    32-chain groups; 4: chain classes, the default) */
} uhann_server_config;

/* Index description for uhann_server_register. host_refs is borrowed until
 * the index is resident (uhann_server_index_state); sample and projection are
 * copied. kind EXACT ignores the sample; FLAT ignores the projection;
 * HIERARCHICAL takes projection_dim x sample_count binary64 values drawn as
 * uhann_build_projection draws them. strategy: DIRECT, or PROJECTION_BF16 for
 * a hierarchical L2 index. */
typedef struct uhann_index_spec {
    int kind;
    int metric;
    int strategy;
    const float* host_refs;
    uint64_t rows;
    uint64_t dim;
    const uhann_sampled_column* sample;
    uint64_t sample_count;
    const double* projection;
    uint64_t projection_dim;
} uhann_index_spec;

#define UHANN_INDEX_PENDING 0   /* waiting for HBM budget */
#define UHANN_INDEX_RESIDENT 1
#define UHANN_INDEX_FAILED 2    /* the build failed; requests report its error */

typedef struct uhann_serving_stats {
    uint64_t requests;           /* completed since the last reset */
    uint64_t queries;
    uint64_t ticks;
    double p50_ms;               /* request latency: submit -> results ready */
    double p99_ms;
    double mean_ms;
    double max_ms;
    double wall_s;               /* first submit -> last completion */
    double queries_per_s;        /* queries / wall_s */
    double gpu_busy_s;           /* summed tick durations on the device */
    double mean_tick_queries;
    uint64_t resident_indexes;
    uint64_t pending_indexes;
    uint64_t resident_bytes;
} uhann_serving_stats;

/* One request of a replayed trace: submitted arrival_s seconds after the
 * replay starts. */
typedef struct uhann_trace_entry {
    double arrival_s;
    uint64_t index_id;
    const float* queries;
    uint64_t ld_queries;
    uint64_t count;
    int on_device;
    uint64_t* out_idx;
    float* out_dist;
} uhann_trace_entry;

int uhann_server_create(const uhann_server_config* config, uhann_server** out);
/* Waits for queued requests, then frees every index and buffer. */
void uhann_server_destroy(uhann_server* server);
int uhann_server_register(uhann_server* server, const uhann_index_spec* spec, uint64_t* index_id);
int uhann_server_index_state(uhann_server* server, uint64_t index_id, int* state, uint64_t* device_bytes);
/* Waits until the index has no request in flight, frees it and admits waiting
 * indexes that now fit. */
int uhann_server_remove(uhann_server* server, uint64_t index_id);
/* Queues count queries (row stride ld_queries floats >= the index dimension;
 * host memory, or device memory on the server's device when on_device) for
 * an index. queries, out_idx (count values) and out_dist (count values, or
 * null) must stay valid until the request completes. Data errors (a
 * non-finite query value the index reads, a non-finite projection or median)
 * are reported by uhann_server_wait with the index's error type. */
int uhann_server_submit(uhann_server* server, uint64_t index_id, const float* queries, int on_device,
                        uint64_t ld_queries, uint64_t count, uint64_t* out_idx, float* out_dist,
                        uint64_t* request_id);
/* Waits up to timeout_s seconds (< 0: no limit). *done = 1 once the request
 * completed; then *request_status is its status (uhann_last_error holds its
 * message) and the id is released. */
int uhann_server_wait(uhann_server* server, uint64_t request_id, double timeout_s, int* done,
                      int* request_status);
/* Submits the entries at their arrival times from the calling thread (open
 * loop) and writes their request ids; returns when the last one is queued. */
int uhann_server_run_trace(uhann_server* server, const uhann_trace_entry* entries, uint64_t count,
                           uint64_t* request_ids);
/* Waits until no request is queued or in flight. */
int uhann_server_drain(uhann_server* server);
int uhann_server_stats(uhann_server* server, uhann_serving_stats* out);
int uhann_server_reset_stats(uhann_server* server);
/* Times the pair-scan query tiles (1, 2, 4 queries per block) on one tick of
 * `rows_per_index` device queries (row stride ld_queries, at least the widest
 * resident index) for every resident index, keeps the fastest and writes it
 * to *pairs_tq. The server must be idle. */
int uhann_server_autotune(uhann_server* server, const float* device_queries, uint64_t ld_queries,
                          uint32_t rows_per_index, int* pairs_tq);

#ifdef __cplusplus
}
#endif

#endif /* UHANN_H */
