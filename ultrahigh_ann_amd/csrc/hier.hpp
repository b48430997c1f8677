// Launch interfaces of the hierarchical-index kernels in hier.hip.
#pragma once

#include <hip/hip_runtime_api.h>

#include <cstdint>

namespace uhann::kern {

// y[q][o] = binary32 projection of x[q] (k values) on proj[o] (k values) in
// the reference's 256-chain order, q < nq, o < p. *invalid = 1 on a
// non-finite output.
void project_fp32(const float* x, uint64_t ld_x, uint32_t nq, uint32_t k, const float* proj,
                  uint64_t ld_proj, uint32_t p, float* y, uint64_t ld_y, int* invalid,
                  hipStream_t stream);

// The same projection with x and proj rounded to bf16 on MFMA (binary32
// accumulation). proj16 is a bf16 matrix of ceil(p/32)*32 rows (rows >= p
// zero) and ld_proj >= ceil(k/16)*16 elements (columns >= k zero, ld_proj a
// multiple of 8); ld_x is a multiple of 4.
void project_bf16(const float* x, uint64_t ld_x, uint32_t nq, uint32_t k, const void* proj16,
                  uint64_t ld_proj, uint32_t p, float* y, uint64_t ld_y, int* invalid, hipStream_t stream);
uint64_t project_bf16_rows(uint64_t p);
uint64_t project_bf16_ld(uint64_t k);

// Median-L1 scan: refs_t is p x rows (transposed projected references), yq
// nq x p (row stride ld_y). keys[q] receives atomicMin((median bits << 32) |
// row) and must hold ~0 before the launch. *invalid = 1 on a non-finite
// difference or median. p up to median_register_dimension() sorts in
// registers; larger p (median_supported: 4 nextpow2(p) bytes within the
// current device's per-block shared memory) sorts in LDS.
void median_l1(const float* refs_t, uint32_t rows, const float* yq, uint64_t ld_y, uint32_t nq,
               uint32_t p, unsigned long long* keys, int* invalid, hipStream_t stream);
uint32_t median_register_dimension();
bool median_supported(uint64_t p);
uint32_t next_pow2(uint32_t v);

// idx[i] = 0, dist[i] = 0 for i < n (either may be null).
void zero_results(uint64_t* idx, float* dist, uint32_t n, hipStream_t stream);

}  // namespace uhann::kern
