// Launch interface of the GPU importance-probability kernels (probabilities.hip).
#pragma once

#include <hip/hip_runtime_api.h>

#include <cstdint>

namespace uhann::kern {

// out_bits[c] <- bit pattern of p_c (binary64, the reference's sequential CPU
// value). refs: rows x dim binary32 on the device, row stride ld. inv_scratch:
// rows x rows doubles (only a < b is used).
void importance_probabilities(bool l1, const float* refs, uint64_t ld, uint32_t rows, uint32_t dim,
                              double* inv_scratch, unsigned long long* out_bits, hipStream_t stream);

// out[q][r] (row stride ldo) = the binary64 distance of query q and
// reference r in the reference's CPU order: L2 sqrt(sum (q - r)^2) (the
// diagnostic distance), L1 sum |q - r|. Queries nq x dim (stride ldq) and
// references nr x dim (stride ldr) binary32 on the device.
void distance_table(bool l1, const float* queries, uint64_t ldq, uint32_t nq, const float* refs, uint64_t ldr,
                    uint32_t nr, uint32_t dim, double* out, uint64_t ldo, hipStream_t stream);

}  // namespace uhann::kern
