// Device gather of the sampled coordinate multiset I (and its per-coordinate
// importance scaling) from device-resident full-width queries.
//
// The reference packs on the host inside the timed query
// (l2_coordinate_embedding.cpp:90-127 out = q[src_j] * scale_j;
// l1_coordinate_embedding.cpp:81-111 and cuda_hierarchical_l2_ann_index.cpp:
// 128-161 copy raw values), checking only the sampled values for finiteness.
// Here one block owns a chunk of up to 1024 sampled columns: it stages the
// column indices and scales in LDS once and reuses them for every query row
// it visits; consecutive threads write consecutive outputs (coalesced) and
// read ascending source columns of the same query row. The scaled value is a
// single binary32 multiply, exactly as the reference computes it.

#include <hip/hip_runtime.h>

#include <stdexcept>
#include <string>

#include "bodies.hpp"
#include "gram.hpp"
#include "kernels.hpp"

namespace uhann::kern {
namespace {

using body::kGatherChunk;
using body::kGatherRows;
using body::kGatherThreads;

__global__ void __launch_bounds__(kGatherThreads)
gather_kernel(const float* __restrict__ src, uint64_t ld_src, uint32_t nq,
              const uint32_t* __restrict__ cols, const float* __restrict__ scales, uint32_t k,
              float* __restrict__ out, uint64_t ld_out, int* invalid)
{
    body::gather_tile(src, ld_src, nq, cols, scales, k, out, ld_out, invalid, blockIdx.x, blockIdx.y);
}

__global__ void check_finite_kernel(const float* __restrict__ values, uint64_t n, int* invalid)
{
    bool bad = false;
    for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
        bad |= !__builtin_isfinite(values[i]);
    }
    if (bad) atomicExch(invalid, 1);
}

__global__ void zero_u32_kernel(unsigned int* p, uint32_t n)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0U;
}


void check_launch(const char* what)
{
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        throw std::runtime_error(std::string("cannot launch ") + what + ": " +
                                 hipGetErrorString(e));
    }
}

}  // namespace

void gather_columns(const float* src, uint64_t ld_src, uint32_t nq, const uint32_t* cols,
                    const float* scales, uint32_t k, float* out, uint64_t ld_out, int* invalid,
                    hipStream_t stream)
{
    if (nq == 0 || k == 0) return;
    const dim3 grid((k + kGatherChunk - 1) / kGatherChunk, (nq + kGatherRows - 1) / kGatherRows);
    if (grid.y > 65535U * 1024U) throw std::length_error("gather grid exceeds the launch limit");
    gather_kernel<<<grid, kGatherThreads, 0, stream>>>(src, ld_src, nq, cols, scales, k, out,
                                                       ld_out, invalid);
    check_launch("coordinate gather kernel");
}

void check_finite(const float* values, uint64_t n, int* invalid, hipStream_t stream)
{
    if (n == 0) return;
    const uint64_t want = (n + 255) / 256;
    const uint32_t blocks = static_cast<uint32_t>(want < 4096 ? want : 4096);
    check_finite_kernel<<<blocks, 256, 0, stream>>>(values, n, invalid);
    check_launch("finiteness check kernel");
}

void zero_u32(unsigned int* p, uint32_t n, hipStream_t stream)
{
    if (n == 0) return;
    zero_u32_kernel<<<(n + 255) / 256, 256, 0, stream>>>(p, n);
    check_launch("zero fill kernel");
}

}  // namespace uhann::kern
