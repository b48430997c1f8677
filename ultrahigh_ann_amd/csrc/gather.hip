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

// dst row i <- src row rows[i] (dim floats), one block per row.
__global__ void gather_rows_kernel(const float* __restrict__ src, uint64_t ld_src,
                                   const unsigned int* __restrict__ rows, uint32_t n, uint32_t dim,
                                   float* __restrict__ dst, uint64_t ld_dst)
{
    const uint32_t i = blockIdx.x;
    if (i >= n) return;
    const float* s = src + static_cast<uint64_t>(rows[i]) * ld_src;
    float* d = dst + static_cast<uint64_t>(i) * ld_dst;
    for (uint32_t c = threadIdx.x; c < dim; c += blockDim.x) d[c] = s[c];
}

__global__ void scatter_results_kernel(const uint64_t* __restrict__ idx, const float* __restrict__ dist,
                                       const unsigned int* __restrict__ rows, uint32_t n,
                                       uint64_t* __restrict__ out_idx, float* __restrict__ out_dist)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (out_idx != nullptr) out_idx[rows[i]] = idx[i];
    if (out_dist != nullptr) out_dist[rows[i]] = dist[i];
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

void gather_rows(const float* src, uint64_t ld_src, const unsigned int* rows, uint32_t n, uint32_t dim,
                 float* dst, uint64_t ld_dst, hipStream_t stream)
{
    if (n == 0) return;
    gather_rows_kernel<<<n, 256, 0, stream>>>(src, ld_src, rows, n, dim, dst, ld_dst);
    check_launch("row gather kernel");
}

void scatter_results(const uint64_t* idx, const float* dist, const unsigned int* rows, uint32_t n,
                     uint64_t* out_idx, float* out_dist, hipStream_t stream)
{
    if (n == 0) return;
    scatter_results_kernel<<<(n + 255) / 256, 256, 0, stream>>>(idx, dist, rows, n, out_idx, out_dist);
    check_launch("result scatter kernel");
}

}  // namespace uhann::kern
