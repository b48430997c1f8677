// Hierarchical-index kernels: the binary32 projection of gathered query
// coordinates and the median-L1 scan over projected vectors.
//
// project_fp32 reproduces the reference's K7 (project_queries_direct,
// cuda_float_projection.cu:234-263) bit for bit: y[q][o] is the reference's
// block_sum of 256 chains acc = fmaf(x[c], P[o][c], acc) over c = t, t+256, ...
// Projection sums can be negative or -0.0, so the second-stage tree keeps the
// reference's two "+0.0f" adds of lanes 8..31 (they turn -0.0 into +0.0).
// Layout: one block of 256 threads (thread t = chain t) computes a
// kProjQ x kProjP tile of outputs; each loaded x value is reused kProjP times
// and each P value kProjQ times.
//
// project_bf16 is the MFMA form of the same projection for the L2 hierarchy:
// y = x P^T with both operands rounded to bf16 (round to nearest even) and
// v_mfma_f32_32x32x16_bf16 accumulating in binary32. A block of four waves
// owns a 32 x 32 output tile; wave w takes the K steps w, w + 4, ... and the
// four partial tiles are added in LDS in a fixed order, so the result is
// deterministic (it is not the binary32 path's bits: the rounding of x and P
// to bf16 changes the projection by about 2^-9 relative).
//
// median_l1 replaces K9/K10 (cuda_median_l1_scan_index.cu:202-301): the
// estimate for a pair is the median of |r_k - q_k| over the p projected
// coordinates (even p: lower + (upper - lower) * 0.5f). The median depends
// only on the multiset, so instead of a shared-memory bitonic sort per pair
// each thread sorts its pair's p values in registers with a compile-time
// bitonic network (padded with +inf to a power of two). A block takes one
// query (its vector staged in LDS) and 256 references read column-wise from
// a transposed (p x rows) copy, so every load is coalesced. The block's best
// (estimate, row) joins the per-query atomicMin key like the direct scan.

#include <hip/hip_runtime.h>

#include <stdexcept>
#include <string>

#include "bodies.hpp"
#include "hier.hpp"

namespace uhann::kern {
namespace {

using body::kMedianThreads;
using body::kProjP;
using body::kProjQ;
using body::kProjThreads;
using body::kProjWaves;

void check_launch(const char* what)
{
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        throw std::runtime_error(std::string("cannot launch ") + what + ": " +
                                 hipGetErrorString(e));
    }
}

__global__ void __launch_bounds__(kProjThreads)
project_fp32_kernel(const float* __restrict__ x, uint64_t ld_x, uint32_t nq, uint32_t k,
                    const float* __restrict__ proj, uint64_t ld_proj, uint32_t p,
                    float* __restrict__ y, uint64_t ld_y, int* invalid)
{
    const uint32_t tiles_p = (p + kProjP - 1) / kProjP;
    const uint32_t q0 = (blockIdx.x / tiles_p) * kProjQ;
    const uint32_t o0 = (blockIdx.x % tiles_p) * kProjP;
    body::project_fp32_tile(x, ld_x, nq, k, proj, ld_proj, p, y, ld_y, invalid, q0, o0);
}

__global__ void __launch_bounds__(64 * kProjWaves)
project_bf16_kernel(const float* __restrict__ x, uint64_t ld_x, uint32_t nq, uint32_t k,
                    const __bf16* __restrict__ proj, uint64_t ld_proj, uint32_t p,
                    float* __restrict__ y, uint64_t ld_y, int* invalid)
{
    body::project_bf16_tile(x, ld_x, nq, k, proj, ld_proj, p, y, ld_y, invalid, blockIdx.x * 32, blockIdx.y * 32);
}

template <int PMAX, bool EXACT>
__global__ void __launch_bounds__(kMedianThreads)
median_l1_kernel(const float* __restrict__ refs_t, uint32_t rows, const float* __restrict__ yq,
                 uint64_t ld_y, uint32_t nq, uint32_t p, unsigned long long* __restrict__ keys,
                 int* invalid)
{
    body::median_block<PMAX, EXACT>(refs_t, rows, yq, ld_y, nq, p, keys, invalid, blockIdx.y, blockIdx.x);
}

__global__ void zero_results_kernel(uint64_t* idx, float* dist, uint32_t n)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        if (idx != nullptr) idx[i] = 0;
        if (dist != nullptr) dist[i] = 0.0F;
    }
}

}  // namespace

void project_fp32(const float* x, uint64_t ld_x, uint32_t nq, uint32_t k, const float* proj,
                  uint64_t ld_proj, uint32_t p, float* y, uint64_t ld_y, int* invalid,
                  hipStream_t stream)
{
    if (nq == 0 || p == 0) return;
    const uint64_t blocks = static_cast<uint64_t>((nq + kProjQ - 1) / kProjQ) * ((p + kProjP - 1) / kProjP);
    if (blocks > 0x7fffffffULL) throw std::length_error("projection grid exceeds the launch limit");
    project_fp32_kernel<<<static_cast<uint32_t>(blocks), kProjThreads, 0, stream>>>(
        x, ld_x, nq, k, proj, ld_proj, p, y, ld_y, invalid);
    check_launch("binary32 projection kernel");
}

uint32_t median_max_dimension() { return 128; }
uint64_t project_bf16_rows(uint64_t p) { return (p + 31U) / 32U * 32U; }
uint64_t project_bf16_ld(uint64_t k) { return (k + 15U) / 16U * 16U + 16U; }

void project_bf16(const float* x, uint64_t ld_x, uint32_t nq, uint32_t k, const void* proj16,
                  uint64_t ld_proj, uint32_t p, float* y, uint64_t ld_y, int* invalid, hipStream_t stream)
{
    if (nq == 0 || p == 0) return;
    if (ld_x % 4 != 0 || ld_proj % 8 != 0) throw std::invalid_argument("bf16 projection operands are misaligned");
    const dim3 grid((nq + 31U) / 32U, (p + 31U) / 32U);
    project_bf16_kernel<<<grid, 64 * kProjWaves, 0, stream>>>(x, ld_x, nq, k, static_cast<const __bf16*>(proj16),
                                                             ld_proj, p, y, ld_y, invalid);
    check_launch("bf16 projection kernel");
}

void median_l1(const float* refs_t, uint32_t rows, const float* yq, uint64_t ld_y, uint32_t nq,
               uint32_t p, unsigned long long* keys, int* invalid, hipStream_t stream)
{
    if (nq == 0 || rows == 0) return;
    if (p == 0 || p > median_max_dimension()) {
        throw std::invalid_argument("median-L1 projection dimension must be in [1, 128]");
    }
    // grid.y carries the query; batches above 65535 go in slices.
    for (uint32_t q0 = 0; q0 < nq; q0 += 65535U) {
        const uint32_t n = min(65535U, nq - q0);
        const dim3 grid((rows + kMedianThreads - 1) / kMedianThreads, n);
        const float* yb = yq + static_cast<uint64_t>(q0) * ld_y;
        unsigned long long* kb = keys + q0;
        // p equal to the network size: the kernel with fixed middle positions.
#define UHANN_MEDIAN(PM)                                                                                   \
    if (p == PM) {                                                                                         \
        median_l1_kernel<PM, true><<<grid, kMedianThreads, 0, stream>>>(refs_t, rows, yb, ld_y, n, p, kb, invalid); \
    } else {                                                                                               \
        median_l1_kernel<PM, false><<<grid, kMedianThreads, 0, stream>>>(refs_t, rows, yb, ld_y, n, p, kb, invalid); \
    }
        if (p <= 8) {
            UHANN_MEDIAN(8)
        } else if (p <= 16) {
            UHANN_MEDIAN(16)
        } else if (p <= 32) {
            UHANN_MEDIAN(32)
        } else if (p <= 64) {
            UHANN_MEDIAN(64)
        } else {
            UHANN_MEDIAN(128)
        }
#undef UHANN_MEDIAN
        check_launch("median-L1 kernel");
    }
}

void zero_results(uint64_t* idx, float* dist, uint32_t n, hipStream_t stream)
{
    if (n == 0) return;
    zero_results_kernel<<<(n + 255) / 256, 256, 0, stream>>>(idx, dist, n);
    check_launch("result fill kernel");
}

}  // namespace uhann::kern
