// Importance probabilities on the GPU, bit-identical to the reference's
// sequential binary64 CPU policy.
//
// Reference: L2 importance_sampling.cpp:18-48, p_j = max over pairs a < b
// with ||a-b||^2 != 0 of min(1, (a_j-b_j)^2 * (1/||a-b||^2)); L1
// l1/importance_sampling.cpp:19-47, p_j = max of |a_j-b_j| * (1/||a-b||_1),
// with the L1 distance summed in libstdc++ transform_reduce's groups of four
// (l1_distance.cpp:12-28). The reference's GPU policy is FP32 and differs;
// this one keeps the CPU policy's value so the sampled index is the same.
//
// 1. pair_inverse: inv[a][b] = 1 / dist(a, b) (0 when dist == 0), for a < b.
//    Each thread owns a 4 x 4 register tile of pairs and sums its pairs'
//    columns strictly in order (c = 0, 1, 2, ...) with separate binary64
//    multiply and add, exactly the CPU loop. Rows are staged in LDS as
//    binary64 (the exact float -> double conversion), 64 x 64 pairs per
//    block, upper-triangular blocks only.
// 2. column_max: thread = column, a block takes 8 consecutive first rows a
//    and sweeps b > a; per pair the candidate probability is computed as on
//    the CPU and the running maximum is published with a 64-bit atomicMax
//    (non-negative doubles order like their bit patterns). max is exact and
//    order-free, so the result is the CPU's, bit for bit.

#include <hip/hip_runtime.h>

#include <stdexcept>
#include <string>

#include "probabilities.hpp"

namespace uhann::kern {
namespace {

constexpr int kPairTile = 64;    // rows per side of a block's pair tile
constexpr int kPairThreads = 256;
constexpr int kChunk = 32;       // columns staged per LDS round
constexpr int kLdsStride = kChunk + 1;

void check_launch(const char* what)
{
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        throw std::runtime_error(std::string("cannot launch ") + what + ": " + hipGetErrorString(e));
    }
}

// The binary64 distances of a 64 x 64 tile of pairs (rows a0.. of A against
// rows b0.. of B), in the CPU order: each thread's 4 x 4 pairs sum their
// columns strictly in order (c = 0, 1, ...) with separate multiply and add
// (L2: (a - b)^2; L1: transform_reduce's init + ((t0 + t1) + (t2 + t3)) per
// complete group of four, then the tail term by term). Rows past the end
// repeat the last row (their sums are not used). Thread (ti, tj) =
// (tid / 16, tid % 16) owns pairs (ti*4 + i, tj*4 + j).
template <bool L1>
__device__ __forceinline__ void pair_tile_fp64(const float* __restrict__ A, uint64_t lda, uint32_t rows_a, uint32_t a0,
                                               const float* __restrict__ B, uint64_t ldb, uint32_t rows_b, uint32_t b0,
                                               uint32_t dim, double (&acc)[4][4])
{
    __shared__ double sa[kPairTile * kLdsStride];
    __shared__ double sb[kPairTile * kLdsStride];
    const int tid = threadIdx.x;
    const int ti = tid / 16;  // 16 x 16 threads, 4 x 4 pairs each
    const int tj = tid % 16;

#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = 0.0;

    // Columns below full_end belong to complete groups of four (L1 grouping);
    // kChunk is a multiple of 4, so a group never straddles two chunks.
    const uint32_t full_end = L1 ? (dim / 4) * 4 : 0;
    for (uint32_t c0 = 0; c0 < dim; c0 += kChunk) {
        const uint32_t n = min(static_cast<uint32_t>(kChunk), dim - c0);
        __syncthreads();
        for (int e = tid; e < kPairTile * kChunk; e += kPairThreads) {
            const int row = e / kChunk;
            const int col = e % kChunk;
            const uint32_t ra = min(a0 + row, rows_a - 1U);
            const uint32_t rb = min(b0 + row, rows_b - 1U);
            const bool in = static_cast<uint32_t>(col) < n;
            sa[row * kLdsStride + col] = in ? static_cast<double>(A[static_cast<uint64_t>(ra) * lda + c0 + col]) : 0.0;
            sb[row * kLdsStride + col] = in ? static_cast<double>(B[static_cast<uint64_t>(rb) * ldb + c0 + col]) : 0.0;
        }
        __syncthreads();
        if constexpr (!L1) {
            for (uint32_t c = 0; c < n; ++c) {
                double va[4], vb[4];
#pragma unroll
                for (int i = 0; i < 4; ++i) va[i] = sa[(ti * 4 + i) * kLdsStride + c];
#pragma unroll
                for (int j = 0; j < 4; ++j) vb[j] = sb[(tj * 4 + j) * kLdsStride + c];
#pragma unroll
                for (int i = 0; i < 4; ++i) {
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        const double d = va[i] - vb[j];
                        const double sq = d * d;
                        acc[i][j] = acc[i][j] + sq;
                    }
                }
            }
        } else {
            uint32_t c = 0;
            // init + ((t0 + t1) + (t2 + t3)) for every complete group of four
            for (; c + 4 <= n && c0 + c + 4 <= full_end; c += 4) {
#pragma unroll
                for (int i = 0; i < 4; ++i) {
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        const double* pa = sa + (ti * 4 + i) * kLdsStride + c;
                        const double* pb = sb + (tj * 4 + j) * kLdsStride + c;
                        const double t0 = fabs(pa[0] - pb[0]);
                        const double t1 = fabs(pa[1] - pb[1]);
                        const double t2 = fabs(pa[2] - pb[2]);
                        const double t3 = fabs(pa[3] - pb[3]);
                        const double v1 = t0 + t1;
                        const double v2 = t2 + t3;
                        acc[i][j] = acc[i][j] + (v1 + v2);
                    }
                }
            }
            // then the tail one term at a time
            for (; c < n; ++c) {
#pragma unroll
                for (int i = 0; i < 4; ++i) {
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        const double t0 = fabs(sa[(ti * 4 + i) * kLdsStride + c] - sb[(tj * 4 + j) * kLdsStride + c]);
                        acc[i][j] = acc[i][j] + t0;
                    }
                }
            }
        }
    }
}

template <bool L1>
__global__ void __launch_bounds__(kPairThreads)
pair_inverse_kernel(const float* __restrict__ refs, uint64_t ld, uint32_t rows, uint32_t dim,
                    double* __restrict__ inv, uint32_t tiles)
{
    // Upper-triangular tile index -> (ta, tb) with ta <= tb.
    uint32_t t = blockIdx.x;
    uint32_t ta = 0;
    while (t >= tiles - ta) {
        t -= tiles - ta;
        ++ta;
    }
    const uint32_t tb = ta + t;
    const uint32_t a0 = ta * kPairTile;
    const uint32_t b0 = tb * kPairTile;
    double acc[4][4];
    pair_tile_fp64<L1>(refs, ld, rows, a0, refs, ld, rows, b0, dim, acc);
    const int ti = threadIdx.x / 16;
    const int tj = threadIdx.x % 16;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const uint32_t a = a0 + ti * 4 + i;
            const uint32_t b = b0 + tj * 4 + j;
            if (a < b && b < rows) {
                const double dist = acc[i][j];
                inv[static_cast<uint64_t>(a) * rows + b] = dist == 0.0 ? 0.0 : 1.0 / dist;
            }
        }
    }
}

// out[q][r] = the reference's exact diagnostic distance of query q and
// reference r (benchmark_metrics.cpp:32-41): L2 sqrt(squared_l2_dist), L1
// l1_dist, both in the CPU order above; sqrt correctly rounded.
template <bool L1>
__global__ void __launch_bounds__(kPairThreads)
distance_table_kernel(const float* __restrict__ Q, uint64_t ldq, uint32_t nq, const float* __restrict__ R,
                      uint64_t ldr, uint32_t nr, uint32_t dim, double* __restrict__ out, uint64_t ldo)
{
    const uint32_t q0 = blockIdx.y * kPairTile;
    const uint32_t r0 = blockIdx.x * kPairTile;
    double acc[4][4];
    pair_tile_fp64<L1>(Q, ldq, nq, q0, R, ldr, nr, r0, dim, acc);
    const int ti = threadIdx.x / 16;
    const int tj = threadIdx.x % 16;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const uint32_t q = q0 + ti * 4 + i;
            const uint32_t r = r0 + tj * 4 + j;
            if (q < nq && r < nr) out[static_cast<uint64_t>(q) * ldo + r] = L1 ? acc[i][j] : __dsqrt_rn(acc[i][j]);
        }
    }
}

constexpr int kColThreads = 256;
constexpr int kRowTile = 8;

template <bool L1>
__global__ void __launch_bounds__(kColThreads)
column_max_kernel(const float* __restrict__ refs, uint64_t ld, uint32_t rows, uint32_t dim,
                  const double* __restrict__ inv, unsigned long long* __restrict__ out)
{
    const uint32_t c = blockIdx.x * kColThreads + threadIdx.x;
    const uint32_t a0 = blockIdx.y * kRowTile;
    if (c >= dim) return;
    double av[kRowTile];
#pragma unroll
    for (int i = 0; i < kRowTile; ++i) {
        av[i] = static_cast<double>(refs[static_cast<uint64_t>(min(a0 + i, rows - 1U)) * ld + c]);
    }
    double best = 0.0;
    for (uint32_t b = a0 + 1; b < rows; ++b) {
        const double bv = static_cast<double>(refs[static_cast<uint64_t>(b) * ld + c]);
#pragma unroll
        for (int i = 0; i < kRowTile; ++i) {
            const uint32_t a = a0 + i;
            if (a < b) {
                const double w = inv[static_cast<uint64_t>(a) * rows + b];
                const double diff = av[i] - bv;
                double prob;
                if constexpr (L1) {
                    prob = fabs(diff) * w;
                } else {
                    const double sq = diff * diff;
                    const double x = sq * w;
                    prob = x < 1.0 ? x : 1.0;
                }
                best = best < prob ? prob : best;
            }
        }
    }
    atomicMax(out + c, static_cast<unsigned long long>(__double_as_longlong(best)));
}

__global__ void zero_u64_kernel(unsigned long long* p, uint64_t n)
{
    const uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0ULL;
}

}  // namespace

void distance_table(bool l1, const float* queries, uint64_t ldq, uint32_t nq, const float* refs, uint64_t ldr,
                    uint32_t nr, uint32_t dim, double* out, uint64_t ldo, hipStream_t stream)
{
    if (nq == 0 || nr == 0) return;
    // grid.y carries the query tiles; batches beyond 65535 of them go in
    // slices (a 64-bit counter, so the last step cannot wrap past 2^32)
    const uint64_t slice = 65535ULL * kPairTile;
    for (uint64_t q0 = 0; q0 < nq; q0 += slice) {
        const uint32_t n = static_cast<uint32_t>(nq - q0 < slice ? nq - q0 : slice);
        const float* qs = queries + q0 * ldq;
        double* os = out + q0 * ldo;
        const dim3 grid((nr + kPairTile - 1) / kPairTile, (n + kPairTile - 1) / kPairTile);
        if (l1) {
            distance_table_kernel<true><<<grid, kPairThreads, 0, stream>>>(qs, ldq, n, refs, ldr, nr, dim, os, ldo);
        } else {
            distance_table_kernel<false><<<grid, kPairThreads, 0, stream>>>(qs, ldq, n, refs, ldr, nr, dim, os, ldo);
        }
        check_launch("distance table kernel");
    }
}

void importance_probabilities(bool l1, const float* refs, uint64_t ld, uint32_t rows, uint32_t dim,
                              double* inv_scratch, unsigned long long* out_bits, hipStream_t stream)
{
    if (dim == 0) return;
    zero_u64_kernel<<<(dim + 255) / 256, 256, 0, stream>>>(out_bits, dim);
    check_launch("probability fill kernel");
    if (rows < 2) return;
    const uint32_t tiles = (rows + kPairTile - 1) / kPairTile;
    const uint64_t blocks = static_cast<uint64_t>(tiles) * (tiles + 1) / 2;
    if (blocks > 0x7fffffffULL) throw std::length_error("probability grid exceeds the launch limit");
    if (l1) {
        pair_inverse_kernel<true><<<static_cast<uint32_t>(blocks), kPairThreads, 0, stream>>>(refs, ld, rows, dim,
                                                                                           inv_scratch, tiles);
    } else {
        pair_inverse_kernel<false><<<static_cast<uint32_t>(blocks), kPairThreads, 0, stream>>>(refs, ld, rows, dim,
                                                                                            inv_scratch, tiles);
    }
    check_launch("pair distance kernel");
    const dim3 grid((dim + kColThreads - 1) / kColThreads, (rows + kRowTile - 1) / kRowTile);
    if (grid.y > 65535U) throw std::length_error("probability grid exceeds the launch limit");
    if (l1) {
        column_max_kernel<true><<<grid, kColThreads, 0, stream>>>(refs, ld, rows, dim, inv_scratch, out_bits);
    } else {
        column_max_kernel<false><<<grid, kColThreads, 0, stream>>>(refs, ld, rows, dim, inv_scratch, out_bits);
    }
    check_launch("column maximum kernel");
}

}  // namespace uhann::kern
