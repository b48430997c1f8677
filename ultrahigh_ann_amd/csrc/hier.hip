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

#include "hier.hpp"

namespace uhann::kern {
namespace {

constexpr int kProjThreads = 256;
constexpr int kProjQ = 4;
constexpr int kProjP = 8;

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
    const int t = threadIdx.x;
    const float* xr[kProjQ];
    const float* pr[kProjP];
#pragma unroll
    for (int i = 0; i < kProjQ; ++i) xr[i] = x + static_cast<uint64_t>(min(q0 + i, nq - 1U)) * ld_x;
#pragma unroll
    for (int j = 0; j < kProjP; ++j) pr[j] = proj + static_cast<uint64_t>(min(o0 + j, p - 1U)) * ld_proj;
    float acc[kProjQ][kProjP];
#pragma unroll
    for (int i = 0; i < kProjQ; ++i)
#pragma unroll
        for (int j = 0; j < kProjP; ++j) acc[i][j] = 0.0F;
    for (uint32_t c = t; c < k; c += kProjThreads) {
        float xv[kProjQ], pv[kProjP];
#pragma unroll
        for (int i = 0; i < kProjQ; ++i) xv[i] = xr[i][c];
#pragma unroll
        for (int j = 0; j < kProjP; ++j) pv[j] = pr[j][c];
#pragma unroll
        for (int i = 0; i < kProjQ; ++i)
#pragma unroll
            for (int j = 0; j < kProjP; ++j) acc[i][j] = __builtin_fmaf(xv[i], pv[j], acc[i][j]);
    }
    __shared__ float warp_sums[8][kProjQ * kProjP];
#pragma unroll
    for (int i = 0; i < kProjQ; ++i) {
#pragma unroll
        for (int j = 0; j < kProjP; ++j) {
            float v = acc[i][j];
            for (int off = 16; off >= 1; off /= 2) v = v + __shfl_down(v, off);
            if (t % 32 == 0) warp_sums[t / 32][i * kProjP + j] = v;
        }
    }
    __syncthreads();
    if (t < kProjQ * kProjP) {
        const int i = t / kProjP;
        const int j = t % kProjP;
        const uint32_t q = q0 + i;
        const uint32_t o = o0 + j;
        if (q < nq && o < p) {
            float w[8];
#pragma unroll
            for (int g = 0; g < 8; ++g) w[g] = (warp_sums[g][t] + 0.0F) + 0.0F;
            const float s = ((w[0] + w[4]) + (w[2] + w[6])) + ((w[1] + w[5]) + (w[3] + w[7]));
            if (!__builtin_isfinite(s)) atomicExch(invalid, 1);
            y[static_cast<uint64_t>(q) * ld_y + o] = s;
        }
    }
}

template <int N>
__device__ __forceinline__ void bitonic_sort(float (&v)[N])
{
#pragma unroll
    for (int size = 2; size <= N; size <<= 1) {
#pragma unroll
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
#pragma unroll
            for (int i = 0; i < N; ++i) {
                const int l = i ^ stride;
                if (l > i) {
                    const float a = v[i];
                    const float b = v[l];
                    const float lo = fminf(a, b);
                    const float hi = fmaxf(a, b);
                    if ((i & size) == 0) {
                        v[i] = lo;
                        v[l] = hi;
                    } else {
                        v[i] = hi;
                        v[l] = lo;
                    }
                }
            }
        }
    }
}

constexpr int kMedianThreads = 256;

template <int PMAX>
__global__ void __launch_bounds__(kMedianThreads)
median_l1_kernel(const float* __restrict__ refs_t, uint32_t rows, const float* __restrict__ yq,
                 uint64_t ld_y, uint32_t nq, uint32_t p, unsigned long long* __restrict__ keys,
                 int* invalid)
{
    __shared__ float qv[PMAX];
    __shared__ unsigned long long best_key[kMedianThreads / 64];
    const uint32_t q = blockIdx.y;
    if (q >= nq) return;
    const int t = threadIdx.x;
    for (uint32_t c = t; c < p; c += kMedianThreads) qv[c] = yq[static_cast<uint64_t>(q) * ld_y + c];
    __syncthreads();
    const uint32_t r = blockIdx.x * kMedianThreads + t;
    unsigned long long key = ~0ULL;
    if (r < rows) {
        float v[PMAX];
        bool bad = false;
#pragma unroll
        for (int c = 0; c < PMAX; ++c) {
            if (static_cast<uint32_t>(c) < p) {
                float d = __builtin_fabsf(refs_t[static_cast<uint64_t>(c) * rows + r] - qv[c]);
                if (!__builtin_isfinite(d)) {
                    bad = true;
                    d = __builtin_inff();
                }
                v[c] = d;
            } else {
                v[c] = __builtin_inff();
            }
        }
        bitonic_sort<PMAX>(v);
        const uint32_t mid = p / 2;
        float med = 0.0F;
        float lower = 0.0F;
#pragma unroll
        for (int c = 0; c < PMAX; ++c) {
            if (static_cast<uint32_t>(c) == mid) med = v[c];
            if (static_cast<uint32_t>(c) + 1U == mid) lower = v[c];
        }
        if (p % 2 == 0) med = lower + (med - lower) * 0.5F;
        if (!__builtin_isfinite(med)) bad = true;
        if (bad) atomicExch(invalid, 1);
        key = (static_cast<unsigned long long>(__float_as_uint(med)) << 32) | r;
    }
    for (int off = 32; off >= 1; off /= 2) {
        const unsigned long long other = __shfl_down(key, off);
        key = other < key ? other : key;
    }
    if (t % 64 == 0) best_key[t / 64] = key;
    __syncthreads();
    if (t == 0) {
        unsigned long long b = best_key[0];
        for (int w = 1; w < kMedianThreads / 64; ++w) b = best_key[w] < b ? best_key[w] : b;
        atomicMin(keys + q, b);
    }
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
        if (p <= 8) {
            median_l1_kernel<8><<<grid, kMedianThreads, 0, stream>>>(refs_t, rows, yb, ld_y, n, p, kb, invalid);
        } else if (p <= 16) {
            median_l1_kernel<16><<<grid, kMedianThreads, 0, stream>>>(refs_t, rows, yb, ld_y, n, p, kb, invalid);
        } else if (p <= 32) {
            median_l1_kernel<32><<<grid, kMedianThreads, 0, stream>>>(refs_t, rows, yb, ld_y, n, p, kb, invalid);
        } else if (p <= 64) {
            median_l1_kernel<64><<<grid, kMedianThreads, 0, stream>>>(refs_t, rows, yb, ld_y, n, p, kb, invalid);
        } else {
            median_l1_kernel<128><<<grid, kMedianThreads, 0, stream>>>(refs_t, rows, yb, ld_y, n, p, kb, invalid);
        }
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
