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

void check_hip_attr(hipError_t e)
{
    if (e != hipSuccess) {
        throw std::runtime_error(std::string("cannot raise the kernel's LDS limit: ") + hipGetErrorString(e));
    }
}

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

// Median-L1 for projection dimensions above the register networks (p > 128):
// one block per (query, run of kWideRefs references); for each reference the
// p absolute differences are padded with +inf to P2 = nextpow2(p) values in
// dynamic LDS (4 P2 bytes, the reference's per-block requirement) and sorted
// there with a bitonic network; the median is read as in the register
// kernels. Query values are read from global memory (each block reads its
// query row once per reference, from cache after the first).
constexpr int kWideThreads = 256;
constexpr uint32_t kWideRefs = 16;

__global__ void __launch_bounds__(kWideThreads)
median_l1_wide_kernel(const float* __restrict__ refs_t, uint32_t rows, const float* __restrict__ yq, uint64_t ld_y,
                      uint32_t nq, uint32_t p, uint32_t p2, unsigned long long* __restrict__ keys, int* invalid)
{
    extern __shared__ float s[];
    const uint32_t q = blockIdx.y;
    if (q >= nq) return;
    const uint32_t t = threadIdx.x;
    const float* qrow = yq + static_cast<uint64_t>(q) * ld_y;
    const uint32_t r0 = blockIdx.x * kWideRefs;
    const uint32_t r1 = min(rows, r0 + kWideRefs);
    unsigned long long best = ~0ULL;
    int bad = 0;
    for (uint32_t r = r0; r < r1; ++r) {
        for (uint32_t c = t; c < p2; c += kWideThreads) {
            float d = __builtin_inff();
            if (c < p) {
                d = __builtin_fabsf(refs_t[static_cast<uint64_t>(c) * rows + r] - qrow[c]);
                if (!__builtin_isfinite(d)) {
                    bad = 1;
                    d = __builtin_inff();
                }
            }
            s[c] = d;
        }
        __syncthreads();
        for (uint32_t size = 2; size <= p2; size <<= 1) {
            for (uint32_t stride = size >> 1; stride > 0; stride >>= 1) {
                for (uint32_t i = t; i < p2 / 2; i += kWideThreads) {
                    const uint32_t a = 2 * stride * (i / stride) + (i % stride);
                    const uint32_t b = a + stride;
                    const float x = s[a];
                    const float y = s[b];
                    const bool up = (a & size) == 0;
                    s[a] = up ? fminf(x, y) : fmaxf(x, y);
                    s[b] = up ? fmaxf(x, y) : fminf(x, y);
                }
                __syncthreads();
            }
        }
        if (t == 0) {
            const uint32_t mid = p / 2;
            float med = s[mid];
            if (p % 2 == 0) med = s[mid - 1] + (med - s[mid - 1]) * 0.5F;
            if (!__builtin_isfinite(med)) bad = 1;
            const unsigned long long key = (static_cast<unsigned long long>(__float_as_uint(med)) << 32) | r;
            best = key < best ? key : best;
        }
        __syncthreads();
    }
    if (bad) atomicExch(invalid, 1);
    if (t == 0 && best != ~0ULL) atomicMin(keys + q, best);
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

uint32_t median_register_dimension() { return 128; }

uint32_t next_pow2(uint32_t v)
{
    uint32_t p2 = 1;
    while (p2 < v) p2 <<= 1;
    return p2;
}

// The reference's rule (cuda_median_l1_scan_index.cu:314-336): the padded
// sort of 4 nextpow2(p) bytes must fit the device's per-block shared memory.
bool median_supported(uint64_t p)
{
    if (p == 0 || p > (1ULL << 30)) return false;
    if (p <= median_register_dimension()) return true;
    int dev = 0;
    hipDeviceProp_t prop{};
    if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess) return false;
    return static_cast<uint64_t>(next_pow2(static_cast<uint32_t>(p))) * sizeof(float) <= prop.sharedMemPerBlock;
}
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
    if (!median_supported(p)) {
        throw std::invalid_argument("median-L1 projection dimension exceeds the device's per-block LDS for its sort");
    }
    if (p > median_register_dimension()) {
        const uint32_t p2 = next_pow2(p);
        const size_t lds = static_cast<size_t>(p2) * sizeof(float);
        if (lds > 65536) {
            check_hip_attr(hipFuncSetAttribute(reinterpret_cast<const void*>(median_l1_wide_kernel),
                                               hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds)));
        }
        for (uint32_t q0 = 0; q0 < nq; q0 += 65535U) {
            const uint32_t n = min(65535U, nq - q0);
            const dim3 grid((rows + kWideRefs - 1) / kWideRefs, n);
            median_l1_wide_kernel<<<grid, kWideThreads, lds, stream>>>(refs_t, rows, yq + static_cast<uint64_t>(q0) * ld_y,
                                                                       ld_y, n, p, p2, keys + q0, invalid);
            check_launch("wide median-L1 kernel");
        }
        return;
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
