// Direct FP32 L2 / L1 scan with a fused lowest-index argmin, for gfx950.
//
// Arithmetic contract (what makes the distances bit-identical to the
// reference's CUDA kernels K2/K5, cuda_dense_l2_scan_index.cu:295-324 and
// cuda_dense_l1_scan_index.cu:218-247): for every (query, reference) pair the
// reference sums 256 independent chains, chain t taking columns
// t, t+256, t+512, ... in increasing order with fmaf(d, d, acc) (L2) or
// acc + |d| / fmaf(w, |d|, acc) (L1), then adds the chains with a fixed tree:
// 32-lane shfl_down (offsets 16,8,4,2,1) inside each group of 32 chains,
// then ((W0+W4)+(W2+W6)) + ((W1+W5)+(W3+W7)) over the eight group sums.
//
// The layout here is ours. A block of 256/RPT threads covers all 256 chains:
// thread t owns the RPT consecutive chains RPT*t .. RPT*t+RPT-1 (one vector
// load per row per 256-column step), and holds a TQ x TR register tile of
// pairs, so each loaded reference value is reused TQ times and each query
// value TR times. The chain order is kept exactly; the tree is replayed with
// wave64 shuffles (offsets that stay inside a group of 32 chains) plus an
// in-thread step for the chains a thread owns, and the eight group sums meet
// in LDS. The argmin never leaves the chip: each block reduces its tile and
// publishes one 64-bit key per query with atomicMin. Distances are >= 0, so
// the IEEE bit pattern orders them, and (bits << 32 | row) orders by distance
// first and then by the lower row -- the reference's candidate_is_better
// (cuda_dense_l2_scan_index.cu:221-229).
//
// Work distribution: blocks are remapped so that each XCD (blocks are dealt
// round-robin over 8 XCDs) gets a contiguous run of logical tiles, and
// logical tiles sweep all reference tiles for `group_q` query tiles at a
// time, so co-resident blocks share reference and query rows in their XCD's L2.

#include <hip/hip_runtime.h>

#include <stdexcept>
#include <string>

#include "kernels.hpp"

namespace uhann::kern {
namespace {

constexpr int kChains = 256;
constexpr int kGroups = 8;  // groups of 32 chains

template <int RPT>
__device__ __forceinline__ void load_vec(const float* __restrict__ p, float (&v)[RPT])
{
    if constexpr (RPT == 1) {
        v[0] = p[0];
    } else if constexpr (RPT == 2) {
        const float2 x = *reinterpret_cast<const float2*>(p);
        v[0] = x.x;
        v[1] = x.y;
    } else {
        static_assert(RPT == 4, "RPT must be 1, 2 or 4");
        const float4 x = *reinterpret_cast<const float4*>(p);
        v[0] = x.x;
        v[1] = x.y;
        v[2] = x.z;
        v[3] = x.w;
    }
}

template <int METRIC, bool WEIGHTED, bool WFMA>
__device__ __forceinline__ float accumulate(float acc, float r, float q, float w)
{
    const float d = r - q;
    if constexpr (METRIC == 0) {
        return __builtin_fmaf(d, d, acc);
    } else {
        const float ad = __builtin_fabsf(d);
        if constexpr (!WEIGHTED) {
            return acc + ad;
        } else if constexpr (WFMA) {
            return __builtin_fmaf(w, ad, acc);
        } else {
            const float prod = w * ad;
            return acc + prod;
        }
    }
}

// Blocks are dealt to the 8 XCDs round-robin; give XCD x the contiguous
// logical range [start_x, start_x + count_x).
__device__ __forceinline__ uint32_t xcd_logical_block(uint32_t b, uint32_t nblocks)
{
    const uint32_t per = nblocks / 8U;
    const uint32_t rem = nblocks % 8U;
    const uint32_t x = b % 8U;
    const uint32_t k = b / 8U;
    return x * per + (x < rem ? x : rem) + k;
}

template <int METRIC, bool WEIGHTED, bool WFMA, int RPT, int TQ, int TR>
__global__ void __launch_bounds__(kChains / RPT)
scan_direct_kernel(ScanArgs a, uint32_t tiles_q, uint32_t tiles_r, uint32_t group_q)
{
    constexpr int NT = kChains / RPT;     // threads per block
    constexpr int GROUP = 32 / RPT;       // threads per group of 32 chains
    constexpr int NP = TQ * TR;           // pairs per block

    const uint32_t L = xcd_logical_block(blockIdx.x, tiles_q * tiles_r);
    const uint32_t span = group_q * tiles_r;
    const uint32_t gid = L / span;
    const uint32_t first_q = gid * group_q;
    const uint32_t gsize = min(group_q, tiles_q - first_q);
    const uint32_t within = L - gid * span;
    const uint32_t q0 = (first_q + within % gsize) * TQ;
    const uint32_t r0 = (within / gsize) * TR;

    const int t = threadIdx.x;
    const float* rrow[TR];
    const float* qrow[TQ];
#pragma unroll
    for (int j = 0; j < TR; ++j) {
        rrow[j] = a.refs + static_cast<uint64_t>(min(r0 + j, a.rows - 1U)) * a.ld_ref;
    }
#pragma unroll
    for (int i = 0; i < TQ; ++i) {
        qrow[i] = a.queries + static_cast<uint64_t>(min(q0 + i, a.nq - 1U)) * a.ld_q;
    }

    float acc[TQ][TR][RPT];
#pragma unroll
    for (int i = 0; i < TQ; ++i)
#pragma unroll
        for (int j = 0; j < TR; ++j)
#pragma unroll
            for (int k = 0; k < RPT; ++k) acc[i][j][k] = 0.0F;

    // One 256-column step at column `col` (this thread's first chain). In the
    // tail step a vector may straddle dim (rows are padded to a multiple of 4
    // floats): components at or past dim are skipped, so pad values are
    // never added.
    auto step = [&](uint32_t col, bool masked) {
        float rv[TR][RPT];
        float qv[TQ][RPT];
        float wv[RPT];
#pragma unroll
        for (int j = 0; j < TR; ++j) load_vec<RPT>(rrow[j] + col, rv[j]);
#pragma unroll
        for (int i = 0; i < TQ; ++i) load_vec<RPT>(qrow[i] + col, qv[i]);
        if constexpr (WEIGHTED) {
            load_vec<RPT>(a.weights + col, wv);
        } else {
#pragma unroll
            for (int k = 0; k < RPT; ++k) wv[k] = 0.0F;
        }
#pragma unroll
        for (int k = 0; k < RPT; ++k) {
            if (masked && col + static_cast<uint32_t>(k) >= a.dim) continue;
#pragma unroll
            for (int i = 0; i < TQ; ++i)
#pragma unroll
                for (int j = 0; j < TR; ++j)
                    acc[i][j][k] = accumulate<METRIC, WEIGHTED, WFMA>(acc[i][j][k], rv[j][k],
                                                                     qv[i][k], wv[k]);
        }
    };

    const uint32_t col_t = static_cast<uint32_t>(RPT * t);
    const uint32_t full_end = (a.dim / kChains) * kChains;
#pragma unroll 2
    for (uint32_t c0 = 0; c0 < full_end; c0 += kChains) step(c0 + col_t, false);
    if (full_end + col_t < a.dim) step(full_end + col_t, true);

    // Replay the reference tree. Cross-thread steps first (chain offsets 16..RPT
    // are thread offsets GROUP/2..1), then the chains inside the thread.
    __shared__ float group_sums[kGroups][NP];
    __shared__ float pair_dist[NP];
#pragma unroll
    for (int i = 0; i < TQ; ++i) {
#pragma unroll
        for (int j = 0; j < TR; ++j) {
            float v[RPT];
#pragma unroll
            for (int k = 0; k < RPT; ++k) v[k] = acc[i][j][k];
#pragma unroll
            for (int off = GROUP / 2; off >= 1; off /= 2) {
#pragma unroll
                for (int k = 0; k < RPT; ++k) v[k] = v[k] + __shfl_down(v[k], off);
            }
            float w;
            if constexpr (RPT == 1) {
                w = v[0];
            } else if constexpr (RPT == 2) {
                w = v[0] + v[1];
            } else {
                w = (v[0] + v[2]) + (v[1] + v[3]);
            }
            if (t % GROUP == 0) group_sums[t / GROUP][i * TR + j] = w;
        }
    }
    __syncthreads();
    for (int p = t; p < NP; p += NT) {
        const float g0 = group_sums[0][p], g1 = group_sums[1][p];
        const float g2 = group_sums[2][p], g3 = group_sums[3][p];
        const float g4 = group_sums[4][p], g5 = group_sums[5][p];
        const float g6 = group_sums[6][p], g7 = group_sums[7][p];
        // Lanes 8..31 of the reference's second warp_sum hold +0.0f; adding
        // them leaves a non-negative sum unchanged, so they are skipped.
        pair_dist[p] = ((g0 + g4) + (g2 + g6)) + ((g1 + g5) + (g3 + g7));
    }
    __syncthreads();
    if (t < TQ) {
        const uint32_t q = q0 + static_cast<uint32_t>(t);
        if (q < a.nq) {
            unsigned long long best = ~0ULL;
#pragma unroll
            for (int j = 0; j < TR; ++j) {
                const uint32_t r = r0 + static_cast<uint32_t>(j);
                if (r < a.rows) {
                    const unsigned long long key =
                        (static_cast<unsigned long long>(__float_as_uint(pair_dist[t * TR + j]))
                         << 32) |
                        r;
                    best = key < best ? key : best;
                }
            }
            atomicMin(a.keys + q, best);
        }
    }
}

__global__ void fill_keys_kernel(unsigned long long* keys, uint32_t n)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) keys[i] = ~0ULL;
}

__global__ void finalize_keys_kernel(const unsigned long long* keys, uint32_t n,
                                     uint64_t* out_idx, float* out_dist)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        const unsigned long long key = keys[i];
        if (out_idx != nullptr) out_idx[i] = key & 0xffffffffULL;
        if (out_dist != nullptr) out_dist[i] = __uint_as_float(static_cast<uint32_t>(key >> 32));
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

using ScanFn = void (*)(ScanArgs, uint32_t, uint32_t, uint32_t);

template <int METRIC, bool WEIGHTED, bool WFMA>
ScanFn pick(int rpt, int tq, int tr)
{
#define UHANN_TILE(R, Q, T)                                                       \
    if (rpt == R && tq == Q && tr == T) return scan_direct_kernel<METRIC, WEIGHTED, WFMA, R, Q, T>;
    UHANN_TILE(1, 8, 8)
    UHANN_TILE(1, 12, 12)
    UHANN_TILE(1, 16, 8)
    UHANN_TILE(1, 8, 16)
    UHANN_TILE(1, 4, 16)
    UHANN_TILE(1, 2, 16)
    UHANN_TILE(1, 1, 16)
    UHANN_TILE(4, 1, 8)
    UHANN_TILE(4, 2, 8)
    UHANN_TILE(4, 4, 8)
    UHANN_TILE(4, 8, 4)
#undef UHANN_TILE
    return nullptr;
}

ScanFn select_scan(Metric metric, bool weighted, bool wfma, const ScanTile& t)
{
    if (metric == Metric::l2) return pick<0, false, false>(t.rpt, t.tq, t.tr);
    if (!weighted) return pick<1, false, false>(t.rpt, t.tq, t.tr);
    if (wfma) return pick<1, true, true>(t.rpt, t.tq, t.tr);
    return pick<1, true, false>(t.rpt, t.tq, t.tr);
}

}  // namespace

bool scan_tile_supported(const ScanTile& tile)
{
    return select_scan(Metric::l2, false, false, tile) != nullptr && tile.group_q >= 1;
}

void fill_keys(unsigned long long* keys, uint32_t n, hipStream_t stream)
{
    if (n == 0) return;
    fill_keys_kernel<<<(n + 255) / 256, 256, 0, stream>>>(keys, n);
    check_launch("key initialisation kernel");
}

void finalize_keys(const unsigned long long* keys, uint32_t n, uint64_t* out_idx,
                   float* out_dist, hipStream_t stream)
{
    if (n == 0) return;
    finalize_keys_kernel<<<(n + 255) / 256, 256, 0, stream>>>(keys, n, out_idx, out_dist);
    check_launch("argmin finalisation kernel");
}

void scan_direct(Metric metric, bool weight_fma, const ScanArgs& args, const ScanTile& tile,
                 hipStream_t stream)
{
    const ScanFn fn = select_scan(metric, args.weights != nullptr, weight_fma, tile);
    if (fn == nullptr) throw std::invalid_argument("unsupported scan tile");
    if (args.nq == 0 || args.rows == 0) return;
    const uint32_t tiles_q = (args.nq + tile.tq - 1) / tile.tq;
    const uint32_t tiles_r = (args.rows + tile.tr - 1) / tile.tr;
    const uint64_t blocks = static_cast<uint64_t>(tiles_q) * tiles_r;
    if (blocks > 0x7fffffffULL) throw std::length_error("scan grid exceeds the launch limit");
    const uint32_t group_q = static_cast<uint32_t>(tile.group_q < 1 ? 1 : tile.group_q);
    hipLaunchKernelGGL(fn, dim3(static_cast<uint32_t>(blocks)), dim3(kChains / tile.rpt), 0,
                       stream, args, tiles_q, tiles_r, group_q);
    check_launch("direct scan kernel");
}

}  // namespace uhann::kern
