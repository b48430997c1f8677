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
#include <vector>

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


// Pair-per-thread form for short rows (flat |I|, hierarchical p): the
// 256-chain tree costs 255 adds per pair whatever the row length, and in the
// tiled kernel above it is a cross-lane tree, which dominates when a chain
// holds one or two terms. Here each thread owns one reference row and TQ
// queries and evaluates the whole reference order in registers: for each
// group w of 32 chains, chain 32w+i sums columns 32w+i, 32w+i+256, ... in
// order, then the 32-lane tree (offsets 16..1) gives W_w, and
// ((W0+W4)+(W2+W6))+((W1+W5)+(W3+W7)) the distance -- the same additions as
// the reference, so the same bits. References are read from a transposed
// (dim x rows) copy, so a wave's loads are contiguous; queries (and L1
// weights) sit in LDS and are broadcast.
template <int METRIC, bool WEIGHTED, bool WFMA, int TQ>
__global__ void __launch_bounds__(256)
scan_pairs_kernel(const float* __restrict__ refs_t, uint64_t ld_t, uint32_t rows,
                  const float* __restrict__ queries, uint64_t ld_q, uint32_t nq,
                  const float* __restrict__ weights, uint32_t dim, unsigned long long* __restrict__ keys)
{
    extern __shared__ float smem[];
    float* qs = smem;              // TQ x dim
    float* ws = smem + TQ * dim;   // dim (weighted L1)
    const uint32_t q0 = blockIdx.y * TQ;
    const int t = threadIdx.x;
    for (uint32_t e = t; e < TQ * dim; e += 256) {
        const uint32_t qi = e / dim;
        const uint32_t c = e % dim;
        qs[e] = queries[static_cast<uint64_t>(min(q0 + qi, nq - 1U)) * ld_q + c];
    }
    if constexpr (WEIGHTED) {
        for (uint32_t c = t; c < dim; c += 256) ws[c] = weights[c];
    }
    __syncthreads();
    const uint32_t r = blockIdx.x * 256 + t;
    const uint32_t rr = min(r, rows - 1U);
    // Groups are visited in bit-reversed order 0,4 | 2,6 | 1,5 | 3,7, so pass
    // p of 4 produces (W0+W4), (W2+W6), (W1+W5), (W3+W7) in turn and the
    // distance is (P0 + P1) + (P2 + P3): the reference's final tree with only
    // four partial sums live (the loops stay rolled to keep registers low).
    float P[TQ][4];
#pragma unroll
    for (int q = 0; q < TQ; ++q)
#pragma unroll
        for (int k = 0; k < 4; ++k) P[q][k] = 0.0F;
#pragma unroll 1
    for (int pass = 0; pass < 4; ++pass) {
        float pair_sum[TQ];
#pragma unroll 1
        for (int h = 0; h < 2; ++h) {
            const int g = 2 * pass + h;
            const int w = ((g & 1) << 2) | (g & 2) | ((g & 4) >> 2);  // bit reversal of g
            float v[TQ][32];
#pragma unroll
            for (int i = 0; i < 32; ++i)
#pragma unroll
                for (int q = 0; q < TQ; ++q) v[q][i] = 0.0F;
            // Term k of chain 32w+i is column 32w + i + 256k: all 32 chains of
            // the group load term k together (32 independent loads in flight),
            // and each chain still adds its terms in increasing k.
            for (uint32_t base = static_cast<uint32_t>(32 * w); base < dim; base += kChains) {
                float rv[32];
#pragma unroll
                for (int i = 0; i < 32; ++i) {
                    const uint32_t c = min(base + i, dim - 1U);
                    rv[i] = refs_t[static_cast<uint64_t>(c) * ld_t + rr];
                }
#pragma unroll
                for (int i = 0; i < 32; ++i) {
                    const uint32_t c = base + i;
                    if (c < dim) {
                        const float wv = WEIGHTED ? ws[c] : 0.0F;
#pragma unroll
                        for (int q = 0; q < TQ; ++q) {
                            v[q][i] = accumulate<METRIC, WEIGHTED, WFMA>(v[q][i], rv[i], qs[q * dim + c], wv);
                        }
                    }
                }
            }
#pragma unroll
            for (int q = 0; q < TQ; ++q) {
#pragma unroll
                for (int off = 16; off >= 1; off /= 2) {
#pragma unroll
                    for (int i = 0; i < off; ++i) v[q][i] = v[q][i] + v[q][i + off];
                }
                pair_sum[q] = h == 0 ? v[q][0] : pair_sum[q] + v[q][0];
            }
        }
#pragma unroll
        for (int q = 0; q < TQ; ++q)
#pragma unroll
            for (int k = 0; k < 4; ++k) P[q][k] = pass == k ? pair_sum[q] : P[q][k];
    }
    __shared__ unsigned long long best[TQ][4];
#pragma unroll
    for (int q = 0; q < TQ; ++q) {
        const float dist = (P[q][0] + P[q][1]) + (P[q][2] + P[q][3]);
        unsigned long long key = r < rows
                                     ? (static_cast<unsigned long long>(__float_as_uint(dist)) << 32) | r
                                     : ~0ULL;
        for (int off = 32; off >= 1; off /= 2) {
            const unsigned long long o = __shfl_down(key, off);
            key = o < key ? o : key;
        }
        if (t % 64 == 0) best[q][t / 64] = key;
    }
    __syncthreads();
    if (t < TQ && q0 + t < nq) {
        unsigned long long b = best[t][0];
        for (int w = 1; w < 4; ++w) b = best[t][w] < b ? best[t][w] : b;
        atomicMin(keys + q0 + t, b);
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
    UHANN_TILE(1, 1, 4)
    UHANN_TILE(1, 2, 8)
    UHANN_TILE(1, 4, 8)
    UHANN_TILE(1, 16, 4)
    UHANN_TILE(4, 1, 4)
    UHANN_TILE(4, 2, 4)
    UHANN_TILE(4, 4, 4)
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

const std::vector<ScanTile>& scan_tiles()
{
    static const std::vector<ScanTile> tiles = {
        {1, 8, 8, 8},  {1, 12, 12, 8}, {1, 16, 8, 4}, {1, 8, 16, 2}, {1, 4, 16, 1}, {1, 2, 16, 1},
        {1, 1, 16, 1}, {4, 1, 8, 1},   {4, 2, 8, 1},  {4, 4, 8, 1},  {4, 8, 4, 2},  {1, 1, 4, 1},
        {1, 2, 8, 1},  {1, 4, 8, 1},   {1, 16, 4, 4}, {4, 1, 4, 1},  {4, 2, 4, 1},  {4, 4, 4, 1},
    };
    return tiles;
}

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

uint32_t scan_pairs_max_dim() { return 4096; }

void scan_pairs(Metric metric, bool weight_fma, const float* refs_t, uint64_t ld_t, uint32_t rows,
                const float* queries, uint64_t ld_q, uint32_t nq, const float* weights, uint32_t dim,
                unsigned long long* keys, int tq, hipStream_t stream)
{
    if (nq == 0 || rows == 0) return;
    if (dim > scan_pairs_max_dim()) throw std::invalid_argument("pair scan supports dim <= 4096");
    const bool weighted = weights != nullptr;
    const size_t lds = (static_cast<size_t>(tq) * dim + (weighted ? dim : 0)) * sizeof(float);
    const dim3 grid((rows + 255) / 256, (nq + tq - 1) / tq);
    if (grid.y > 65535U) throw std::length_error("pair scan batch exceeds the grid limit");
#define UHANN_PAIRS(M, W, F, Q)                                                                   \
    hipLaunchKernelGGL((scan_pairs_kernel<M, W, F, Q>), grid, dim3(256), lds, stream, refs_t, ld_t, \
                       rows, queries, ld_q, nq, weights, dim, keys)
    const int m = metric == Metric::l2 ? 0 : 1;
    if (tq == 1) {
        if (m == 0) UHANN_PAIRS(0, false, false, 1);
        else if (!weighted) UHANN_PAIRS(1, false, false, 1);
        else if (weight_fma) UHANN_PAIRS(1, true, true, 1);
        else UHANN_PAIRS(1, true, false, 1);
    } else if (tq == 2) {
        if (m == 0) UHANN_PAIRS(0, false, false, 2);
        else if (!weighted) UHANN_PAIRS(1, false, false, 2);
        else if (weight_fma) UHANN_PAIRS(1, true, true, 2);
        else UHANN_PAIRS(1, true, false, 2);
    } else if (tq == 4) {
        if (m == 0) UHANN_PAIRS(0, false, false, 4);
        else if (!weighted) UHANN_PAIRS(1, false, false, 4);
        else if (weight_fma) UHANN_PAIRS(1, true, true, 4);
        else UHANN_PAIRS(1, true, false, 4);
    } else {
        throw std::invalid_argument("pair scan supports 1, 2 or 4 queries per block");
    }
#undef UHANN_PAIRS
    check_launch("pair scan kernel");
}

}  // namespace uhann::kern
