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

#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "bodies.hpp"
#include "kernels.hpp"

namespace uhann::kern {
namespace {

using body::kChains;

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
    const uint32_t L = xcd_logical_block(blockIdx.x, tiles_q * tiles_r);
    const uint32_t span = group_q * tiles_r;
    const uint32_t gid = L / span;
    const uint32_t first_q = gid * group_q;
    const uint32_t gsize = min(group_q, tiles_q - first_q);
    const uint32_t within = L - gid * span;
    const uint32_t q0 = (first_q + within % gsize) * TQ;
    const uint32_t r0 = (within / gsize) * TR;

    body::scan_tile<METRIC, WEIGHTED, WFMA, RPT, TQ, TR>(a, q0, r0);
}

// Shared-query tile (body::scan_tile_sq): G reference groups of TR rows and TQ
// queries per block of 256 G threads; query tiles are the fast index, so the
// blocks of one reference tile run together (XCD-contiguous ranges).
template <int METRIC, bool WEIGHTED, bool WFMA, int TQ, int TR, int G>
__global__ void __launch_bounds__(kChains * G)
scan_sq_kernel(ScanArgs a, uint32_t tiles_q, uint32_t tiles_r)
{
    const uint32_t L = xcd_logical_block(blockIdx.x, tiles_q * tiles_r);
    const uint32_t q0 = (L % tiles_q) * TQ;
    const uint32_t r0 = (L / tiles_q) * (TR * G);
    body::scan_tile_sq<METRIC, WEIGHTED, WFMA, TQ, TR, G>(a, q0, r0, a.stream_refs);
}

// Pair-per-thread form for short rows (body::scan_pairs): grid.x covers
// the references 256 at a time, grid.y the queries TQ at a time.
template <int METRIC, bool WEIGHTED, bool WFMA, int TQ>
__global__ void __launch_bounds__(256)
scan_pairs_kernel(const float* __restrict__ refs_t, uint64_t ld_t, uint32_t rows,
                  const float* __restrict__ queries, uint64_t ld_q, uint32_t nq,
                  const float* __restrict__ weights, uint32_t dim, unsigned long long* __restrict__ keys)
{
    extern __shared__ float smem[];
    body::scan_pairs<METRIC, WEIGHTED, WFMA, TQ>(refs_t, ld_t, rows, queries, ld_q, nq, weights, dim, keys,
                                                 blockIdx.y * TQ, blockIdx.x, smem);
}

// Chain-class form of the pair-per-thread scan (body::scan_pairs_cls): grid
// as scan_pairs_kernel; WHOLE (dim <= body::kClsWholeRowDim) stages the whole
// query rows in LDS, otherwise two groups' columns per pass.
template <int METRIC, bool WEIGHTED, bool WFMA, int TQ, bool WHOLE>
__global__ void __launch_bounds__(256)
scan_pairs_cls_kernel(const float* __restrict__ refs_t, uint64_t ld_t, uint32_t rows,
                      const float* __restrict__ queries, uint64_t ld_q, uint32_t nq,
                      const float* __restrict__ weights, uint32_t dim, unsigned long long* __restrict__ keys)
{
    extern __shared__ float smem[];
    body::scan_pairs_cls<METRIC, WEIGHTED, WFMA, TQ, WHOLE>(refs_t, ld_t, rows, queries, ld_q, nq, weights, dim,
                                                            keys, blockIdx.y * TQ, blockIdx.x, smem);
}

// Device-side escalation of the mixed-precision strategy (no host round
// trip): the queries its selection listed (list[0 .. *count)) are copied into
// a contiguous block (their keys reset on the way), scanned exactly by a
// fixed grid that walks (8-query tile, 8-row tile) work items with the
// direct scan's tile body, and the results are written back to the listed
// positions. With nothing listed every block returns at once.
constexpr int kEscTQ = 8;
constexpr int kEscTR = 8;
constexpr uint32_t kEscBlocks = 1024;

__global__ void __launch_bounds__(256)
escalate_gather_kernel(const float* __restrict__ src, uint64_t ld_src, uint32_t dim,
                       const unsigned int* __restrict__ count, const unsigned int* __restrict__ list, uint32_t cap,
                       float* __restrict__ dst, uint64_t ld_dst, unsigned long long* __restrict__ keys)
{
    const uint32_t n = min(*count, cap);
    for (uint32_t i = blockIdx.x; i < n; i += gridDim.x) {
        const float* s = src + static_cast<uint64_t>(list[i]) * ld_src;
        float* d = dst + static_cast<uint64_t>(i) * ld_dst;
        for (uint32_t c = threadIdx.x; c < ld_dst; c += blockDim.x) d[c] = c < dim ? s[c] : 0.0F;
        if (threadIdx.x == 0) keys[i] = ~0ULL;
    }
}

__global__ void __launch_bounds__(kChains)
escalate_scan_kernel(ScanArgs a, const unsigned int* __restrict__ count, uint32_t cap)
{
    a.nq = min(*count, cap);
    const uint32_t tiles_r = (a.rows + kEscTR - 1) / kEscTR;
    const uint64_t work = static_cast<uint64_t>((a.nq + kEscTQ - 1) / kEscTQ) * tiles_r;
    for (uint64_t w = blockIdx.x; w < work; w += gridDim.x) {
        __syncthreads();
        body::scan_tile<0, false, false, 1, kEscTQ, kEscTR>(a, static_cast<uint32_t>(w / tiles_r) * kEscTQ,
                                                            static_cast<uint32_t>(w % tiles_r) * kEscTR);
    }
}

__global__ void __launch_bounds__(256)
escalate_finalize_kernel(const unsigned long long* __restrict__ keys, const unsigned int* __restrict__ count,
                         const unsigned int* __restrict__ list, uint32_t cap, uint64_t* __restrict__ out_idx,
                         float* __restrict__ out_dist)
{
    const uint32_t n = min(*count, cap);
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const unsigned long long key = keys[i];
        const uint32_t q = list[i];
        out_idx[q] = key & 0xffffffffULL;
        out_dist[q] = __uint_as_float(static_cast<uint32_t>(key >> 32));
    }
}

// Thread-per-query form for short rows (body::scan_short): grid.x covers
// the references kShortRefs at a time, grid.y the queries 256 at a time.
template <int METRIC, bool WEIGHTED, bool WFMA, int DMAX>
__global__ void __launch_bounds__(body::kShortThreads)
scan_short_kernel(const float* __restrict__ refs, uint64_t ld_ref, uint32_t rows, const float* __restrict__ queries,
                  uint64_t ld_q, uint32_t nq, const float* __restrict__ weights, uint32_t dim,
                  unsigned long long* __restrict__ keys)
{
    body::scan_short<METRIC, WEIGHTED, WFMA, DMAX>(refs, ld_ref, rows, queries, ld_q, nq, weights, dim, keys,
                                                   blockIdx.y * body::kShortThreads, blockIdx.x * body::kShortRefs);
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
using SqFn = void (*)(ScanArgs, uint32_t, uint32_t);

// Shared-query tiles: ScanTile{2, TQ, TR, G}.
template <int METRIC, bool WEIGHTED, bool WFMA>
SqFn pick_sq(int tq, int tr, int g)
{
#define UHANN_SQ(Q, T, GG)                                                        \
    if (tq == Q && tr == T && g == GG) return scan_sq_kernel<METRIC, WEIGHTED, WFMA, Q, T, GG>;
    UHANN_SQ(16, 4, 4)
    UHANN_SQ(8, 4, 4)
    UHANN_SQ(8, 4, 2)
#undef UHANN_SQ
    return nullptr;
}

SqFn select_sq(Metric metric, bool weighted, bool wfma, const ScanTile& t)
{
    if (t.rpt != 2) return nullptr;
    if (metric == Metric::l2) return pick_sq<0, false, false>(t.tq, t.tr, t.group_q);
    if (!weighted) return pick_sq<1, false, false>(t.tq, t.tr, t.group_q);
    if (wfma) return pick_sq<1, true, true>(t.tq, t.tr, t.group_q);
    return pick_sq<1, true, false>(t.tq, t.tr, t.group_q);
}

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
    UHANN_TILE(4, 1, 2)
    UHANN_TILE(4, 1, 1)
    UHANN_TILE(4, 2, 2)
    UHANN_TILE(4, 4, 2)
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
        {4, 1, 2, 1},  {4, 1, 1, 1},   {4, 2, 2, 1},  {4, 4, 2, 1},   {2, 16, 4, 4}, {2, 8, 4, 4},
        {2, 8, 4, 2},
    };
    return tiles;
}

bool scan_tile_supported(const ScanTile& tile)
{
    if (tile.rpt == 2) return select_sq(Metric::l2, false, false, tile) != nullptr;
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
    if (tile.rpt == 2) {
        const SqFn sq = select_sq(metric, args.weights != nullptr, weight_fma, tile);
        if (sq == nullptr) throw std::invalid_argument("unsupported scan tile");
        if (args.nq == 0 || args.rows == 0) return;
        const uint32_t tiles_q = (args.nq + tile.tq - 1) / tile.tq;
        const uint32_t per_block = static_cast<uint32_t>(tile.tr * tile.group_q);
        const uint32_t tiles_r = (args.rows + per_block - 1) / per_block;
        const uint64_t blocks = static_cast<uint64_t>(tiles_q) * tiles_r;
        if (blocks > 0x7fffffffULL) throw std::length_error("scan grid exceeds the launch limit");
        hipLaunchKernelGGL(sq, dim3(static_cast<uint32_t>(blocks)), dim3(kChains * tile.group_q), 0, stream, args,
                           tiles_q, tiles_r);
        check_launch("shared-query scan kernel");
        return;
    }
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

void escalate_scan(const ScanArgs& base, const unsigned int* count, const unsigned int* list, uint32_t cap,
                   float* staged, uint64_t ld_staged, uint64_t* out_idx, float* out_dist, hipStream_t stream)
{
    if (cap == 0 || base.rows == 0) return;
    escalate_gather_kernel<<<min(cap, kEscBlocks), 256, 0, stream>>>(base.queries, base.ld_q, base.dim, count, list,
                                                                     cap, staged, ld_staged, base.keys);
    ScanArgs a = base;
    a.queries = staged;
    a.ld_q = ld_staged;
    a.weights = nullptr;
    a.stream_refs = false;
    escalate_scan_kernel<<<kEscBlocks, kChains, 0, stream>>>(a, count, cap);
    escalate_finalize_kernel<<<(cap + 255U) / 256U, 256, 0, stream>>>(base.keys, count, list, cap, out_idx,
                                                                      out_dist);
    check_launch("escalation scan kernels");
}

uint32_t scan_short_max_dim() { return 128; }

void scan_short(Metric metric, bool weight_fma, const float* refs, uint64_t ld_ref, uint32_t rows,
                const float* queries, uint64_t ld_q, uint32_t nq, const float* weights, uint32_t dim,
                unsigned long long* keys, hipStream_t stream)
{
    if (nq == 0 || rows == 0) return;
    if (dim == 0 || dim > scan_short_max_dim()) throw std::invalid_argument("short-row scan supports 1 <= dim <= 128");
    const bool weighted = weights != nullptr;
    const int m = metric == Metric::l2 ? 0 : 1;
    // grid.y carries the query blocks; batches beyond 65535 of them go in
    // slices (a 64-bit counter, so the last step cannot wrap past 2^32)
    const uint64_t slice = 65535ULL * static_cast<uint64_t>(body::kShortThreads);
    for (uint64_t q0 = 0; q0 < nq; q0 += slice) {
        const uint32_t n = static_cast<uint32_t>(nq - q0 < slice ? nq - q0 : slice);
        const float* qs = queries + q0 * ld_q;
        unsigned long long* ks = keys + q0;
        const dim3 grid((rows + body::kShortRefs - 1) / body::kShortRefs,
                        (n + body::kShortThreads - 1) / body::kShortThreads);
#define UHANN_SHORT(M, W, F, D)                                                                                 \
    hipLaunchKernelGGL((scan_short_kernel<M, W, F, D>), grid, dim3(body::kShortThreads), 0, stream, refs, ld_ref, \
                       rows, qs, ld_q, n, weights, dim, ks)
#define UHANN_SHORT_D(D)                                 \
    if (m == 0) UHANN_SHORT(0, false, false, D);         \
    else if (!weighted) UHANN_SHORT(1, false, false, D); \
    else if (weight_fma) UHANN_SHORT(1, true, true, D);  \
    else UHANN_SHORT(1, true, false, D);
        if (dim <= 16) {
            UHANN_SHORT_D(16)
        } else if (dim <= 32) {
            UHANN_SHORT_D(32)
        } else if (dim <= 64) {
            UHANN_SHORT_D(64)
        } else {
            UHANN_SHORT_D(128)
        }
        check_launch("short-row scan kernel");
    }
#undef UHANN_SHORT_D
#undef UHANN_SHORT
}

uint32_t scan_pairs_max_dim() { return 4096; }

void scan_pairs(Metric metric, bool weight_fma, const float* refs_t, uint64_t ld_t, uint32_t rows,
                const float* queries, uint64_t ld_q, uint32_t nq, const float* weights, uint32_t dim,
                unsigned long long* keys, int tq, bool classes, hipStream_t stream)
{
    if (nq == 0 || rows == 0) return;
    if (dim > scan_pairs_max_dim()) throw std::invalid_argument("pair scan supports dim <= 4096");
    const bool weighted = weights != nullptr;
    if (rows >= kPairsMaxRows) throw std::length_error("pair scan supports fewer than 2^25 reference rows");
    // query (and weight) columns in LDS: the whole row padded to a multiple
    // of 4, or two groups' columns at a time for the chain-class body
    const size_t ldqs = (static_cast<size_t>(dim) + 3) & ~static_cast<size_t>(3);
    const size_t lds = classes ? body::cls_lds_floats(tq, weighted, dim) * sizeof(float)
                               : (static_cast<size_t>(tq) * ldqs + (weighted ? ldqs : 0)) * sizeof(float);
    if (classes ? (tq != 4 && tq != 8) : (tq != 1 && tq != 2 && tq != 4)) {
        throw std::invalid_argument(classes ? "the chain-class pair scan supports 4 or 8 queries per block"
                                            : "pair scan supports 1, 2 or 4 queries per block");
    }
    const bool cls = classes;
    // grid.y carries the query tile; batches beyond 65535 tiles go in slices.
    // (a 64-bit counter, so the last step cannot wrap past 2^32)
    const uint64_t slice = 65535ULL * static_cast<uint64_t>(tq);
    for (uint64_t q0 = 0; q0 < nq; q0 += slice) {
        const uint32_t n = static_cast<uint32_t>(nq - q0 < slice ? nq - q0 : slice);
        const float* qs = queries + q0 * ld_q;
        unsigned long long* ks = keys + q0;
        const dim3 grid((rows + 255) / 256, (n + tq - 1) / tq);
#define UHANN_PAIRS(M, W, F, Q)                                                                       \
    do {                                                                                              \
        if constexpr (Q >= 4) {                                                                       \
            if (cls) {                                                                                \
                if (dim <= body::kClsWholeRowDim)                                                     \
                    hipLaunchKernelGGL((scan_pairs_cls_kernel<M, W, F, Q, true>), grid, dim3(256), lds, stream, \
                                       refs_t, ld_t, rows, qs, ld_q, n, weights, dim, ks);            \
                else                                                                                  \
                    hipLaunchKernelGGL((scan_pairs_cls_kernel<M, W, F, Q, false>), grid, dim3(256), lds, stream, \
                                       refs_t, ld_t, rows, qs, ld_q, n, weights, dim, ks);            \
                break;                                                                                \
            }                                                                                         \
        }                                                                                             \
        if constexpr (Q <= 4) {                                                                       \
            hipLaunchKernelGGL((scan_pairs_kernel<M, W, F, Q>), grid, dim3(256), lds, stream, refs_t,     \
                               ld_t, rows, qs, ld_q, n, weights, dim, ks);                            \
        }                                                                                             \
    } while (0)
    const int m = metric == Metric::l2 ? 0 : 1;
    if (tq == 1 && !cls) {
        if (m == 0) UHANN_PAIRS(0, false, false, 1);
        else if (!weighted) UHANN_PAIRS(1, false, false, 1);
        else if (weight_fma) UHANN_PAIRS(1, true, true, 1);
        else UHANN_PAIRS(1, true, false, 1);
    } else if (tq == 2 && !cls) {
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
        if (m == 0) UHANN_PAIRS(0, false, false, 8);
        else if (!weighted) UHANN_PAIRS(1, false, false, 8);
        else if (weight_fma) UHANN_PAIRS(1, true, true, 8);
        else UHANN_PAIRS(1, true, false, 8);
    }
#undef UHANN_PAIRS
    check_launch("pair scan kernel");
    }
}


// Argmin keys of the allocation-free direct query: a file-scope device array
// of a fixed size, one copy per GPU, held until the process ends. Host code
// never passes `scratch` itself to a kernel (in host code the name is only a
// stand-in for the symbol): noalloc_keys resolves the array's address on
// each device once, with that device current, and callers use that address.
constexpr uint32_t kNoAllocQueries = 65536;
constexpr int kNoAllocDevices = 64;
static __device__ __align__(256) unsigned long long scratch[kNoAllocQueries];

uint32_t noalloc_max_queries() { return kNoAllocQueries; }

unsigned long long* noalloc_keys(int device)
{
    static std::mutex mu;
    static unsigned long long* ptrs[kNoAllocDevices] = {};
    if (device < 0 || device >= kNoAllocDevices) throw std::invalid_argument("device id out of range");
    std::lock_guard<std::mutex> lock(mu);
    if (ptrs[device] == nullptr) {
        int current = -1;
        if (hipGetDevice(&current) != hipSuccess || current != device) {
            throw std::runtime_error("the scratch keys must be resolved with their device current");
        }
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || p == nullptr) {
            throw std::runtime_error(std::string("cannot resolve the scratch key array: ") + hipGetErrorString(e));
        }
        ptrs[device] = static_cast<unsigned long long*>(p);
    }
    return ptrs[device];
}

}  // namespace uhann::kern
