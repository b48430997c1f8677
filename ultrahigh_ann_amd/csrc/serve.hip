// Grouped, persistent kernels of a serving tick (layout in serve.hpp).
//
// Each stage kernel is launched with a fixed grid; block b takes the stage's
// units b, b + gridDim.x, ... . A unit belongs to the entry with the largest
// `begin` not above it (binary search, uniform across the block), and its
// coordinates inside the item's unit grid are (u % fast, u / fast). The tile
// bodies are the one-index kernels' bodies, so every request gets the bits
// it would get alone: the binary32 scans and the K7 projection are the
// reference's order, the bf16 projection is the one-index MFMA tile.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <stdexcept>
#include <string>

#include "bodies.hpp"
#include "serve.hpp"

namespace uhann::kern {
namespace {

constexpr int kThreads = 256;

void check_launch(const char* what)
{
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        throw std::runtime_error(std::string("cannot launch ") + what + ": " + hipGetErrorString(e));
    }
}

struct Unit {
    const ServeItem* item;
    uint32_t a;  // u % fast
    uint32_t b;  // u / fast
};

__device__ __forceinline__ const ServeHeader& header(const void* blob)
{
    return *static_cast<const ServeHeader*>(blob);
}

__device__ __forceinline__ Unit locate(const void* blob, int stage, uint32_t u)
{
    const ServeHeader& h = header(blob);
    const auto* base = static_cast<const unsigned char*>(blob);
    const auto* e = reinterpret_cast<const ServeEntry*>(base + h.offset[stage]);
    const auto* items = reinterpret_cast<const ServeItem*>(base + h.items_offset);
    uint32_t lo = 0;
    uint32_t hi = h.count[stage] - 1U;
    while (lo < hi) {
        const uint32_t mid = (lo + hi + 1U) / 2U;
        if (e[mid].begin <= u) {
            lo = mid;
        } else {
            hi = mid - 1U;
        }
    }
    const uint32_t local = u - e[lo].begin;
    return {items + e[lo].item, local % e[lo].fast, local / e[lo].fast};
}

__global__ void __launch_bounds__(kThreads)
serve_reset_kernel(const void* blob, unsigned long long* keys, int* flags)
{
    const ServeHeader& h = header(blob);
    for (uint32_t i = blockIdx.x * kThreads + threadIdx.x; i < h.nq; i += gridDim.x * kThreads) keys[i] = ~0ULL;
    for (uint32_t i = blockIdx.x * kThreads + threadIdx.x; i < 4U * h.items; i += gridDim.x * kThreads) flags[i] = 0;
}

__global__ void __launch_bounds__(kThreads)
serve_finalize_kernel(const void* blob, const unsigned long long* keys, uint64_t* out_idx, float* out_dist)
{
    const ServeHeader& h = header(blob);
    for (uint32_t i = blockIdx.x * kThreads + threadIdx.x; i < h.nq; i += gridDim.x * kThreads) {
        const unsigned long long key = keys[i];
        out_idx[i] = key & 0xffffffffULL;
        out_dist[i] = __uint_as_float(static_cast<uint32_t>(key >> 32));
    }
}

// Unit = one query row; every value of the row must be finite.
__global__ void __launch_bounds__(kThreads) serve_check_kernel(const void* blob)
{
    const uint32_t total = header(blob).units[kServeCheck];
    for (uint32_t u = blockIdx.x; u < total; u += gridDim.x) {
        const Unit w = locate(blob, kServeCheck, u);
        const ServeItem& it = *w.item;
        const float* row = it.queries + static_cast<uint64_t>(w.b) * it.ld_q;
        bool bad = false;
        for (uint32_t c = threadIdx.x; c < it.dim; c += kThreads) bad |= !__builtin_isfinite(row[c]);
        if (bad) atomicExch(it.flags, 1);
    }
}

__global__ void __launch_bounds__(kThreads) serve_gather_kernel(const void* blob)
{
    const uint32_t total = header(blob).units[kServeGather];
    for (uint32_t u = blockIdx.x; u < total; u += gridDim.x) {
        const Unit w = locate(blob, kServeGather, u);
        const ServeItem& it = *w.item;
        __syncthreads();
        body::gather_tile(it.queries, it.ld_q, it.n, it.cols, it.scales, it.k, it.gathered, it.ld_g, it.flags,
                          w.a, w.b);
    }
}

__global__ void __launch_bounds__(kThreads) serve_project32_kernel(const void* blob)
{
    const uint32_t total = header(blob).units[kServeProject32];
    for (uint32_t u = blockIdx.x; u < total; u += gridDim.x) {
        const Unit w = locate(blob, kServeProject32, u);
        const ServeItem& it = *w.item;
        __syncthreads();
        body::project_fp32_tile(it.gathered, it.ld_g, it.n, it.k, it.proj, it.ld_proj, it.p, it.projected, it.ld_p,
                                it.flags + 1, w.b * body::kProjQ, w.a * body::kProjP);
    }
}

__global__ void __launch_bounds__(kThreads) serve_project16_kernel(const void* blob)
{
    const uint32_t total = header(blob).units[kServeProject16];
    for (uint32_t u = blockIdx.x; u < total; u += gridDim.x) {
        const Unit w = locate(blob, kServeProject16, u);
        const ServeItem& it = *w.item;
        __syncthreads();
        body::project_bf16_tile(it.gathered, it.ld_g, it.n, it.k, static_cast<const __bf16*>(it.proj16),
                                it.ld_proj16, it.p, it.projected, it.ld_p, it.flags + 1, w.b * 32U, w.a * 32U);
    }
}

template <int METRIC, bool WEIGHTED, int TQ>
__global__ void __launch_bounds__(kThreads)
serve_pairs_kernel(const void* blob, int stage, unsigned long long* keys)
{
    extern __shared__ float smem[];
    const uint32_t total = header(blob).units[stage];
    for (uint32_t u = blockIdx.x; u < total; u += gridDim.x) {
        const Unit w = locate(blob, stage, u);
        const ServeItem& it = *w.item;
        __syncthreads();
        // a: query tile (the fast index), b: 256-row reference block
        body::scan_pairs_serve<METRIC, WEIGHTED, true, TQ>(it.refs_t, it.rows, it.rows, it.scan_q, it.ld_sq, it.n,
                                                            it.weights, it.dim, keys + it.qoff, w.a * TQ, w.b, smem);
    }
}

// Pair stages at 4 queries per unit: the chain-class body (reference copies
// carry kPairsPadCols zero rows past dim); WHOLE for the short pair stages
// (rows up to kServePairsShortDim <= body::kClsWholeRowDim).
template <int METRIC, bool WEIGHTED, bool WHOLE>
__global__ void __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(4)))
serve_pairs_cls_kernel(const void* blob, int stage, unsigned long long* keys)
{
    extern __shared__ float smem[];
    const uint32_t total = header(blob).units[stage];
    for (uint32_t u = blockIdx.x; u < total; u += gridDim.x) {
        const Unit w = locate(blob, stage, u);
        const ServeItem& it = *w.item;
        __syncthreads();
        // a: query tile (the fast index), b: 256-row reference block
        body::scan_pairs_cls<METRIC, WEIGHTED, true, 4, WHOLE>(it.refs_t, it.rows, it.rows, it.scan_q, it.ld_sq, it.n,
                                                               it.weights, it.dim, keys + it.qoff, w.a * 4U, w.b,
                                                               smem);
    }
}

template <int METRIC, bool WEIGHTED>
__global__ void __launch_bounds__(kThreads)
serve_tile_kernel(const void* blob, int stage, unsigned long long* keys)
{
    const uint32_t total = header(blob).units[stage];
    for (uint32_t u = blockIdx.x; u < total; u += gridDim.x) {
        const Unit w = locate(blob, stage, u);
        const ServeItem& it = *w.item;
        ScanArgs a;
        a.refs = it.refs;
        a.ld_ref = it.ld_ref;
        a.rows = it.rows;
        a.queries = it.scan_q;
        a.ld_q = it.ld_sq;
        a.nq = it.n;
        a.weights = it.weights;
        a.dim = it.dim;
        a.keys = keys + it.qoff;
        __syncthreads();
        body::scan_tile<METRIC, WEIGHTED, true, 1, kServeTileQ, kServeTileR>(a, w.b * kServeTileQ,
                                                                            w.a * kServeTileR);
    }
}

template <int PMAX>
__global__ void __launch_bounds__(kThreads)
serve_median_kernel(const void* blob, int stage, unsigned long long* keys)
{
    const uint32_t total = header(blob).units[stage];
    for (uint32_t u = blockIdx.x; u < total; u += gridDim.x) {
        const Unit w = locate(blob, stage, u);
        const ServeItem& it = *w.item;
        __syncthreads();
        body::median_block<PMAX, false>(it.refs_t, it.rows, it.projected, it.ld_p, it.n, it.p, keys + it.qoff,
                                        it.flags + 2, w.b, w.a);
    }
}

template <int TQ>
void launch_pairs(const ServeLaunch& l, int stage, hipStream_t stream)
{
    const bool weighted = stage == kServePairsL1W || stage == kServePairsShortL1W;
    const bool plain_l1 = stage == kServePairsL1 || stage == kServePairsShortL1;
    const bool short_rows = stage == kServePairsShortL2 || stage == kServePairsShortL1 || stage == kServePairsShortL1W;
    const uint32_t max_dim = short_rows ? kServePairsShortDim : kServePairsMaxDim;
    if constexpr (TQ == 4) {
        static_assert(kServePairsShortDim <= body::kClsWholeRowDim && kServePairsMaxDim > body::kClsWholeRowDim,
                      "short pair stages stage whole rows, long ones two groups at a time");
        const size_t lds = body::cls_lds_floats(TQ, weighted, max_dim) * sizeof(float);
        if (short_rows) {
            if (weighted) {
                serve_pairs_cls_kernel<1, true, true><<<l.grid, kThreads, lds, stream>>>(l.blob, stage, l.keys);
            } else if (plain_l1) {
                serve_pairs_cls_kernel<1, false, true><<<l.grid, kThreads, lds, stream>>>(l.blob, stage, l.keys);
            } else {
                serve_pairs_cls_kernel<0, false, true><<<l.grid, kThreads, lds, stream>>>(l.blob, stage, l.keys);
            }
        } else if (weighted) {
            serve_pairs_cls_kernel<1, true, false><<<l.grid, kThreads, lds, stream>>>(l.blob, stage, l.keys);
        } else if (plain_l1) {
            serve_pairs_cls_kernel<1, false, false><<<l.grid, kThreads, lds, stream>>>(l.blob, stage, l.keys);
        } else {
            serve_pairs_cls_kernel<0, false, false><<<l.grid, kThreads, lds, stream>>>(l.blob, stage, l.keys);
        }
    } else {
        const size_t lds = static_cast<size_t>(TQ + (weighted ? 1 : 0)) * max_dim * sizeof(float);
        if (weighted) {
            serve_pairs_kernel<1, true, TQ><<<l.grid, kThreads, lds, stream>>>(l.blob, stage, l.keys);
        } else if (plain_l1) {
            serve_pairs_kernel<1, false, TQ><<<l.grid, kThreads, lds, stream>>>(l.blob, stage, l.keys);
        } else {
            serve_pairs_kernel<0, false, TQ><<<l.grid, kThreads, lds, stream>>>(l.blob, stage, l.keys);
        }
    }
}

}  // namespace

bool serve_pairs_tq_supported(int tq) { return tq == 1 || tq == 2 || tq == 4; }

void serve_tick(const ServeLaunch& l, uint32_t mask, hipStream_t stream)
{
    if (!serve_pairs_tq_supported(l.pairs_tq)) throw std::invalid_argument("unsupported pair-scan query tile");
    const uint32_t small = (l.max_q + kThreads - 1) / kThreads > 0 ? (l.max_q + kThreads - 1) / kThreads : 1;
    serve_reset_kernel<<<small < 1024U ? small : 1024U, kThreads, 0, stream>>>(l.blob, l.keys, l.flags);
    check_launch("serving reset kernel");
    auto on = [&](int s) { return (mask >> s) & 1U; };
    if (on(kServeCheck)) serve_check_kernel<<<l.grid, kThreads, 0, stream>>>(l.blob);
    if (on(kServeGather)) serve_gather_kernel<<<l.grid, kThreads, 0, stream>>>(l.blob);
    if (on(kServeProject32)) serve_project32_kernel<<<l.grid, kThreads, 0, stream>>>(l.blob);
    if (on(kServeProject16)) serve_project16_kernel<<<l.grid, kThreads, 0, stream>>>(l.blob);
    for (int s : {kServePairsShortL2, kServePairsShortL1, kServePairsShortL1W, kServePairsL2, kServePairsL1,
                  kServePairsL1W}) {
        if (!on(s)) continue;
        if (l.pairs_tq == 1) {
            launch_pairs<1>(l, s, stream);
        } else if (l.pairs_tq == 2) {
            launch_pairs<2>(l, s, stream);
        } else {
            launch_pairs<4>(l, s, stream);
        }
    }
    if (on(kServeTileL2)) serve_tile_kernel<0, false><<<l.grid, kThreads, 0, stream>>>(l.blob, kServeTileL2, l.keys);
    if (on(kServeTileL1)) serve_tile_kernel<1, false><<<l.grid, kThreads, 0, stream>>>(l.blob, kServeTileL1, l.keys);
    if (on(kServeTileL1W)) serve_tile_kernel<1, true><<<l.grid, kThreads, 0, stream>>>(l.blob, kServeTileL1W, l.keys);
    if (on(kServeMedian8)) serve_median_kernel<8><<<l.grid, kThreads, 0, stream>>>(l.blob, kServeMedian8, l.keys);
    if (on(kServeMedian16)) serve_median_kernel<16><<<l.grid, kThreads, 0, stream>>>(l.blob, kServeMedian16, l.keys);
    if (on(kServeMedian32)) serve_median_kernel<32><<<l.grid, kThreads, 0, stream>>>(l.blob, kServeMedian32, l.keys);
    if (on(kServeMedian64)) serve_median_kernel<64><<<l.grid, kThreads, 0, stream>>>(l.blob, kServeMedian64, l.keys);
    if (on(kServeMedian128)) {
        serve_median_kernel<128><<<l.grid, kThreads, 0, stream>>>(l.blob, kServeMedian128, l.keys);
    }
    check_launch("serving stage kernel");
    serve_finalize_kernel<<<small < 1024U ? small : 1024U, kThreads, 0, stream>>>(l.blob, l.keys, l.out_idx,
                                                                                  l.out_dist);
    check_launch("serving finalisation kernel");
}

}  // namespace uhann::kern
