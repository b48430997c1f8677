// Device bodies of the scan, gather, projection and median kernels. Each
// body computes one tile given its coordinates; the one-index kernels in
// scan.hip, gather.hip and hier.hip pass their block coordinates, and the
// grouped serving kernels in serve.hip pass the coordinates of a work unit
// of one request, so both run the same arithmetic in the same order.
// Bodies use the whole block (barriers inside); callers that loop over
// several tiles put a barrier between two calls.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "kernels.hpp"

namespace uhann::kern::body {

constexpr int kChains = 256;
// Word 3 of a raw (stride 0) buffer resource on gfx9-family parts: 32-bit
// data format, no swizzle.
constexpr int kBufferWord3 = 0x00020000;
constexpr int kGroups = 8;  // groups of 32 chains


typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx16 __attribute__((ext_vector_type(16)));

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

typedef float floatx4_t __attribute__((ext_vector_type(4)));
typedef float floatx2_t __attribute__((ext_vector_type(2)));

// load_vec with the non-temporal cache policy (streamed data).
template <int RPT>
__device__ __forceinline__ void load_vec_stream(const float* __restrict__ p, float (&v)[RPT])
{
    if constexpr (RPT == 1) {
        v[0] = __builtin_nontemporal_load(p);
    } else if constexpr (RPT == 2) {
        const floatx2_t x = __builtin_nontemporal_load(reinterpret_cast<const floatx2_t*>(p));
        v[0] = x[0];
        v[1] = x[1];
    } else {
        const floatx4_t x = __builtin_nontemporal_load(reinterpret_cast<const floatx4_t*>(p));
        v[0] = x[0];
        v[1] = x[1];
        v[2] = x[2];
        v[3] = x[3];
    }
}

// load_vec through a raw buffer resource: vector byte offset voff (the
// thread's column), scalar byte offset soff (the row); AUX is the cache
// policy (2: non-temporal).
template <int RPT, int AUX>
__device__ __forceinline__ void load_buf(__amdgpu_buffer_rsrc_t rs, uint32_t voff, uint32_t soff, float (&v)[RPT])
{
    if constexpr (RPT == 1) {
        v[0] = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(rs, voff, soff, AUX));
    } else if constexpr (RPT == 2) {
        const auto x = __builtin_amdgcn_raw_buffer_load_b64(rs, voff, soff, AUX);
        v[0] = __uint_as_float(x[0]);
        v[1] = __uint_as_float(x[1]);
    } else {
        const auto x = __builtin_amdgcn_raw_buffer_load_b128(rs, voff, soff, AUX);
        v[0] = __uint_as_float(x[0]);
        v[1] = __uint_as_float(x[1]);
        v[2] = __uint_as_float(x[2]);
        v[3] = __uint_as_float(x[3]);
    }
}

template <bool V>
struct BoolTag {
    static constexpr bool value = V;
};

// One chain step from the difference d = r - q (binary32).
template <int METRIC, bool WEIGHTED, bool WFMA>
__device__ __forceinline__ float accumulate_diff(float acc, float d, float w)
{
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

template <int METRIC, bool WEIGHTED, bool WFMA>
__device__ __forceinline__ float accumulate(float acc, float r, float q, float w)
{
    return accumulate_diff<METRIC, WEIGHTED, WFMA>(acc, r - q, w);
}

// An empty asm statement that reads and writes v[0 .. N): every value is
// formed before it and used after it, so the instructions that produce a
// group of values issue together and their consumers follow together (a
// dependent pair issued back to back waits for the first result).
template <int N>
__device__ __forceinline__ void issue_fence(float* v)
{
    if constexpr (N >= 8) {
        asm volatile("" : "+v"(v[0]), "+v"(v[1]), "+v"(v[2]), "+v"(v[3]), "+v"(v[4]), "+v"(v[5]), "+v"(v[6]), "+v"(v[7]));
        issue_fence<N - 8>(v + 8);
    } else if constexpr (N >= 4) {
        asm volatile("" : "+v"(v[0]), "+v"(v[1]), "+v"(v[2]), "+v"(v[3]));
        issue_fence<N - 4>(v + 4);
    } else if constexpr (N >= 2) {
        asm volatile("" : "+v"(v[0]), "+v"(v[1]));
        issue_fence<N - 2>(v + 2);
    } else if constexpr (N == 1) {
        asm volatile("" : "+v"(v[0]));
    }
}

// One TQ x TR tile of the direct scan (queries q0.., references r0..).
template <int METRIC, bool WEIGHTED, bool WFMA, int RPT, int TQ, int TR>
__device__ __forceinline__ void scan_tile(const ScanArgs& a, uint32_t q0, uint32_t r0)
{
    constexpr int NT = kChains / RPT;     // threads per block
    constexpr int GROUP = 32 / RPT;       // threads per group of 32 chains
    constexpr int NP = TQ * TR;           // pairs per block
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

    // Buffer form of the loads (a.buffers): one resource at the tile's first
    // reference row and one at its first query row; reference j and query i
    // are scalar byte offsets from them (clamped to the last row like the
    // pointers above) and the column is the thread's one vector offset, so
    // no 64-bit address is formed per load.
    const __amdgpu_buffer_rsrc_t rres = __builtin_amdgcn_make_buffer_rsrc(
        const_cast<float*>(a.refs + static_cast<uint64_t>(r0) * a.ld_ref), 0, -1, kBufferWord3);
    const __amdgpu_buffer_rsrc_t qres = __builtin_amdgcn_make_buffer_rsrc(
        const_cast<float*>(a.queries + static_cast<uint64_t>(min(q0, a.nq - 1U)) * a.ld_q), 0, -1, kBufferWord3);
    uint32_t roff[TR];
    uint32_t qoff[TQ];
#pragma unroll
    for (int j = 0; j < TR; ++j) {
        roff[j] = min(static_cast<uint32_t>(j), a.rows - 1U - r0) * static_cast<uint32_t>(a.ld_ref) * 4U;
    }
#pragma unroll
    for (int i = 0; i < TQ; ++i) {
        qoff[i] = q0 < a.nq ? min(static_cast<uint32_t>(i), a.nq - 1U - q0) * static_cast<uint32_t>(a.ld_q) * 4U : 0U;
    }

    // One 256-column step at column `col` (this thread's first chain). In the
    // tail step a vector may straddle dim (rows are padded to a multiple of 4
    // floats): components at or past dim are skipped, so pad values are
    // never added.
    auto step = [&](uint32_t col, bool masked, bool stream, auto buf_tag) {
        constexpr bool BUF = decltype(buf_tag)::value;
        float rv[TR][RPT];
        float qv[TQ][RPT];
        float wv[RPT];
        if constexpr (BUF) {
            const uint32_t voff = col * 4U;
            if (stream) {
#pragma unroll
                for (int j = 0; j < TR; ++j) load_buf<RPT, 2>(rres, voff, roff[j], rv[j]);
            } else {
#pragma unroll
                for (int j = 0; j < TR; ++j) load_buf<RPT, 0>(rres, voff, roff[j], rv[j]);
            }
#pragma unroll
            for (int i = 0; i < TQ; ++i) load_buf<RPT, 0>(qres, voff, qoff[i], qv[i]);
        } else {
            if (stream) {
#pragma unroll
                for (int j = 0; j < TR; ++j) load_vec_stream<RPT>(rrow[j] + col, rv[j]);
            } else {
#pragma unroll
                for (int j = 0; j < TR; ++j) load_vec<RPT>(rrow[j] + col, rv[j]);
            }
#pragma unroll
            for (int i = 0; i < TQ; ++i) load_vec<RPT>(qrow[i] + col, qv[i]);
        }
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
            for (int i = 0; i < TQ; ++i) {
                // The TR differences first, then their TR chain steps.
                float dd[TR];
#pragma unroll
                for (int j = 0; j < TR; ++j) dd[j] = rv[j][k] - qv[i][k];
                issue_fence<TR>(dd);
#pragma unroll
                for (int j = 0; j < TR; ++j)
                    acc[i][j][k] = accumulate_diff<METRIC, WEIGHTED, WFMA>(acc[i][j][k], dd[j], wv[k]);
            }
        }
    };

    const uint32_t col_t = static_cast<uint32_t>(RPT * t);
    const uint32_t full_end = (a.dim / kChains) * kChains;
    // Replay the reference tree for the chains of one pair (v: this thread's
    // RPT chain sums): cross-thread steps first (chain offsets 16..RPT are
    // thread offsets GROUP/2..1), then the chains inside the thread; lane 0 of
    // each group of 32 chains stores the group sum for pair p.
    __shared__ float group_sums[kGroups][NP];
    __shared__ float pair_dist[NP];
    auto pair_tree = [&](float (&v)[RPT], int p) {
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
        if (t % GROUP == 0) group_sums[t / GROUP][p] = w;
    };

    bool rows_done = false;
    if constexpr (RPT == 4 && TQ == 1 && TR > 1) {
        if (a.stream_refs) {
            // One query streaming its references: the rows are read one after
            // the other (one sequential stream per wave, which HBM serves
            // better than TR interleaved ones), eight 256-column steps of
            // loads in flight, and each row's tree runs before the next row.
            // Each pair's chains see the same columns in the same order as in
            // the interleaved form.
#pragma unroll 1
            for (int j = 0; j < TR; ++j) {
                const float* rr = a.refs + static_cast<uint64_t>(min(r0 + j, a.rows - 1U)) * a.ld_ref;
                float v[RPT];
#pragma unroll
                for (int k = 0; k < RPT; ++k) v[k] = 0.0F;
                auto step1 = [&](uint32_t col, bool masked) {
                    float rv[RPT];
                    float qv[RPT];
                    float wv[RPT];
                    if (masked) {
                        load_vec<RPT>(rr + col, rv);
                    } else {
                        load_vec_stream<RPT>(rr + col, rv);
                    }
                    load_vec<RPT>(qrow[0] + col, qv);
                    if constexpr (WEIGHTED) {
                        load_vec<RPT>(a.weights + col, wv);
                    } else {
#pragma unroll
                        for (int k = 0; k < RPT; ++k) wv[k] = 0.0F;
                    }
#pragma unroll
                    for (int k = 0; k < RPT; ++k) {
                        if (masked && col + static_cast<uint32_t>(k) >= a.dim) continue;
                        v[k] = accumulate<METRIC, WEIGHTED, WFMA>(v[k], rv[k], qv[k], wv[k]);
                    }
                };
#pragma unroll 4
                for (uint32_t c0 = 0; c0 < full_end; c0 += kChains) step1(c0 + col_t, false);
                if (full_end + col_t < a.dim) step1(full_end + col_t, true);
                pair_tree(v, j);
            }
            rows_done = true;
        }
    }
    if (!rows_done) {
        // Single-pass streams keep more 256-column steps of loads in flight.
        constexpr int kStreamUnroll = RPT == 4 && TQ * TR <= 8 ? 4 : 2;
        if (a.buffers) {
            if (a.stream_refs) {
#pragma unroll kStreamUnroll
                for (uint32_t c0 = 0; c0 < full_end; c0 += kChains) step(c0 + col_t, false, true, BoolTag<true>{});
            } else {
#pragma unroll 2
                for (uint32_t c0 = 0; c0 < full_end; c0 += kChains) step(c0 + col_t, false, false, BoolTag<true>{});
            }
        } else if (a.stream_refs) {
#pragma unroll kStreamUnroll
            for (uint32_t c0 = 0; c0 < full_end; c0 += kChains) step(c0 + col_t, false, true, BoolTag<false>{});
        } else {
#pragma unroll 2
            for (uint32_t c0 = 0; c0 < full_end; c0 += kChains) step(c0 + col_t, false, false, BoolTag<false>{});
        }
        if (full_end + col_t < a.dim) step(full_end + col_t, true, false, BoolTag<false>{});
#pragma unroll
        for (int i = 0; i < TQ; ++i) {
#pragma unroll
            for (int j = 0; j < TR; ++j) {
                float v[RPT];
#pragma unroll
                for (int k = 0; k < RPT; ++k) v[k] = acc[i][j][k];
                pair_tree(v, i * TR + j);
            }
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


// Shared-query tile for small batches: G reference groups of TR rows each
// and TQ queries per block, one chain per thread (thread t: chain t % 256,
// group t / 256), so a thread holds TQ x TR partial sums. The block stages
// each 256-column step's query values in LDS once (two buffers of one or two
// steps each, filled one buffer ahead of use) and all G groups read them
// there, so the query traffic from L2 is one TQ x 256 stripe per step for
// G TR reference rows.
// Each chain adds its terms in column order and the tree is the reference's
// (32-lane shuffles, then ((W0+W4)+(W2+W6))+((W1+W5)+(W3+W7))), so the bits
// are those of scan_tile.
constexpr int kSqDepth = 4;  // steps of reference values in flight (the step loop's unroll)
template <int METRIC, bool WEIGHTED, bool WFMA, int TQ, int TR, int G>
__device__ __forceinline__ void scan_tile_sq(const ScanArgs& a, uint32_t q0, uint32_t r0, bool stream)
{
    constexpr int NT = kChains * G;
    constexpr int NP = G * TQ * TR;
    static_assert(TQ % G == 0, "the query stripe must split evenly over the reference groups");
    constexpr int SQ = TQ / G;  // query values each thread stages per step
    // Without weights two steps share an LDS buffer (one barrier per two
    // steps); the weighted form keeps one step per buffer, as its extra
    // registers for two steps would spill.
    constexpr bool kPairs = !WEIGHTED;
    constexpr int kSlots = kPairs ? 4 : 2;
    __shared__ float qs[kSlots][TQ][kChains];
    __shared__ float wsd[kSlots][kChains];
    __shared__ float group_sums[kGroups][NP];
    __shared__ float pair_dist[NP];
    const int t = threadIdx.x;
    const uint32_t c = static_cast<uint32_t>(t % kChains);
    // the group is the same for the whole wave: row addresses in scalar registers
    const int g = __builtin_amdgcn_readfirstlane(t / kChains);
    const float* rrow[TR];
#pragma unroll
    for (int j = 0; j < TR; ++j) {
        rrow[j] = a.refs + static_cast<uint64_t>(min(r0 + static_cast<uint32_t>(g * TR + j), a.rows - 1U)) * a.ld_ref;
    }
    // this thread stages queries g, g + G, ... at its chain's column
    const float* qrow[SQ];
#pragma unroll
    for (int k = 0; k < SQ; ++k) {
        qrow[k] = a.queries + static_cast<uint64_t>(min(q0 + static_cast<uint32_t>(g + k * G), a.nq - 1U)) * a.ld_q;
    }
    float acc[TQ][TR];
#pragma unroll
    for (int i = 0; i < TQ; ++i)
#pragma unroll
        for (int j = 0; j < TR; ++j) acc[i][j] = 0.0F;

    const uint32_t steps = (a.dim + kChains - 1) / kChains;
    // Columns past the row are read at the row's last column: the step that
    // holds them adds nothing for them, so their values are never used, and
    // the loads need no branch.
    const uint32_t last = a.dim - 1U;
    auto load_refs = [&](uint32_t col, float (&rv)[TR]) __attribute__((always_inline)) {
        const uint32_t cc = min(col, last);
#pragma unroll
        for (int j = 0; j < TR; ++j) rv[j] = stream ? __builtin_nontemporal_load(rrow[j] + cc) : rrow[j][cc];
    };
    auto load_queries = [&](uint32_t col, float (&qv)[SQ], float& wv) __attribute__((always_inline)) {
        const uint32_t cc = min(col, last);
#pragma unroll
        for (int k = 0; k < SQ; ++k) qv[k] = qrow[k][cc];
        wv = WEIGHTED ? a.weights[cc] : 0.0F;
    };
    auto store_queries = [&](int buf, const float (&qv)[SQ], float wv) __attribute__((always_inline)) {
#pragma unroll
        for (int k = 0; k < SQ; ++k) qs[buf][g + k * G][c] = qv[k];
        if (WEIGHTED && g == 0) wsd[buf][c] = wv;
    };
    // kSqDepth sets of reference registers and two of query registers
    // rotate with the steps, so nothing is copied between steps: at the end
    // of step s its reference set takes the values of step s + kSqDepth; the
    // query values reach LDS one buffer (one or two steps) ahead of use and
    // are loaded one buffer before that. Rows of no coordinates add no terms
    // (every distance is zero).
    if (steps != 0) {
        if constexpr (kPairs) {
            // pair P = steps 2P, 2P + 1: its query values take LDS slots 2 (P & 1)
            // and 2 (P & 1) + 1, and one barrier serves both steps
            float rv[kSqDepth][TR], qa[2][SQ], qb[2][SQ], wa[2], wb[2];
            auto load_pair = [&](uint32_t col, float (&qv)[2][SQ], float (&wv)[2]) __attribute__((always_inline)) {
                load_queries(col, qv[0], wv[0]);
                load_queries(col + kChains, qv[1], wv[1]);
            };
            {
                float q0v[2][SQ], w0v[2];
                load_pair(c, q0v, w0v);
                store_queries(0, q0v[0], w0v[0]);
                store_queries(1, q0v[1], w0v[1]);
            }
            load_pair(c + 2 * kChains, qa, wa);
            load_pair(c + 4 * kChains, qb, wb);
#pragma unroll
            for (int u = 0; u < kSqDepth; ++u) load_refs(c + static_cast<uint32_t>(u) * kChains, rv[u]);
            __syncthreads();
            auto step = [&](uint32_t s, int slot, float (&rs)[TR]) __attribute__((always_inline)) {
                const uint32_t col = s * kChains + c;
                if (col < a.dim) {
                    const float wl = WEIGHTED ? wsd[slot][c] : 0.0F;
                    float ql[TQ];
#pragma unroll
                    for (int i = 0; i < TQ; ++i) ql[i] = qs[slot][i][c];
#pragma unroll
                    for (int i = 0; i < TQ; ++i) {
                        float dd[TR];
#pragma unroll
                        for (int j = 0; j < TR; ++j) dd[j] = rs[j] - ql[i];
                        issue_fence<TR>(dd);
#pragma unroll
                        for (int j = 0; j < TR; ++j) acc[i][j] = accumulate_diff<METRIC, WEIGHTED, WFMA>(acc[i][j], dd[j], wl);
                    }
                }
                load_refs(col + kSqDepth * kChains, rs);
            };
            auto pair_end = [&](uint32_t s, int buf, float (&qv)[2][SQ], float (&wv)[2]) __attribute__((always_inline)) {
                store_queries(2 * (buf ^ 1), qv[0], wv[0]);
                store_queries(2 * (buf ^ 1) + 1, qv[1], wv[1]);
                __syncthreads();
                load_pair(s * kChains + c + 6 * kChains, qv, wv);
            };
#pragma unroll 1
            for (uint32_t s = 0; s < steps; s += 4) {
                step(s, 0, rv[0]);
                step(s + 1, 1, rv[1]);
                pair_end(s, 0, qa, wa);
                step(s + 2, 2, rv[2]);
                step(s + 3, 3, rv[3]);
                pair_end(s + 2, 1, qb, wb);
            }
        } else {
            float rv[kSqDepth][TR], qa[SQ], qb[SQ], wa, wb;
            {
                float q0v[SQ], w0v;
                load_queries(c, q0v, w0v);
                store_queries(0, q0v, w0v);
            }
            load_queries(c + kChains, qa, wa);
            load_queries(c + 2 * kChains, qb, wb);
#pragma unroll
            for (int u = 0; u < kSqDepth; ++u) load_refs(c + static_cast<uint32_t>(u) * kChains, rv[u]);
            __syncthreads();
            auto step = [&](uint32_t s, float (&rs)[TR], float (&qv)[SQ], float& wv) __attribute__((always_inline)) {
                const uint32_t col = s * kChains + c;
                const int buf = static_cast<int>(s & 1U);
                if (col < a.dim) {
                    const float wl = WEIGHTED ? wsd[buf][c] : 0.0F;
                    float ql[TQ];
#pragma unroll
                    for (int i = 0; i < TQ; ++i) ql[i] = qs[buf][i][c];
#pragma unroll
                    for (int i = 0; i < TQ; ++i) {
                        float dd[TR];
#pragma unroll
                        for (int j = 0; j < TR; ++j) dd[j] = rs[j] - ql[i];
                        issue_fence<TR>(dd);
#pragma unroll
                        for (int j = 0; j < TR; ++j) acc[i][j] = accumulate_diff<METRIC, WEIGHTED, WFMA>(acc[i][j], dd[j], wl);
                    }
                }
                store_queries(buf ^ 1, qv, wv);
                __syncthreads();
                load_refs(col + kSqDepth * kChains, rs);
                load_queries(col + 3 * kChains, qv, wv);
            };
            // steps past the row (fewer than kSqDepth) add nothing
#pragma unroll 1
            for (uint32_t s = 0; s < steps; s += kSqDepth) {
#pragma unroll
                for (int u = 0; u < kSqDepth; u += 2) {
                    step(s + u, rv[u], qa, wa);
                    step(s + u + 1, rv[u + 1], qb, wb);
                }
            }
        }
    }
    // The reference's tree for each pair: chain offsets 16..1 are lane
    // offsets inside a 32-lane group; lane 0 of each group stores W.
#pragma unroll
    for (int i = 0; i < TQ; ++i) {
#pragma unroll
        for (int j = 0; j < TR; ++j) {
            float v = acc[i][j];
#pragma unroll
            for (int off = 16; off >= 1; off /= 2) v = v + __shfl_down(v, off);
            if (c % 32U == 0) group_sums[c / 32U][(g * TQ + i) * TR + j] = v;
        }
    }
    __syncthreads();
    for (int p = t; p < NP; p += NT) {
        const float g0 = group_sums[0][p], g1 = group_sums[1][p];
        const float g2 = group_sums[2][p], g3 = group_sums[3][p];
        const float g4 = group_sums[4][p], g5 = group_sums[5][p];
        const float g6 = group_sums[6][p], g7 = group_sums[7][p];
        pair_dist[p] = ((g0 + g4) + (g2 + g6)) + ((g1 + g5) + (g3 + g7));
    }
    __syncthreads();
    if (t < TQ) {
        const uint32_t q = q0 + static_cast<uint32_t>(t);
        if (q < a.nq) {
            unsigned long long best = ~0ULL;
#pragma unroll
            for (int gg = 0; gg < G; ++gg) {
#pragma unroll
                for (int j = 0; j < TR; ++j) {
                    const uint32_t r = r0 + static_cast<uint32_t>(gg * TR + j);
                    if (r < a.rows) {
                        const unsigned long long key =
                            (static_cast<unsigned long long>(__float_as_uint(pair_dist[(gg * TQ + t) * TR + j])) << 32) |
                            r;
                        best = key < best ? key : best;
                    }
                }
            }
            atomicMin(a.keys + q, best);
        }
    }
}

// Thread-per-query form for rows of at most 128 coordinates (hierarchical p,
// small flat |I|): thread t owns query q0 + t with its row in registers, the
// block's kShortRefs references (and the L1 weights) sit in LDS, zero-padded
// to DMAX columns, and the thread walks them keeping its own (distance, row)
// minimum, so a pair costs no cross-lane reduction. With at most 256 columns
// every chain holds at most one term: chain c is the reference's first
// accumulation step from +0 (fmaf(d, d, 0) for L2, 0 + |d| or the weighted
// step for L1). The 32-chain trees (offsets 16, 8, 4, 2, 1) of the groups
// that hold columns follow, then ((W0+W4)+(W2+W6))+((W1+W5)+(W3+W7)); chains
// and groups without columns contribute +0, and adding +0 to a non-negative
// sum changes no bit, so they are left out. Same distances as the
// reference's K2/K5.
constexpr int kShortRefs = 128;
constexpr int kShortThreads = 256;

template <int METRIC, bool WEIGHTED, bool WFMA, int DMAX>
__device__ __forceinline__ void scan_short(const float* __restrict__ refs, uint64_t ld_ref, uint32_t rows,
                                           const float* __restrict__ queries, uint64_t ld_q, uint32_t nq,
                                           const float* __restrict__ weights, uint32_t dim,
                                           unsigned long long* __restrict__ keys, uint32_t q0, uint32_t r0)
{
    static_assert(DMAX == 16 || DMAX == 32 || DMAX == 64 || DMAX == 128, "DMAX must be 16, 32, 64 or 128");
    constexpr int NG = DMAX < 32 ? 1 : DMAX / 32;  // groups holding columns
    constexpr int GW = DMAX < 32 ? DMAX : 32;      // chains per group that can hold a column
    __shared__ __attribute__((aligned(16))) float ls[kShortRefs * DMAX];
    __shared__ __attribute__((aligned(16))) float lw[DMAX];
    const int t = threadIdx.x;
    for (uint32_t e = t; e < static_cast<uint32_t>(kShortRefs * DMAX); e += kShortThreads) {
        const uint32_t j = e / DMAX;
        const uint32_t c = e % DMAX;
        const uint32_t r = r0 + j;
        ls[e] = (r < rows && c < dim) ? refs[static_cast<uint64_t>(r) * ld_ref + c] : 0.0F;
    }
    if constexpr (WEIGHTED) {
        for (uint32_t c = t; c < static_cast<uint32_t>(DMAX); c += kShortThreads) lw[c] = c < dim ? weights[c] : 0.0F;
    }
    const uint32_t q = q0 + static_cast<uint32_t>(t);
    const float* qrow = queries + static_cast<uint64_t>(min(q, nq - 1U)) * ld_q;
    float qv[DMAX];
#pragma unroll
    for (int c = 0; c < DMAX; ++c) qv[c] = static_cast<uint32_t>(c) < dim ? qrow[c] : 0.0F;
    __syncthreads();

    unsigned long long best = ~0ULL;
    const uint32_t nref = min(static_cast<uint32_t>(kShortRefs), rows - r0);
    for (uint32_t j = 0; j < nref; ++j) {
        const float* rj = ls + j * DMAX;
        float W[NG];
#pragma unroll
        for (int g = 0; g < NG; ++g) {
            float x[GW];
#pragma unroll
            for (int i = 0; i < GW; i += 4) {
                const float4 rv = *reinterpret_cast<const float4*>(rj + 32 * g + i);
                float4 wv = make_float4(0.0F, 0.0F, 0.0F, 0.0F);
                if constexpr (WEIGHTED) wv = *reinterpret_cast<const float4*>(lw + 32 * g + i);
                x[i] = accumulate<METRIC, WEIGHTED, WFMA>(0.0F, rv.x, qv[32 * g + i], wv.x);
                x[i + 1] = accumulate<METRIC, WEIGHTED, WFMA>(0.0F, rv.y, qv[32 * g + i + 1], wv.y);
                x[i + 2] = accumulate<METRIC, WEIGHTED, WFMA>(0.0F, rv.z, qv[32 * g + i + 2], wv.z);
                x[i + 3] = accumulate<METRIC, WEIGHTED, WFMA>(0.0F, rv.w, qv[32 * g + i + 3], wv.w);
            }
            // Offsets 16 .. 1; with GW = 16 the offset-16 level only adds the
            // empty chains 16 .. 31.
#pragma unroll
            for (int off = 16; off >= 1; off /= 2) {
                if (off >= GW) continue;
#pragma unroll
                for (int i = 0; i < off; ++i) x[i] = x[i] + x[i + off];
            }
            W[g] = x[0];
        }
        float dist;
        if constexpr (NG == 1) {
            dist = W[0];
        } else if constexpr (NG == 2) {
            dist = W[0] + W[1];
        } else {
            dist = (W[0] + W[2]) + (W[1] + W[3]);
        }
        const unsigned long long key = (static_cast<unsigned long long>(__float_as_uint(dist)) << 32) | (r0 + j);
        best = key < best ? key : best;
    }
    if (q < nq) atomicMin(keys + q, best);
}


// A wave-uniform 64-bit value moved to scalar registers.
__device__ __forceinline__ uint64_t uniform_u64(uint64_t x)
{
    const uint32_t lo = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(x));
    const uint32_t hi = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(x >> 32));
    return (static_cast<uint64_t>(hi) << 32) | lo;
}

// Terms base .. base+31 of the 32 chains of one group (chain base%256 + i
// takes column base + i), for TQ queries, added to v. Columns at or past dim
// leave their chain unchanged. At TQ <= 2, chunks that lie wholly below dim
// skip the per-column checks. The reference column is read through a buffer
// resource at the group's first column: scalar column offset plus the
// thread's 32-bit byte offset roff.
template <int METRIC, bool WEIGHTED, bool WFMA, int TQ>
__device__ __forceinline__ void pair_group_terms(float (&v)[TQ][32], const float* __restrict__ refs_t, uint64_t ld_t,
                                                 uint32_t roff, const float* qs, uint32_t ldqs, const float* ws,
                                                 uint32_t base, uint32_t dim)
{
    // Column base + i sits i * ld_t * 4 bytes past the resource base; with
    // rows < 2^25 (ld_t == rows) every offset stays below 2^32.
    const __amdgpu_buffer_rsrc_t cols = __builtin_amdgcn_make_buffer_rsrc(
        const_cast<float*>(refs_t + static_cast<uint64_t>(base) * ld_t), 0, -1, kBufferWord3);
    const uint32_t col_bytes = static_cast<uint32_t>(ld_t) * 4U;
    float rv[32];
    // At TQ = 4 the unchecked form needs more registers than two waves per
    // SIMD allow, so it keeps the checked form.
    if (TQ <= 2 && base + 32U <= dim) {
#pragma unroll
        for (int i = 0; i < 32; ++i) {
            rv[i] = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(cols, roff, i * col_bytes, 0));
        }
        // Eight columns at a time; the fence keeps each chunk's query reads
        // next to their use.
#pragma unroll
        for (int j = 0; j < 4; ++j) {
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                const int i = 8 * j + k;
                const float wv = WEIGHTED ? ws[base + i] : 0.0F;
#pragma unroll
                for (int q = 0; q < TQ; ++q) {
                    v[q][i] = accumulate<METRIC, WEIGHTED, WFMA>(v[q][i], rv[i], qs[q * ldqs + base + i], wv);
                }
            }
            __builtin_amdgcn_sched_barrier(0);
        }
    } else {
#pragma unroll
        for (int i = 0; i < 32; ++i) {
            // base < dim: column 0 needs no clamp (a clamped offset is
            // computed on the vector unit, and a vector soffset would cost
            // a readfirstlane loop around the load)
            const uint32_t c = i == 0 ? base : min(base + i, dim - 1U);
            rv[i] = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(cols, roff, (c - base) * col_bytes, 0));
        }
#pragma unroll
        for (int i = 0; i < 32; ++i) {
            const uint32_t c = base + i;
            const float wv = WEIGHTED && c < dim ? ws[c] : 0.0F;
#pragma unroll
            for (int q = 0; q < TQ; ++q) {
                if (c < dim) v[q][i] = accumulate<METRIC, WEIGHTED, WFMA>(v[q][i], rv[i], qs[q * ldqs + c], wv);
            }
        }
    }
}

// Query values of column c for TQ queries from the column-major LDS copy
// (all lanes read the same address: a broadcast).
template <int TQ>
__device__ __forceinline__ void lds_query_column(const float* qs, uint32_t c, float (&qv)[TQ])
{
    const float* p = qs + static_cast<uint64_t>(c) * TQ;
    if constexpr (TQ % 4 == 0) {
#pragma unroll
        for (int k = 0; k < TQ / 4; ++k) {
            const float4 x = *reinterpret_cast<const float4*>(p + 4 * k);
            qv[4 * k] = x.x;
            qv[4 * k + 1] = x.y;
            qv[4 * k + 2] = x.z;
            qv[4 * k + 3] = x.w;
        }
    } else if constexpr (TQ == 2) {
        const float2 x = *reinterpret_cast<const float2*>(p);
        qv[0] = x.x;
        qv[1] = x.y;
    } else {
        qv[0] = p[0];
    }
}

// One chain class of a group in the pair-per-thread scan: chains 32w + m + 4s
// (s = 0..7) of this thread's reference row, every term of each chain in
// column order, then their part of the group's 32-lane tree (offsets 16, 8,
// 4 between chains are s offsets 4, 2, 1). cls[q] receives the class value
// (the group sum is (d0 + d2) + (d1 + d3) over the classes m).
template <int METRIC, bool WEIGHTED, bool WFMA, int TQ>
__device__ __forceinline__ void pair_class(float (&cls)[TQ], const float* __restrict__ refs_t, uint64_t ld_t,
                                           uint32_t roff, const float* qs, const float* ws, uint32_t first,
                                           uint32_t dim)
{
    float v[TQ][8];
#pragma unroll
    for (int q = 0; q < TQ; ++q)
#pragma unroll
        for (int s = 0; s < 8; ++s) v[q][s] = 0.0F;
    const uint32_t col_bytes = static_cast<uint32_t>(ld_t) * 4U;
    for (uint32_t base = first; base < dim; base += kChains) {
        // Columns base + 4s: the resource sits at column base, column
        // base + 4s is 4s rows of refs_t past it (rows < 2^25 keep every
        // offset below 2^32); a column at or past dim is read clamped and
        // not added.
        const __amdgpu_buffer_rsrc_t cols = __builtin_amdgcn_make_buffer_rsrc(
            const_cast<float*>(refs_t + static_cast<uint64_t>(base) * ld_t), 0, -1, kBufferWord3);
        float rv[8];
        const bool full = base + 28U < dim;
#pragma unroll
        for (int s = 0; s < 8; ++s) {
            const uint32_t c = full || s == 0 ? base + 4U * s : min(base + 4U * s, dim - 1U);
            rv[s] = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(cols, roff, (c - base) * col_bytes, 0));
        }
#pragma unroll
        for (int s = 0; s < 8; ++s) {
            const uint32_t c = base + 4U * s;
            if (full || c < dim) {
                float qv[TQ];
                lds_query_column<TQ>(qs, c, qv);
                const float wv = WEIGHTED ? ws[c] : 0.0F;
                float dd[TQ];
#pragma unroll
                for (int q = 0; q < TQ; ++q) dd[q] = rv[s] - qv[q];
                issue_fence<TQ>(dd);
#pragma unroll
                for (int q = 0; q < TQ; ++q) v[q][s] = accumulate_diff<METRIC, WEIGHTED, WFMA>(v[q][s], dd[q], wv);
            }
        }
    }
#pragma unroll
    for (int q = 0; q < TQ; ++q) {
#pragma unroll
        for (int off = 4; off >= 1; off /= 2) {
#pragma unroll
            for (int s = 0; s < off; ++s) v[q][s] = v[q][s] + v[q][s + off];
        }
        cls[q] = v[q][0];
    }
}

// One unit of a chain class: columns base + 4s (s = 0..7) of the eight
// chains, their reference values in rv, their query values (and weights) at
// local columns lbase + 4s of the group's LDS copy; INIT starts the chains
// (the first term onto +0). Columns past dim read zeros (padded reference
// and query copies, weight 0) and add +0, which changes no bit of a
// non-negative sum.
template <int METRIC, bool WEIGHTED, bool WFMA, int TQ, bool INIT>
__device__ __forceinline__ void cls_unit(float (&v)[TQ][8], const float (&rv)[8], const float* qs, const float* ws,
                                         uint32_t lbase)
{
#pragma unroll
    for (int s = 0; s < 8; ++s) {
        const uint32_t c = lbase + 4U * s;
        float qv[TQ];
        lds_query_column<TQ>(qs, c, qv);
        const float wv = WEIGHTED ? ws[c] : 0.0F;
#pragma unroll
        for (int q = 0; q < TQ; ++q)
            v[q][s] = accumulate_diff<METRIC, WEIGHTED, WFMA>(INIT ? 0.0F : v[q][s], rv[s] - qv[q], wv);
    }
}

// One group's sum W by chain classes, visited in the order m = 0, 2, 1, 3:
// W = (d0 + d2) + (d1 + d3), a class without columns (m >= dim - gbase)
// being +0. Each unit's eight reference loads are issued one unit ahead,
// across class boundaries too, so a wave has loads in flight while it adds.
// refs_t holds zero rows past dim (a unit reads up to column dim + 27); qs
// and ws hold column c at local column c (the whole row) or, LOCAL, column
// gbase + j + 256 t at local column 32 t + j (the group's columns only).
// Needs gbase < dim.
template <int METRIC, bool WEIGHTED, bool WFMA, int TQ, bool LOCAL>
__device__ __forceinline__ void pair_group_cls(float (&gsum)[TQ], const float* __restrict__ refs_t, uint64_t ld_t,
                                               uint32_t roff, const float* qs, const float* ws, uint32_t gbase,
                                               uint32_t dim)
{
    const uint32_t l0 = LOCAL ? 0U : gbase;
    const uint32_t col_bytes = static_cast<uint32_t>(ld_t) * 4U;
    // Columns base + 4s: the resource sits at column base, column base + 4s
    // is 4s rows of refs_t past it (rows < 2^25 keep the offsets below 2^32).
    auto load_unit = [&](uint32_t base, float (&r)[8]) __attribute__((always_inline)) {
        const __amdgpu_buffer_rsrc_t cols = __builtin_amdgcn_make_buffer_rsrc(
            const_cast<float*>(refs_t + static_cast<uint64_t>(base) * ld_t), 0, -1, kBufferWord3);
#pragma unroll
        for (int s = 0; s < 8; ++s)
            r[s] = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(cols, roff, 4U * s * col_bytes, 0));
    };
    const uint32_t span = dim - gbase;  // columns from gbase on (>= 1)
    if (span >= 4U && span <= kChains) {
        // one unit per class: two reference buffers in turn, no copies
        auto tree = [&](float (&v)[TQ][8], float (&d)[TQ]) __attribute__((always_inline)) {
#pragma unroll
            for (int q = 0; q < TQ; ++q) {
#pragma unroll
                for (int off = 4; off >= 1; off /= 2) {
#pragma unroll
                    for (int s = 0; s < off; ++s) v[q][s] = v[q][s] + v[q][s + off];
                }
                d[q] = v[q][0];
            }
        };
        float ra[8], rb[8], v[TQ][8], d0[TQ], d2[TQ], d1[TQ], d3[TQ];
        load_unit(gbase, ra);
        load_unit(gbase + 2U, rb);
        cls_unit<METRIC, WEIGHTED, WFMA, TQ, true>(v, ra, qs, ws, l0);
        tree(v, d0);
        load_unit(gbase + 1U, ra);
        cls_unit<METRIC, WEIGHTED, WFMA, TQ, true>(v, rb, qs, ws, l0 + 2U);
        tree(v, d2);
        load_unit(gbase + 3U, rb);
        cls_unit<METRIC, WEIGHTED, WFMA, TQ, true>(v, ra, qs, ws, l0 + 1U);
        tree(v, d1);
        cls_unit<METRIC, WEIGHTED, WFMA, TQ, true>(v, rb, qs, ws, l0 + 3U);
        tree(v, d3);
#pragma unroll
        for (int q = 0; q < TQ; ++q) gsum[q] = (d0[q] + d2[q]) + (d1[q] + d3[q]);
        return;
    }
    float rn[8];
    load_unit(gbase, rn);
    // Class m (first column gbase + m < dim, its data in rn); the loads run
    // one unit ahead, into `next` (the first column of the next class in the
    // order that has columns, or dim for none) after the class's last unit.
    auto run_class = [&](uint32_t m, uint32_t next, float (&d)[TQ]) __attribute__((always_inline)) {
        float v[TQ][8];
        uint32_t cur = gbase + m, lcur = m;
        auto step = [&](auto init) __attribute__((always_inline)) {
            float rv[8];
#pragma unroll
            for (int s = 0; s < 8; ++s) rv[s] = rn[s];
            const uint32_t nxt = cur + kChains < dim ? cur + kChains : next;
            if (nxt < dim) load_unit(nxt, rn);
            cls_unit<METRIC, WEIGHTED, WFMA, TQ, decltype(init)::value>(v, rv, qs, ws, LOCAL ? lcur : cur);
        };
        step(BoolTag<true>{});
        while (cur + kChains < dim) {
            cur += kChains;
            if constexpr (LOCAL) lcur += 32U;
            step(BoolTag<false>{});
        }
        // the class's part of the 32-lane tree (chain offsets 16, 8, 4)
#pragma unroll
        for (int q = 0; q < TQ; ++q) {
#pragma unroll
            for (int off = 4; off >= 1; off /= 2) {
#pragma unroll
                for (int s = 0; s < off; ++s) v[q][s] = v[q][s] + v[q][s + off];
            }
            d[q] = v[q][0];
        }
    };
    float d0[TQ], d2[TQ], d1[TQ], d3[TQ];
#pragma unroll
    for (int q = 0; q < TQ; ++q) d2[q] = d1[q] = d3[q] = 0.0F;
    run_class(0U, span > 2U ? gbase + 2U : (span > 1U ? gbase + 1U : dim), d0);
    if (span > 2U) run_class(2U, gbase + 1U, d2);
    if (span > 1U) run_class(1U, span > 3U ? gbase + 3U : dim, d1);
    if (span > 3U) run_class(3U, dim, d3);
#pragma unroll
    for (int q = 0; q < TQ; ++q) gsum[q] = (d0[q] + d2[q]) + (d1[q] + d3[q]);
}

// Pair-per-thread scan by chain classes (the default for short rows): each
// thread owns one reference row and TQ <= 8 queries. A group of 32 chains is
// taken a chain class at a time (chains 32w + m + 4s, m in the tree order
// 0, 2, 1, 3), so 8 x TQ partial sums stay live instead of 32 x TQ and a
// loaded reference value serves up to 8 queries; each chain still adds its
// terms in column order and the additions of the tree are the reference's
// (W_w = (d0 + d2) + (d1 + d3) with d_m the class sums, then
// ((W0+W4)+(W2+W6))+((W1+W5)+(W3+W7))), so the bits are the reference's.
// Queries (and L1 weights) sit in LDS column-major (the TQ values of a
// column together, read as broadcast vectors), zeros past dim. WHOLE: the
// whole row once (rows up to kClsWholeRowDim coordinates); otherwise a
// pass's two groups at a time (group w: columns 32w + j + 256 t at local
// column 32 t + j), which keeps the block's LDS small enough for several
// blocks per CU. A block needs cls_lds_floats(TQ, WEIGHTED, dim) floats.
constexpr uint32_t kClsWholeRowDim = 1536;
__host__ __device__ constexpr uint32_t cls_lds_floats(int tq, bool weighted, uint32_t dim)
{
    return dim <= kClsWholeRowDim
               ? static_cast<uint32_t>(tq + (weighted ? 1 : 0)) * ((dim + 31U) & ~3U)
               : 2U * static_cast<uint32_t>(tq + (weighted ? 1 : 0)) * 32U * ((dim + 255U) / 256U);
}

template <int METRIC, bool WEIGHTED, bool WFMA, int TQ, bool WHOLE>
__device__ __forceinline__ void scan_pairs_cls(const float* __restrict__ refs_t, uint64_t ld_t, uint32_t rows,
                                               const float* __restrict__ queries, uint64_t ld_q, uint32_t nq,
                                               const float* __restrict__ weights, uint32_t dim,
                                               unsigned long long* __restrict__ keys, uint32_t q0, uint32_t rblock,
                                               float* smem)
{
    static_assert(TQ % 4 == 0, "the chain-class body takes 4 or 8 queries per block");
    refs_t = reinterpret_cast<const float*>(uniform_u64(reinterpret_cast<uint64_t>(refs_t)));
    ld_t = uniform_u64(ld_t);
    rows = __builtin_amdgcn_readfirstlane(rows);
    dim = __builtin_amdgcn_readfirstlane(dim);
    const int t = threadIdx.x;
    // one column's TQ query values (coalesced row reads across the block)
    // as one or two 16-byte stores, and its weight
    auto stage_column = [&](uint32_t c, float* qs, float* ws) __attribute__((always_inline)) {
        float x[TQ];
#pragma unroll
        for (int qi = 0; qi < TQ; ++qi)
            x[qi] = c < dim ? queries[static_cast<uint64_t>(min(q0 + qi, nq - 1U)) * ld_q + c] : 0.0F;
#pragma unroll
        for (int k = 0; k < TQ / 4; ++k)
            *reinterpret_cast<float4*>(qs + 4 * k) = make_float4(x[4 * k], x[4 * k + 1], x[4 * k + 2], x[4 * k + 3]);
        if constexpr (WEIGHTED) *ws = c < dim ? weights[c] : 0.0F;
    };
    // Groups in bit-reversed order 0,4 | 2,6 | 1,5 | 3,7: pass p produces
    // (W0+W4), (W2+W6), (W1+W5), (W3+W7) in turn, and the distance is
    // (P0 + P1) + (P2 + P3).
    float P[TQ][4];
    uint32_t r;
    if constexpr (WHOLE) {
        const uint32_t ldqs = (dim + 31U) & ~3U;  // zero columns up to dim + 27
        float* qs = smem;               // ldqs x TQ, column-major
        float* ws = smem + TQ * ldqs;   // ldqs (weighted L1)
        for (uint32_t c = t; c < ldqs; c += 256) stage_column(c, qs + c * TQ, ws + c);
        __syncthreads();
        r = rblock * 256 + t;
        const uint32_t roff = min(r, rows - 1U) * 4U;
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
                const uint32_t gbase = static_cast<uint32_t>(32 * w);
                float gsum[TQ];
                if (gbase >= dim) {
#pragma unroll
                    for (int q = 0; q < TQ; ++q) gsum[q] = 0.0F;
                } else {
                    pair_group_cls<METRIC, WEIGHTED, WFMA, TQ, false>(gsum, refs_t, ld_t, roff, qs, ws, gbase, dim);
                }
#pragma unroll
                for (int q = 0; q < TQ; ++q) pair_sum[q] = h == 0 ? gsum[q] : pair_sum[q] + gsum[q];
            }
#pragma unroll
            for (int q = 0; q < TQ; ++q)
#pragma unroll
                for (int k = 0; k < 4; ++k) P[q][k] = pass == k ? pair_sum[q] : P[q][k];
        }
    } else {
        // two slots of gq floats (group w: column 32w + j + 256t at local
        // column 32t + j), the weights after them
        const uint32_t tmax = (dim + 255U) / 256U;  // terms of chain 0
        const uint32_t gq = TQ * 32U * tmax;
        float* wsbase = smem + 2U * gq;
        r = rblock * 256 + t;
        const uint32_t roff = min(r, rows - 1U) * 4U;
#pragma unroll
        for (int q = 0; q < TQ; ++q)
#pragma unroll
            for (int k = 0; k < 4; ++k) P[q][k] = 0.0F;
#pragma unroll 1
        for (int pass = 0; pass < 4; ++pass) {
            // the pass's two groups (after the previous pass's reads)
            __syncthreads();
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const int g = 2 * pass + h;
                const uint32_t w = static_cast<uint32_t>(((g & 1) << 2) | (g & 2) | ((g & 4) >> 2));
                if (32U * w >= dim) continue;
                const uint32_t n = 32U * ((dim - 32U * w + 255U) / 256U);
                for (uint32_t e = t; e < n; e += 256)
                    stage_column(32U * w + (e & 31U) + 256U * (e >> 5), smem + h * gq + e * TQ,
                                 wsbase + h * 32U * tmax + e);
            }
            __syncthreads();
            float pair_sum[TQ];
#pragma unroll 1
            for (int h = 0; h < 2; ++h) {
                const int g = 2 * pass + h;
                const int w = ((g & 1) << 2) | (g & 2) | ((g & 4) >> 2);  // bit reversal of g
                const uint32_t gbase = static_cast<uint32_t>(32 * w);
                float gsum[TQ];
                if (gbase >= dim) {
#pragma unroll
                    for (int q = 0; q < TQ; ++q) gsum[q] = 0.0F;
                } else {
                    pair_group_cls<METRIC, WEIGHTED, WFMA, TQ, true>(gsum, refs_t, ld_t, roff, smem + h * gq,
                                                                     wsbase + h * 32U * tmax, gbase, dim);
                }
#pragma unroll
                for (int q = 0; q < TQ; ++q) pair_sum[q] = h == 0 ? gsum[q] : pair_sum[q] + gsum[q];
            }
#pragma unroll
            for (int q = 0; q < TQ; ++q)
#pragma unroll
                for (int k = 0; k < 4; ++k) P[q][k] = pass == k ? pair_sum[q] : P[q][k];
        }
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

// Pair-per-thread form for short rows (flat |I|, hierarchical p): the
// 256-chain tree costs 255 adds per pair whatever the row length, and in the
// tiled kernel above it is a cross-lane tree, which dominates when a chain
// holds one or two terms. Here each thread owns one reference row and TQ
// queries and evaluates the whole reference order in registers: for each
// group w of 32 chains, chain 32w+i sums columns 32w+i, 32w+i+256, ... in
// order, then the 32-lane tree (offsets 16..1) gives W_w, and
// ((W0+W4)+(W2+W6))+((W1+W5)+(W3+W7)) the distance -- the same additions as
// the reference, so the same bits. A group with no columns (32w >= dim) has
// W_w = +0, and adding +0 to a sum of non-negative terms changes no bit, so
// it is skipped. References are read from a transposed (dim x rows) copy, so
// a wave's loads are contiguous (rows < 2^25, so a group's 32 columns lie
// within 2^32 bytes); queries (and L1 weights) sit in LDS with a row stride padded to a
// multiple of four and are read as broadcast 16-byte vectors.
template <int METRIC, bool WEIGHTED, bool WFMA, int TQ>
__device__ __forceinline__ void scan_pairs(const float* __restrict__ refs_t, uint64_t ld_t, uint32_t rows,
                                           const float* __restrict__ queries, uint64_t ld_q, uint32_t nq,
                                           const float* __restrict__ weights, uint32_t dim,
                                           unsigned long long* __restrict__ keys, uint32_t q0, uint32_t rblock,
                                           float* smem)
{
    // The buffer resource and the column offsets must be scalar (a caller
    // that read these from memory would otherwise get a per-lane loop around
    // every load); they are uniform in every caller.
    refs_t = reinterpret_cast<const float*>(uniform_u64(reinterpret_cast<uint64_t>(refs_t)));
    ld_t = uniform_u64(ld_t);
    rows = __builtin_amdgcn_readfirstlane(rows);
    dim = __builtin_amdgcn_readfirstlane(dim);
    const uint32_t ldqs = (dim + 3U) & ~3U;
    float* qs = smem;               // TQ x ldqs
    float* ws = smem + TQ * ldqs;   // ldqs (weighted L1)
    const int t = threadIdx.x;
    for (uint32_t e = t; e < TQ * ldqs; e += 256) {
        const uint32_t qi = e / ldqs;
        const uint32_t c = e % ldqs;
        qs[e] = c < dim ? queries[static_cast<uint64_t>(min(q0 + qi, nq - 1U)) * ld_q + c] : 0.0F;
    }
    if constexpr (WEIGHTED) {
        for (uint32_t c = t; c < ldqs; c += 256) ws[c] = c < dim ? weights[c] : 0.0F;
    }
    __syncthreads();
    const uint32_t r = rblock * 256 + t;
    const uint32_t roff = min(r, rows - 1U) * 4U;
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
            const uint32_t gbase = static_cast<uint32_t>(32 * w);
            if (gbase >= dim) {
                if (h == 0) {
#pragma unroll
                    for (int q = 0; q < TQ; ++q) pair_sum[q] = 0.0F;
                }
                continue;
            }
            float v[TQ][32];
#pragma unroll
            for (int i = 0; i < 32; ++i)
#pragma unroll
                for (int q = 0; q < TQ; ++q) v[q][i] = 0.0F;
            // Term k of chain 32w+i is column 32w + i + 256k: all 32 chains of
            // the group load term k together (32 independent loads in flight),
            // and each chain still adds its terms in increasing k.
            for (uint32_t base = gbase; base < dim; base += kChains) {
                pair_group_terms<METRIC, WEIGHTED, WFMA, TQ>(v, refs_t, ld_t, roff, qs, ldqs, ws, base, dim);
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

// Pair-per-thread form used by the serving kernel (rows of many small
// indexes, read through global loads; measured faster there than the
// buffer-load form below). Same chains and tree as scan_pairs: the
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
__device__ __forceinline__ void scan_pairs_serve(const float* __restrict__ refs_t, uint64_t ld_t, uint32_t rows,
                                           const float* __restrict__ queries, uint64_t ld_q, uint32_t nq,
                                           const float* __restrict__ weights, uint32_t dim,
                                           unsigned long long* __restrict__ keys, uint32_t q0, uint32_t rblock,
                                           float* smem)
{
    float* qs = smem;              // TQ x dim
    float* ws = smem + TQ * dim;   // dim (weighted L1)
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
    const uint32_t r = rblock * 256 + t;
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
            if (static_cast<uint32_t>(32 * w) >= dim) {  // no columns: W_w = +0 changes no bit
                if (h == 0) {
#pragma unroll
                    for (int q = 0; q < TQ; ++q) pair_sum[q] = 0.0F;
                }
                continue;
            }
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

// Gather tile: sampled-column chunk `chunk` (kGatherChunk columns) of query
// rows row_block*kGatherRows ... 
__device__ __forceinline__ void gather_tile(const float* __restrict__ src, uint64_t ld_src, uint32_t nq,
                                            const uint32_t* __restrict__ cols, const float* __restrict__ scales,
                                            uint32_t k, float* __restrict__ out, uint64_t ld_out, int* invalid,
                                            uint32_t chunk, uint32_t row_block)
{
    __shared__ uint32_t s_cols[kGatherChunk];
    __shared__ float s_scales[kGatherChunk];
    const uint32_t c0 = chunk * kGatherChunk;
    const uint32_t n = min(static_cast<uint32_t>(kGatherChunk), k - c0);
    for (uint32_t i = threadIdx.x; i < n; i += kGatherThreads) {
        s_cols[i] = cols[c0 + i];
        s_scales[i] = scales != nullptr ? scales[c0 + i] : 1.0F;
    }
    __syncthreads();
    bool bad = false;
    const uint32_t q_begin = row_block * kGatherRows;
    const uint32_t q_end = min(q_begin + kGatherRows, nq);
    // All of the thread's gathered values (kGatherRows rows x kPer columns)
    // are loaded before any is stored, so their loads are in flight together.
    constexpr int kPer = kGatherChunk / kGatherThreads;
    float v[kGatherRows][kPer];
#pragma unroll
    for (int rq = 0; rq < kGatherRows; ++rq) {
        const uint32_t q = q_begin + static_cast<uint32_t>(rq);
        const float* row = src + static_cast<uint64_t>(min(q, nq - 1U)) * ld_src;
#pragma unroll
        for (int u = 0; u < kPer; ++u) {
            const uint32_t i = threadIdx.x + static_cast<uint32_t>(u) * kGatherThreads;
            v[rq][u] = (q < q_end && i < n) ? row[s_cols[i]] : 0.0F;
        }
    }
#pragma unroll
    for (int rq = 0; rq < kGatherRows; ++rq) {
        const uint32_t q = q_begin + static_cast<uint32_t>(rq);
        if (q >= q_end) break;
        float* dst = out + static_cast<uint64_t>(q) * ld_out + c0;
#pragma unroll
        for (int u = 0; u < kPer; ++u) {
            const uint32_t i = threadIdx.x + static_cast<uint32_t>(u) * kGatherThreads;
            if (i < n) {
                const float x = v[rq][u];
                const float scaled = scales != nullptr ? x * s_scales[i] : x;
                bad |= !__builtin_isfinite(x) || !__builtin_isfinite(scaled);
                dst[i] = scaled;
            }
        }
    }
    if (bad) atomicExch(invalid, 1);
}

// One kProjQ x kProjP tile of the binary32 projection (K7 order).
__device__ __forceinline__ void project_fp32_tile(const float* __restrict__ x, uint64_t ld_x, uint32_t nq,
                                                  uint32_t k, const float* __restrict__ proj, uint64_t ld_proj,
                                                  uint32_t p, float* __restrict__ y, uint64_t ld_y, int* invalid,
                                                  uint32_t q0, uint32_t o0)
{
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

// One 32 x 32 tile of the bf16 MFMA projection.
__device__ __forceinline__ void project_bf16_tile(const float* __restrict__ x, uint64_t ld_x, uint32_t nq,
                                                  uint32_t k, const __bf16* __restrict__ proj, uint64_t ld_proj,
                                                  uint32_t p, float* __restrict__ y, uint64_t ld_y, int* invalid,
                                                  uint32_t q0, uint32_t o0)
{
    const int lane = threadIdx.x % 64;
    const int wave = threadIdx.x / 64;
    // MFMA operand fragments: lane holds row (lane % 32) and K values
    // 8*(lane/32) .. +7 of each 16-wide K step. Rows of x are a multiple of
    // four floats long, so a float4 that starts below k stays in the row.
    const float* xr = x + static_cast<uint64_t>(min(q0 + lane % 32, nq - 1U)) * ld_x + 8 * (lane / 32);
    const __bf16* pr = proj + static_cast<uint64_t>(o0 + lane % 32) * ld_proj + 8 * (lane / 32);
    const uint32_t steps = (k + 15U) / 16U;
    floatx16 acc;
#pragma unroll
    for (int v = 0; v < 16; ++v) acc[v] = 0.0F;
    for (uint32_t s = wave; s < steps; s += kProjWaves) {
        const uint32_t c0 = s * 16U + 8U * (lane / 32);
        float xv[8];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            float4 f = make_float4(0.0F, 0.0F, 0.0F, 0.0F);
            if (c0 + 4U * h < k) f = *reinterpret_cast<const float4*>(xr + s * 16U + 4U * h);
            xv[4 * h + 0] = f.x;
            xv[4 * h + 1] = f.y;
            xv[4 * h + 2] = f.z;
            xv[4 * h + 3] = f.w;
        }
        bf16x8 a;
#pragma unroll
        for (int e = 0; e < 8; ++e) a[e] = static_cast<__bf16>(c0 + e < k ? xv[e] : 0.0F);
        const bf16x8 b = *reinterpret_cast<const bf16x8*>(pr + s * 16U);
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
    }
    __shared__ float part[kProjWaves][32 * 32];
    // D layout: lane holds column (lane % 32) and rows 8*(v/4) + 4*(lane/32) + v%4.
#pragma unroll
    for (int v = 0; v < 16; ++v) {
        const int row = 8 * (v / 4) + 4 * (lane / 32) + (v % 4);
        part[wave][row * 32 + lane % 32] = acc[v];
    }
    __syncthreads();
    for (int e = threadIdx.x; e < 32 * 32; e += 64 * kProjWaves) {
        const uint32_t q = q0 + e / 32;
        const uint32_t o = o0 + e % 32;
        if (q < nq && o < p) {
            const float s = (part[0][e] + part[1][e]) + (part[2][e] + part[3][e]);
            if (!__builtin_isfinite(s)) atomicExch(invalid, 1);
            y[static_cast<uint64_t>(q) * ld_y + o] = s;
        }
    }
}

// Batcher's odd-even merge sort (ascending) as a compile-time network of
// compare-exchanges in registers: 63 / 191 / 543 / 1471 for N = 16 / 32 /
// 64 / 128, against the bitonic network's 80 / 240 / 672 / 1792.
template <int N>
__device__ __forceinline__ void oddeven_merge_sort(float (&v)[N])
{
#pragma unroll
    for (int p = 1; p < N; p *= 2) {
#pragma unroll
        for (int k = p; k >= 1; k /= 2) {
#pragma unroll
            for (int j = k % p; j + k < N; j += 2 * k) {
#pragma unroll
                for (int i = 0; i < k; ++i) {
                    if (i < N - j - k && (i + j) / (2 * p) == (i + j + k) / (2 * p)) {
                        const float a = v[i + j];
                        const float b = v[i + j + k];
                        v[i + j] = fminf(a, b);
                        v[i + j + k] = fmaxf(a, b);
                    }
                }
            }
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

// Median-L1 estimates of query q against references rblock*256 ... +255.
// EXACT: p == PMAX, so the middle positions are compile-time constants.
template <int PMAX, bool EXACT>
__device__ __forceinline__ void median_block(const float* __restrict__ refs_t, uint32_t rows,
                                             const float* __restrict__ yq, uint64_t ld_y, uint32_t nq, uint32_t p,
                                             unsigned long long* __restrict__ keys, int* invalid, uint32_t q,
                                             uint32_t rblock)
{
    __shared__ float qv[PMAX];
    __shared__ unsigned long long best_key[kMedianThreads / 64];
    if (q >= nq) return;
    const int t = threadIdx.x;
    for (uint32_t c = t; c < p; c += kMedianThreads) qv[c] = yq[static_cast<uint64_t>(q) * ld_y + c];
    __syncthreads();
    const uint32_t r = rblock * kMedianThreads + t;
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
        oddeven_merge_sort<PMAX>(v);
        float med = 0.0F;
        if constexpr (EXACT) {
            // Fixed middle positions: the compiler keeps only the
            // compare-exchanges that feed them (in the last merge stage 2, 4,
            // ..., then all of the two widest levels).
            med = v[PMAX / 2 - 1] + (v[PMAX / 2] - v[PMAX / 2 - 1]) * 0.5F;
        } else {
            const uint32_t mid = p / 2;
            float lower = 0.0F;
#pragma unroll
            for (int c = 0; c < PMAX; ++c) {
                if (static_cast<uint32_t>(c) == mid) med = v[c];
                if (static_cast<uint32_t>(c) + 1U == mid) lower = v[c];
            }
            if (p % 2 == 0) med = lower + (med - lower) * 0.5F;
        }
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

}  // namespace uhann::kern::body
