// Mixed-precision exact L2: BF16 Gram GEMM on MFMA, a bounded candidate set
// per query, and an exact re-rank.
//
// 1. to_bf16_rows: binary32 rows -> bf16 rows (round to nearest even), K
//    padded with zeros to a multiple of 64, plus per-row statistics:
//    ||x||^2 (binary64 sum, stored as binary32), ||x||_4, ||x^||_4, ||x^||_2.
// 2. gram_bf16: approx[q][r] = ||q||^2 + ||v||^2 - 2 <q^, v^> with the inner
//    product on v_mfma_f32_32x32x16_bf16 (FP32 accumulation). Blocks are a
//    grid of 64 x 64 wave tiles (128 x 128 up to 256 x 256), K step 64,
//    operands double-buffered in LDS with a register prefetch of the next K
//    step.
// 3. select_rerank (one block per query):
//    * per-pair bound tau(q,r) on |approx - exact|: the bf16 rounding error
//      of the inner product is sum_i (q_i dv_i + v^_i dq_i); with the rounding
//      signs modelled as independent and symmetric, Hoeffding bounds it by
//      K (||q||_4 ||dv||_4 + ||v^||_4 ||dq||_4) except with probability
//      2 exp(-K^2/2) (K = 7: 4.6e-11 per pair), where dv, dq are the rows'
//      actual rounding errors (stored per row); the FP32 accumulation over
//      ceil(d/16) MFMA steps adds K u_32 sqrt(ceil(d/16)) ||q^||_2 ||v^||_2,
//      and the norms and the three-term epilogue add 4 u_32 (||q||^2 + ||v||^2);
//    * U = min_r (approx + tau); candidates = {r : approx - tau <= U}. The
//      exact argmin is a candidate, and no non-candidate can tie it;
//    * the candidates are re-ranked with the reference's binary32 order
//      (256 chains + fixed tree, as in scan.hip), so the result is the plain
//      path's result, bit for bit;
//    * when several candidates lie within the binary32 error band of the
//      best (gamma_k, k = ceil(d/256) + 10), they are re-ranked once more in
//      binary64 and the binary64 minimum (lowest row on ties) wins;
//    * more than kMaxCandidates candidates -> the query is listed for an
//      exact direct scan (escalation).

#include <hip/hip_runtime.h>

#include <algorithm>
#include <stdexcept>
#include <string>

#include "gram.hpp"

namespace uhann::kern {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
typedef float floatx16 __attribute__((ext_vector_type(16)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

constexpr int kBK = 64;
constexpr int kPad = 8;                 // bf16 elements of row padding in LDS
constexpr int kLdsRow = kBK + kPad;     // 72 bf16 = 144 bytes
constexpr float kHoeffdingK = 7.0F;
constexpr float kU32 = 5.9604645e-08F;  // 2^-24

void check_launch(const char* what)
{
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        throw std::runtime_error(std::string("cannot launch ") + what + ": " +
                                 hipGetErrorString(e));
    }
}

__device__ __forceinline__ double block_sum_f64(double v, double* scratch)
{
    for (int off = 32; off >= 1; off /= 2) v += __shfl_down(v, off);
    const int lane = threadIdx.x % 64;
    const int wave = threadIdx.x / 64;
    __syncthreads();
    if (lane == 0) scratch[wave] = v;
    __syncthreads();
    double total = 0.0;
    const int waves = blockDim.x / 64;
    for (int w = 0; w < waves; ++w) total += scratch[w];
    return total;
}

// Four block sums at once (one pair of barriers), in the same order as
// block_sum_f64: a shuffle tree per wave, then the waves in order.
__device__ __forceinline__ double4 block_sum4_f64(double4 v, double4* scratch)
{
    for (int off = 32; off >= 1; off /= 2) {
        v.x += __shfl_down(v.x, off);
        v.y += __shfl_down(v.y, off);
        v.z += __shfl_down(v.z, off);
        v.w += __shfl_down(v.w, off);
    }
    const int lane = threadIdx.x % 64;
    const int wave = threadIdx.x / 64;
    __syncthreads();
    if (lane == 0) scratch[wave] = v;
    __syncthreads();
    double4 t = make_double4(0.0, 0.0, 0.0, 0.0);
    const int waves = blockDim.x / 64;
    for (int w = 0; w < waves; ++w) {
        t.x += scratch[w].x;
        t.y += scratch[w].y;
        t.z += scratch[w].z;
        t.w += scratch[w].w;
    }
    return t;
}

// One block of 256 threads per row.
__global__ void __launch_bounds__(256)
to_bf16_rows_kernel(const float* __restrict__ src, uint64_t ld_src, uint32_t rows, uint32_t dim,
                    __bf16* __restrict__ dst, uint64_t ld_dst, RowStats* __restrict__ stats)
{
    __shared__ double scratch[8];
    const uint32_t row = blockIdx.x;
    if (row >= rows) return;
    const float* s = src + static_cast<uint64_t>(row) * ld_src;
    __bf16* o = dst + static_cast<uint64_t>(row) * ld_dst;
    double n2 = 0.0, n4 = 0.0, h2 = 0.0, h4 = 0.0, e4 = 0.0;
    auto one = [&](float x) {
        const __bf16 h = static_cast<__bf16>(x);
        const double xd = x;
        const double hd = static_cast<float>(h);
        const double ed = xd - hd;
        n2 += xd * xd;
        n4 += (xd * xd) * (xd * xd);
        h2 += hd * hd;
        h4 += (hd * hd) * (hd * hd);
        e4 += (ed * ed) * (ed * ed);
        return h;
    };
    // Rows of 16-byte aligned source (ld_src % 4 == 0) move four values per
    // thread and step: one 16-byte load, one 8-byte store. ld_dst is a
    // multiple of 64, so the destination quads never straddle a row.
    const bool vec = ld_src % 4 == 0 && (reinterpret_cast<uintptr_t>(src) & 15U) == 0;
    uint32_t c0 = 0;
    if (vec) {
        const uint32_t full = dim / 4 * 4;
        const uint32_t step = blockDim.x * 4;
        auto quad = [&](const float4& x, uint32_t c) {
            bf16x4 h;
            h[0] = one(x.x);
            h[1] = one(x.y);
            h[2] = one(x.z);
            h[3] = one(x.w);
            *reinterpret_cast<bf16x4*>(o + c) = h;
        };
        // Four loads in flight per thread; each thread still visits its
        // columns in increasing order.
        uint32_t c = threadIdx.x * 4;
        for (; c + 3 * step < full; c += 4 * step) {
            float4 x[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) x[u] = *reinterpret_cast<const float4*>(s + c + u * step);
#pragma unroll
            for (int u = 0; u < 4; ++u) quad(x[u], c + u * step);
        }
        for (; c < full; c += step) quad(*reinterpret_cast<const float4*>(s + c), c);
        c0 = full;
    }
    for (uint32_t c = c0 + threadIdx.x; c < ld_dst; c += blockDim.x) o[c] = one(c < dim ? s[c] : 0.0F);
    n2 = block_sum_f64(n2, scratch);
    n4 = block_sum_f64(n4, scratch);
    h2 = block_sum_f64(h2, scratch);
    h4 = block_sum_f64(h4, scratch);
    e4 = block_sum_f64(e4, scratch);
    if (threadIdx.x == 0) {
        RowStats st{};
        st.norm2 = static_cast<float>(n2);
        // Rounded up a little so the stored bounds never understate.
        st.l4 = static_cast<float>(sqrt(sqrt(n4))) * 1.0001F;
        st.l4_hat = static_cast<float>(sqrt(sqrt(h4))) * 1.0001F;
        st.l2_hat = static_cast<float>(sqrt(h2)) * 1.0001F;
        st.d4 = static_cast<float>(sqrt(sqrt(e4))) * 1.0001F;
        st.scale = 1.0F;
        stats[row] = st;
    }
}

__device__ __forceinline__ uint32_t xcd_logical(uint32_t b, uint32_t nblocks)
{
    const uint32_t per = nblocks / 8U;
    const uint32_t rem = nblocks % 8U;
    const uint32_t x = b % 8U;
    return x * per + (x < rem ? x : rem) + b / 8U;
}

// The work of one Gram block: logical tile `tile` (grouped order, see
// tile_origin), K range [k_begin, k_end), and for a split tile its K part
// (-1 for a whole tile). Blocks [0, full) take whole tiles; blocks past that
// take K parts of the tail tiles, part-major so that blocks running together
// share the same K range of neighbouring tiles.
struct GramBlock {
    uint32_t tile;
    uint32_t k_begin;
    uint32_t k_end;
    int part;
};

__device__ __forceinline__ GramBlock gram_block(uint32_t b, const GramSplit& sp, uint32_t K, uint32_t kstep)
{
    if (b < sp.full) return {xcd_logical(b, sp.full), 0U, K, -1};
    const uint32_t u = xcd_logical(b - sp.full, sp.tail * sp.parts);
    const uint32_t part = u / sp.tail;
    const uint32_t kb = part * sp.part_steps * kstep;
    return {sp.full + u % sp.tail, kb, min(K, kb + sp.part_steps * kstep), static_cast<int>(part)};
}

// Output origin (in tiles) of logical tile L: groups of group_m tile rows are
// swept column by column, so consecutive logical tiles share operand tiles.
__device__ __forceinline__ uint2 tile_origin(uint32_t L, uint32_t tiles_m, uint32_t tiles_n, uint32_t group_m)
{
    const uint32_t span = group_m * tiles_n;
    const uint32_t gid = L / span;
    const uint32_t first_m = gid * group_m;
    const uint32_t gsize = min(group_m, tiles_m - first_m);
    const uint32_t within = L - gid * span;
    return make_uint2(first_m + within % gsize, within / gsize);
}

// A split tile's partial accumulators: unit (tail tile, part) holds NT
// threads x TM*TN*4 float4 chunks, chunk-major so that stores and loads are
// coalesced. Chunk c = (i*TN + j)*4 + v/4 holds acc[i][j][v .. v+3].
template <int NT, int TM, int TN>
__device__ __forceinline__ void store_partials(float* __restrict__ dst, const floatx16 (&acc)[TM][TN], int tid)
{
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j)
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                const int chunk = (i * TN + j) * 4 + c;
                *reinterpret_cast<float4*>(dst + (static_cast<uint64_t>(chunk) * NT + tid) * 4) =
                    make_float4(acc[i][j][4 * c], acc[i][j][4 * c + 1], acc[i][j][4 * c + 2], acc[i][j][4 * c + 3]);
            }
}

// out[m][n] = ||q||^2 + ||v||^2 - 2 acc (SCALED: - 2 (s_q s_v acc)) for the
// wave's TM x TN 32x32 accumulators at rows mw.., columns nw... D layout of
// the 32x32 MFMA: lane holds column (lane % 32) and rows
// 8*(v/4) + 4*(lane/32) + v%4 for its 16 values v.
template <int TM, int TN, bool SCALED>
__device__ __forceinline__ void gram_epilogue(const floatx16 (&acc)[TM][TN], uint32_t mw, uint32_t nw, int lane,
                                              uint32_t M, uint32_t N, const RowStats* __restrict__ qs,
                                              const RowStats* __restrict__ vs, float* __restrict__ out,
                                              uint64_t ldo)
{
#pragma unroll
    for (int i = 0; i < TM; ++i) {
#pragma unroll
        for (int j = 0; j < TN; ++j) {
            const uint32_t n = nw + j * 32 + (lane % 32);
            if (n >= N) continue;
            const float nv = vs[n].norm2;
            const float sv = SCALED ? vs[n].scale : 1.0F;
#pragma unroll
            for (int v = 0; v < 16; ++v) {
                const uint32_t m = mw + i * 32 + 8 * (v / 4) + 4 * (lane / 32) + (v % 4);
                if (m < M) {
                    const float base = qs[m].norm2 + nv;
                    float r;
                    if constexpr (SCALED) {
                        r = base - 2.0F * ((qs[m].scale * sv) * acc[i][j][v]);
                    } else {
                        r = base - 2.0F * acc[i][j][v];
                    }
                    out[static_cast<uint64_t>(m) * ldo + n] = r;
                }
            }
        }
    }
}

// Sums the K parts of one split tile in part order and applies the epilogue
// (one block per tail tile, the same thread layout as the Gram kernel).
template <int WM, int WN, int TM, int TN, bool SCALED>
__global__ void __launch_bounds__(64 * WM * WN)
gram_fixup_kernel(GramSplit sp, uint32_t M, uint32_t N, const RowStats* __restrict__ qs,
                  const RowStats* __restrict__ vs, float* __restrict__ out, uint64_t ldo, uint32_t tiles_m,
                  uint32_t tiles_n, uint32_t group_m)
{
    constexpr int NT = 64 * WM * WN;
    constexpr int BM = 32 * TM * WM;
    constexpr int BN = 32 * TN * WN;
    const int tid = threadIdx.x;
    const int lane = tid % 64;
    const int wave = tid / 64;
    const uint2 o = tile_origin(sp.full + blockIdx.x, tiles_m, tiles_n, group_m);
    floatx16 acc[TM][TN];
    for (uint32_t p = 0; p < sp.parts; ++p) {
        const float* src = sp.partials + (static_cast<uint64_t>(blockIdx.x) * sp.parts + p) * (BM * BN);
#pragma unroll
        for (int i = 0; i < TM; ++i)
#pragma unroll
            for (int j = 0; j < TN; ++j)
#pragma unroll
                for (int c = 0; c < 4; ++c) {
                    const int chunk = (i * TN + j) * 4 + c;
                    const float4 x = *reinterpret_cast<const float4*>(src + (static_cast<uint64_t>(chunk) * NT + tid) * 4);
                    if (p == 0) {
                        acc[i][j][4 * c] = x.x;
                        acc[i][j][4 * c + 1] = x.y;
                        acc[i][j][4 * c + 2] = x.z;
                        acc[i][j][4 * c + 3] = x.w;
                    } else {
                        acc[i][j][4 * c] += x.x;
                        acc[i][j][4 * c + 1] += x.y;
                        acc[i][j][4 * c + 2] += x.z;
                        acc[i][j][4 * c + 3] += x.w;
                    }
                }
    }
    gram_epilogue<TM, TN, SCALED>(acc, o.x * BM + (wave / WN) * 32 * TM, o.y * BN + (wave % WN) * 32 * TN, lane, M,
                                  N, qs, vs, out, ldo);
}

// Global -> register staging of one K step: an operand tile is ROWS x 64
// bf16 = ROWS * 8 chunks of 16 bytes; thread tid moves chunks tid + NT*c.
#define UHANN_GRAM_LOAD(k0, RA, RB)                                                 \
    _Pragma("unroll") for (int c = 0; c < CA; ++c) RA[c] =                          \
        *reinterpret_cast<const u32x4*>(qbase + c * qstep + (k0));                  \
    _Pragma("unroll") for (int c = 0; c < CB; ++c) RB[c] =                          \
        *reinterpret_cast<const u32x4*>(vbase + c * vstep + (k0));                  \
    __builtin_amdgcn_sched_barrier(0);
#define UHANN_GRAM_STORE(buf, RA, RB)                                               \
    _Pragma("unroll") for (int c = 0; c < CA; ++c)                                  \
        *reinterpret_cast<u32x4*>(As + (buf) * BM * kLdsRow + lds0 + c * kLdsStep) = RA[c]; \
    _Pragma("unroll") for (int c = 0; c < CB; ++c)                                  \
        *reinterpret_cast<u32x4*>(Bs + (buf) * BN * kLdsRow + lds0 + c * kLdsStep) = RB[c];
// Four K=16 MFMA steps over one LDS stage for the wave's TM x TN 32x32 tiles:
// TM + TN fragment reads feed TM * TN MFMAs.
#define UHANN_GRAM_COMPUTE(buf)                                                     \
    _Pragma("unroll") for (int kk = 0; kk < kBK; kk += 16)                          \
    {                                                                               \
        const __bf16* ab = As + (buf) * BM * kLdsRow + a_off + kk;                  \
        const __bf16* bb = Bs + (buf) * BN * kLdsRow + b_off + kk;                  \
        bf16x8 af[TM], bfr[TN];                                                     \
        _Pragma("unroll") for (int i = 0; i < TM; ++i) af[i] =                      \
            *reinterpret_cast<const bf16x8*>(ab + i * 32 * kLdsRow);                \
        _Pragma("unroll") for (int j = 0; j < TN; ++j) bfr[j] =                     \
            *reinterpret_cast<const bf16x8*>(bb + j * 32 * kLdsRow);                \
        _Pragma("unroll") for (int i = 0; i < TM; ++i)                              \
            _Pragma("unroll") for (int j = 0; j < TN; ++j) acc[i][j] =              \
                __builtin_amdgcn_mfma_f32_32x32x16_bf16(af[i], bfr[j], acc[i][j], 0, 0, 0); \
    }

// approx[m][n] = nq[m] + nv[n] - 2 <Q[m], V[n]> for m < M, n < N. Q is M x K
// and V is N x K bf16, row-major with leading dimensions ldq / ldv, K a
// multiple of kBK (zero padded). A block is a WM x WN grid of waves, each
// computing a (32 TM) x (32 TN) tile with TM x TN v_mfma_f32_32x32x16_bf16
// accumulators. Two LDS stages: while the MFMAs consume one, the next K step
// travels global -> registers -> the other, so each K step needs a single
// barrier.
template <int WM, int WN, int TM, int TN>
__global__ void __launch_bounds__(64 * WM * WN)
gram_bf16_kernel(const __bf16* __restrict__ Q, uint64_t ldq, uint32_t M,
                 const __bf16* __restrict__ V, uint64_t ldv, uint32_t N, uint32_t K,
                 const RowStats* __restrict__ qs, const RowStats* __restrict__ vs,
                 float* __restrict__ out, uint64_t ldo, uint32_t tiles_m, uint32_t tiles_n,
                 uint32_t group_m, GramSplit sp)
{
    constexpr int NT = 64 * WM * WN;
    constexpr int BM = 32 * TM * WM;
    constexpr int BN = 32 * TN * WN;
    static_assert((BM * 8) % NT == 0 && (BN * 8) % NT == 0, "tile rows must split evenly over the threads");
    constexpr int CA = BM * 8 / NT;  // 16-byte chunks per thread per K step
    constexpr int CB = BN * 8 / NT;
    __shared__ __attribute__((aligned(16))) __bf16 As[2 * BM * kLdsRow];
    __shared__ __attribute__((aligned(16))) __bf16 Bs[2 * BN * kLdsRow];

    const GramBlock gb = gram_block(blockIdx.x, sp, K, kBK);
    const uint2 o = tile_origin(gb.tile, tiles_m, tiles_n, group_m);
    const uint32_t m0 = o.x * BM;
    const uint32_t n0 = o.y * BN;

    const int tid = threadIdx.x;
    const int lane = tid % 64;
    const int wave = tid / 64;
    const int wm = (wave / WN) * 32 * TM;  // wave's offsets inside the block tile
    const int wn = (wave % WN) * 32 * TN;

    u32x4 ra[CA], rb[CB];
    // Chunk c of this thread is row tid/8 + c*NT/8, 16-byte column chunk
    // tid % 8, of the block's tile: one base address per operand plus
    // compile-time row strides. Q and V hold whole tiles (rows padded).
    const __bf16* qbase = Q + static_cast<uint64_t>(m0 + tid / 8) * ldq + (tid % 8) * 8;
    const __bf16* vbase = V + static_cast<uint64_t>(n0 + tid / 8) * ldv + (tid % 8) * 8;
    const uint64_t qstep = static_cast<uint64_t>(NT / 8) * ldq;
    const uint64_t vstep = static_cast<uint64_t>(NT / 8) * ldv;
    const int lds0 = (tid / 8) * kLdsRow + (tid % 8) * 8;
    constexpr int kLdsStep = (NT / 8) * kLdsRow;
    // MFMA operand fragments: lane holds row (lane % 32), K values
    // 8*(lane/32) .. +7 of each 16-wide K step.
    const int a_off = (wm + lane % 32) * kLdsRow + (lane / 32) * 8;
    const int b_off = (wn + lane % 32) * kLdsRow + (lane / 32) * 8;

    floatx16 acc[TM][TN];
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j)
#pragma unroll
            for (int v = 0; v < 16; ++v) acc[i][j][v] = 0.0F;

    // The loads of step s + 1 are issued as soon as step s sits in LDS (before
    // the barrier), so they have the barrier wait and the MFMAs of step s to
    // arrive. The last iteration re-loads the last step (unused), which keeps
    // the loads unconditional.
    const uint32_t k_last = gb.k_end - kBK;
    UHANN_GRAM_LOAD(gb.k_begin, ra, rb)
    UHANN_GRAM_STORE(0, ra, rb)
    UHANN_GRAM_LOAD(min(gb.k_begin + kBK, k_last), ra, rb)
    __syncthreads();
    int cur = 0;
    for (uint32_t k0 = gb.k_begin + kBK; k0 < gb.k_end; k0 += kBK) {
        UHANN_GRAM_COMPUTE(cur)
        UHANN_GRAM_STORE(cur ^ 1, ra, rb)
        UHANN_GRAM_LOAD(min(k0 + kBK, k_last), ra, rb)
        __syncthreads();
        cur ^= 1;
    }
    UHANN_GRAM_COMPUTE(cur)

    if (gb.part >= 0) {
        const uint64_t unit = static_cast<uint64_t>(gb.tile - sp.full) * sp.parts + gb.part;
        store_partials<NT, TM, TN>(sp.partials + unit * (BM * BN), acc, tid);
        return;
    }
    gram_epilogue<TM, TN, false>(acc, m0 + wm, n0 + wn, lane, M, N, qs, vs, out, ldo);
}

#undef UHANN_GRAM_LOAD
#undef UHANN_GRAM_STORE
#undef UHANN_GRAM_COMPUTE

// <q,v> - <q^,v^> = sum_i (q_i dv_i + v^_i dq_i) with dv = v - v^, dq = q - q^.
// Modelling the rounding signs as independent and symmetric, Hoeffding with
// c_i = |q_i dv_i| + |v^_i dq_i| and Minkowski + Cauchy-Schwarz
// (sum c_i^2 <= (||q||_4 ||dv||_4 + ||v^||_4 ||dq||_4)^2) bound the error by
// K (||q||_4 ||dv||_4 + ||v^||_4 ||dq||_4) except with probability
// 2 exp(-K^2/2); dv and dq are the rows' actual rounding errors. The MFMA
// accumulation adds acc_coef u_32 ||q^||_2 ||v^||_2 (acc_coef set by the
// caller for the instruction's accumulation behaviour).
__device__ __forceinline__ float pair_tau(const RowStats& q, const RowStats& v, float acc_coef)
{
    const float dot_bound = kHoeffdingK * (q.l4 * v.d4 + v.l4_hat * q.d4) + acc_coef * kU32 * q.l2_hat * v.l2_hat;
    return 2.0F * dot_bound + 4.0F * kU32 * (q.norm2 + v.norm2);
}

// Order-preserving map of binary32 to uint32 (negative values included).
__device__ __forceinline__ uint32_t ordered(float f)
{
    const uint32_t b = __float_as_uint(f);
    return (b & 0x80000000U) ? ~b : (b | 0x80000000U);
}
__device__ __forceinline__ float unordered(uint32_t u)
{
    return __uint_as_float((u & 0x80000000U) ? (u & 0x7fffffffU) : ~u);
}

constexpr int kSelThreads = 256;
constexpr int kSelWaves = kSelThreads / 64;
// Candidate rows re-ranked per pass over the query row (rows / kSelWaves
// per wave): 8 lets one wave take the first two candidates with one read of
// the query row (the INT8 re-rank on the benchmark batch, 2.09 candidates per
// query: 0.55 ms against 0.59 with 4).
constexpr int kRerankRows = 8;
constexpr int kKeepLo = 32;  // approx - tau values kept in registers per thread

// Binary32 distances of one query row to R reference rows in the reference's
// order, computed by ONE wave. The 256 chains (chain c adds columns c,
// c + 256, ... in order with fmaf) sit four to a lane: lane l holds chains
// 4l .. 4l + 3 and loads them as one 16-byte vector per 256-column stripe.
// The reduction pairs the chains exactly as the 256-thread block tree does:
// inside each 32-chain group offsets 16, 8, 4 (lanes l + 4, l + 2, l + 1),
// then 2 and 1 (components of one lane), then the eight group sums
// ((s0 + s4) + (s2 + s6)) + ((s1 + s5) + (s3 + s7)). Rows must be 16-byte
// aligned; the last partial stripe is read element by element (nothing past
// column dim - 1 is touched). The distances are returned in every lane.
template <int R>
__device__ __forceinline__ void wave_distances(const float* __restrict__ qrow, const float* const (&rrow)[R],
                                               uint32_t dim, float (&out)[R])
{
    const int lane = threadIdx.x % 64;
    float acc[R][4];
#pragma unroll
    for (int j = 0; j < R; ++j)
#pragma unroll
        for (int c = 0; c < 4; ++c) acc[j][c] = 0.0F;
    const uint32_t stripes = dim / 256U;
    const float4* q4 = reinterpret_cast<const float4*>(qrow) + lane;
    const float4* r4[R];
#pragma unroll
    for (int j = 0; j < R; ++j) r4[j] = reinterpret_cast<const float4*>(rrow[j]) + lane;
    auto step = [&](const float4& qv, const float4& rv, float (&a)[4]) {
        float d = rv.x - qv.x;
        a[0] = __builtin_fmaf(d, d, a[0]);
        d = rv.y - qv.y;
        a[1] = __builtin_fmaf(d, d, a[1]);
        d = rv.z - qv.z;
        a[2] = __builtin_fmaf(d, d, a[2]);
        d = rv.w - qv.w;
        a[3] = __builtin_fmaf(d, d, a[3]);
    };
    constexpr int kU = 4;  // stripes loaded together
    uint32_t s = 0;
    for (; s + kU <= stripes; s += kU) {
        float4 qv[kU], rv[kU][R];
#pragma unroll
        for (int u = 0; u < kU; ++u) {
            qv[u] = q4[(s + u) * 64U];
#pragma unroll
            for (int j = 0; j < R; ++j) rv[u][j] = r4[j][(s + u) * 64U];
        }
#pragma unroll
        for (int u = 0; u < kU; ++u)
#pragma unroll
            for (int j = 0; j < R; ++j) step(qv[u], rv[u][j], acc[j]);
    }
    for (; s < stripes; ++s) {
        const float4 qv = q4[s * 64U];
#pragma unroll
        for (int j = 0; j < R; ++j) step(qv, r4[j][s * 64U], acc[j]);
    }
    const uint32_t c0 = stripes * 256U + 4U * static_cast<uint32_t>(lane);
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        if (c0 + c < dim) {
            const float qv = qrow[c0 + c];
#pragma unroll
            for (int j = 0; j < R; ++j) {
                const float d = rrow[j][c0 + c] - qv;
                acc[j][c] = __builtin_fmaf(d, d, acc[j][c]);
            }
        }
    }
#pragma unroll
    for (int j = 0; j < R; ++j) {
        float v[4];
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            float x = acc[j][c];
            x = x + __shfl_down(x, 4);  // chain offset 16
            x = x + __shfl_down(x, 2);  // chain offset 8
            x = x + __shfl_down(x, 1);  // chain offset 4
            v[c] = x;
        }
        const float group = (v[0] + v[2]) + (v[1] + v[3]);  // offsets 2, 1; valid in lanes 8G
        float s8[8];
#pragma unroll
        for (int g = 0; g < 8; ++g) s8[g] = __shfl(group, 8 * g);
        out[j] = ((s8[0] + s8[4]) + (s8[2] + s8[6])) + ((s8[1] + s8[5]) + (s8[3] + s8[7]));
    }
}

// Binary64 squared distance of two binary32 rows computed by one wave: lane l
// adds columns 256 s + 4 l + c (16-byte rows) or l + 64 s (any rows) in order,
// then a butterfly over the wave; the value is returned in every lane. Used
// only to order candidates inside the binary32 error band.
template <bool VEC>
__device__ __forceinline__ double wave_distance_f64(const float* __restrict__ qrow, const float* __restrict__ rrow,
                                                    uint32_t dim)
{
    const uint32_t lane = threadIdx.x % 64U;
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    auto sq = [](float r, float q) {
        const double d = static_cast<double>(r) - static_cast<double>(q);
        return d * d;
    };
    if constexpr (VEC) {
        const uint32_t stripes = dim / 256U;
        const float4* q4 = reinterpret_cast<const float4*>(qrow) + lane;
        const float4* r4 = reinterpret_cast<const float4*>(rrow) + lane;
        constexpr int kU = 4;
        uint32_t s = 0;
        for (; s + kU <= stripes; s += kU) {
            float4 qv[kU], rv[kU];
#pragma unroll
            for (int u = 0; u < kU; ++u) {
                qv[u] = q4[(s + u) * 64U];
                rv[u] = r4[(s + u) * 64U];
            }
#pragma unroll
            for (int u = 0; u < kU; ++u) {
                acc[0] += sq(rv[u].x, qv[u].x);
                acc[1] += sq(rv[u].y, qv[u].y);
                acc[2] += sq(rv[u].z, qv[u].z);
                acc[3] += sq(rv[u].w, qv[u].w);
            }
        }
        for (; s < stripes; ++s) {
            const float4 qv = q4[s * 64U];
            const float4 rv = r4[s * 64U];
            acc[0] += sq(rv.x, qv.x);
            acc[1] += sq(rv.y, qv.y);
            acc[2] += sq(rv.z, qv.z);
            acc[3] += sq(rv.w, qv.w);
        }
        const uint32_t c0 = stripes * 256U + 4U * lane;
#pragma unroll
        for (int c = 0; c < 4; ++c)
            if (c0 + c < dim) acc[c] += sq(rrow[c0 + c], qrow[c0 + c]);
    } else {
        uint32_t c = lane;
        for (; c + 3U * 64U < dim; c += 4U * 64U) {
            float qv[4], rv[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                qv[u] = qrow[c + u * 64U];
                rv[u] = rrow[c + u * 64U];
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) acc[u] += sq(rv[u], qv[u]);
        }
        for (; c < dim; c += 64U) acc[0] += sq(rrow[c], qrow[c]);
    }
    double v = (acc[0] + acc[1]) + (acc[2] + acc[3]);
    for (int off = 32; off >= 1; off /= 2) v += __shfl_xor(v, off);
    return v;
}

// One block per query. See the file comment for the algorithm.
// Exact re-rank of `count` candidate rows s_cand of query q: binary32
// distances in the reference's order into s_dist (16-byte aligned rows: each
// wave takes ROWS / 4 candidates per pass over the query row, four chains per
// lane; otherwise thread t is chain t, ROWS candidates per pass), the lowest
// (distance, row), then the binary64 minimum among the candidates inside the
// binary32 error band (one candidate per wave). Uses the whole block.
template <int ROWS>
__device__ void rerank_pick(const SelectArgs& a, uint32_t q, const uint32_t* s_cand, uint32_t count, float* s_dist,
                            float (*s_group)[ROWS], double* s_d64, uint32_t* s_best, float* s_best_d,
                            uint32_t& best, float& best_d)
{
    constexpr int kRerankRows = ROWS;  // candidate rows re-ranked per pass over the query row
    const int t = threadIdx.x;
    const float* qrow = a.queries + static_cast<uint64_t>(q) * a.ld_q;
    // 16-byte aligned rows: each wave re-ranks its own candidates (ROWS /
    // kSelWaves per pass over the query row), four chains per lane.
    const bool vec = a.ld_q % 4U == 0 && a.ld_ref % 4U == 0 && (reinterpret_cast<uintptr_t>(a.queries) & 15U) == 0 &&
                     (reinterpret_cast<uintptr_t>(a.refs) & 15U) == 0;
    constexpr int kWaveRows = ROWS / kSelWaves > 0 ? ROWS / kSelWaves : 1;
    if (vec && count <= static_cast<uint32_t>(kSelWaves)) {
        // few candidates (small batches have one block per query and CU):
        // one per wave, so they stream in parallel
        const uint32_t wave = static_cast<uint32_t>(t) / 64U;
        if (wave < count) {
            const float* rrow[1] = {a.refs + static_cast<uint64_t>(s_cand[wave]) * a.ld_ref};
            float d[1];
            wave_distances<1>(qrow, rrow, a.dim, d);
            if (t % 64 == 0) s_dist[wave] = d[0];
        }
        __syncthreads();
    } else if (vec) {
        const uint32_t wave = static_cast<uint32_t>(t) / 64U;
        for (uint32_t g = wave * kWaveRows; g < count; g += kSelWaves * kWaveRows) {
            const float* rrow[kWaveRows];
#pragma unroll
            for (int j = 0; j < kWaveRows; ++j) {
                const uint32_t r = s_cand[min(g + static_cast<uint32_t>(j), count - 1U)];
                rrow[j] = a.refs + static_cast<uint64_t>(r) * a.ld_ref;
            }
            float d[kWaveRows];
            wave_distances<kWaveRows>(qrow, rrow, a.dim, d);
            if (t % 64 == 0) {
#pragma unroll
                for (int j = 0; j < kWaveRows; ++j)
                    if (g + j < count) s_dist[g + j] = d[j];
            }
        }
        __syncthreads();
    }
    // Otherwise the exact binary32 re-rank in the reference's order with
    // thread t as chain t.
    for (uint32_t g = 0; !vec && g < count; g += kRerankRows) {
        const uint32_t ng = min(static_cast<uint32_t>(kRerankRows), count - g);
        const float* rrow[kRerankRows];
#pragma unroll
        for (int j = 0; j < kRerankRows; ++j) {
            const uint32_t r = s_cand[g + min(static_cast<uint32_t>(j), ng - 1U)];
            rrow[j] = a.refs + static_cast<uint64_t>(r) * a.ld_ref;
        }
        float acc[kRerankRows];
#pragma unroll
        for (int j = 0; j < kRerankRows; ++j) acc[j] = 0.0F;
        // Chain t adds columns t, t + 256, ... in order; four steps load
        // together so more loads are in flight per thread.
        uint32_t c = t;
        for (; c + 3U * kSelThreads < a.dim; c += 4U * kSelThreads) {
            float qv[4], rv[4][kRerankRows];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                qv[u] = qrow[c + u * kSelThreads];
#pragma unroll
                for (int j = 0; j < kRerankRows; ++j) rv[u][j] = rrow[j][c + u * kSelThreads];
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) {
#pragma unroll
                for (int j = 0; j < kRerankRows; ++j) {
                    const float d = rv[u][j] - qv[u];
                    acc[j] = __builtin_fmaf(d, d, acc[j]);
                }
            }
        }
        for (; c < a.dim; c += kSelThreads) {
            const float qv = qrow[c];
#pragma unroll
            for (int j = 0; j < kRerankRows; ++j) {
                const float d = rrow[j][c] - qv;
                acc[j] = __builtin_fmaf(d, d, acc[j]);
            }
        }
#pragma unroll
        for (int j = 0; j < kRerankRows; ++j) {
            float v = acc[j];
            for (int off = 16; off >= 1; off /= 2) v = v + __shfl_down(v, off);
            if (t % 32 == 0) s_group[t / 32][j] = v;
        }
        __syncthreads();
        if (t < static_cast<int>(ng)) {
            const float g0 = s_group[0][t], g1 = s_group[1][t], g2 = s_group[2][t], g3 = s_group[3][t];
            const float g4 = s_group[4][t], g5 = s_group[5][t], g6 = s_group[6][t], g7 = s_group[7][t];
            s_dist[g + t] = ((g0 + g4) + (g2 + g6)) + ((g1 + g5) + (g3 + g7));
        }
        __syncthreads();
    }

    if (t == 0) {
        uint32_t b = 0xffffffffU;
        float b_d = __uint_as_float(0x7f800000U);
        for (uint32_t i = 0; i < count; ++i) {
            const float d = s_dist[i];
            const uint32_t r = s_cand[i];
            if (d < b_d || (d == b_d && r < b)) {
                b_d = d;
                b = r;
            }
        }
        *s_best = b;
        *s_best_d = b_d;
    }
    __syncthreads();
    best = *s_best;
    best_d = *s_best_d;

    // Binary64 re-rank of candidates inside the binary32 error band: the
    // waves take the band's candidates in turn, then the binary64 minimum
    // (lowest row on ties) wins.
    const double band = static_cast<double>(best_d) * a.fp32_band;
    uint32_t near = 0;
    for (uint32_t i = 0; i < count; ++i) near += (static_cast<double>(s_dist[i]) <= band) ? 1U : 0U;
    if (near >= 2) {
        const uint32_t wave = static_cast<uint32_t>(t) / 64U;
        uint32_t k = 0;
        for (uint32_t i = 0; i < count; ++i) {
            if (static_cast<double>(s_dist[i]) > band) continue;
            if (k++ % kSelWaves != wave) continue;
            const float* rr = a.refs + static_cast<uint64_t>(s_cand[i]) * a.ld_ref;
            const double d64 = vec ? wave_distance_f64<true>(qrow, rr, a.dim) : wave_distance_f64<false>(qrow, rr, a.dim);
            if (t % 64 == 0) s_d64[i] = d64;
        }
        __syncthreads();
        double best64 = __longlong_as_double(0x7ff0000000000000LL);
        uint32_t best64_r = 0xffffffffU;
        float best64_f = 0.0F;
        for (uint32_t i = 0; i < count; ++i) {
            if (static_cast<double>(s_dist[i]) > band) continue;
            const uint32_t r = s_cand[i];
            const double d64 = s_d64[i];
            if (d64 < best64 || (d64 == best64 && r < best64_r)) {
                best64 = d64;
                best64_r = r;
                best64_f = s_dist[i];
            }
        }
        best = best64_r;
        best_d = best64_f;
        if (t == 0 && a.stats != nullptr) atomicAdd(&a.stats->fp64_reranked, 1U);
    }
}

__global__ void __launch_bounds__(kSelThreads)
select_rerank_kernel(SelectArgs a)
{
    __shared__ uint32_t s_min;
    __shared__ uint32_t s_count;
    __shared__ uint32_t s_cand[kMaxCandidates];
    __shared__ float s_dist[kMaxCandidates];
    __shared__ float s_group[8][kRerankRows];
    __shared__ double s_d64[kMaxCandidates];
    __shared__ uint32_t s_best;
    __shared__ float s_best_d;

    const uint32_t q = blockIdx.x;
    if (q >= a.nq) return;
    const int t = threadIdx.x;
    const RowStats qs = a.qstats[q];
    const float* arow = a.approx + static_cast<uint64_t>(q) * a.ld_approx;

    if (t == 0) {
        s_min = 0xffffffffU;
        s_count = 0;
    }
    __syncthreads();
    // U = min_r (approx + tau). The lower ends approx - tau of the first
    // kKeepLo * kSelThreads references stay in registers for the second pass.
    uint32_t local = 0xffffffffU;
    float lo_keep[kKeepLo];
#pragma unroll
    for (int i = 0; i < kKeepLo; ++i) {
        const uint32_t r = t + static_cast<uint32_t>(i) * kSelThreads;
        lo_keep[i] = __builtin_inff();
        if (r < a.rows) {
            const float e = arow[r];
            const float tau = pair_tau(qs, a.vstats[r], a.acc_coef);
            local = min(local, ordered(e + tau));
            lo_keep[i] = e - tau;
        }
    }
    for (uint32_t r = t + kKeepLo * kSelThreads; r < a.rows; r += kSelThreads) {
        const float up = arow[r] + pair_tau(qs, a.vstats[r], a.acc_coef);
        local = min(local, ordered(up));
    }
    for (int off = 32; off >= 1; off /= 2) local = min(local, __shfl_down(local, off));
    if (t % 64 == 0) atomicMin(&s_min, local);
    __syncthreads();
    const float upper = unordered(s_min);
#pragma unroll
    for (int i = 0; i < kKeepLo; ++i) {
        if (lo_keep[i] <= upper) {
            const uint32_t slot = atomicAdd(&s_count, 1U);
            if (slot < kMaxCandidates) s_cand[slot] = t + static_cast<uint32_t>(i) * kSelThreads;
        }
    }
    for (uint32_t r = t + kKeepLo * kSelThreads; r < a.rows; r += kSelThreads) {
        const float lo = arow[r] - pair_tau(qs, a.vstats[r], a.acc_coef);
        if (lo <= upper) {
            const uint32_t slot = atomicAdd(&s_count, 1U);
            if (slot < kMaxCandidates) s_cand[slot] = r;
        }
    }
    __syncthreads();
    const uint32_t count = s_count;
    // No candidate only happens with non-finite estimates (overflowing
    // norms); both cases fall back to the exact scan.
    if (count > kMaxCandidates || count == 0) {
        if (t == 0) {
            const uint32_t slot = atomicAdd(a.escalate_count, 1U);
            a.escalate_list[slot] = q;
            a.out_idx[q] = 0;
            a.out_dist[q] = 0.0F;
            if (a.stats != nullptr) atomicAdd(&a.stats->escalated, 1U);
        }
        return;
    }

    uint32_t best;
    float best_d;
    rerank_pick<kRerankRows>(a, q, s_cand, count, s_dist, s_group, s_d64, &s_best, &s_best_d, best, best_d);
    if (t == 0) {
        a.out_idx[q] = best;
        a.out_dist[q] = best_d;
        if (a.stats != nullptr) atomicAdd(&a.stats->candidates, count);
    }
}


// ---------------------------------------------------------------------------
// Per-row-scaled FP8 (e4m3) Gram + top-k shortlist re-rank.

constexpr int kBK8 = 128;             // e4m3 values (bytes) per K step
constexpr int kLdsRow8 = kBK8 + 16;   // LDS row stride in bytes (padding)

// e4m3 (OCP: bias 7, max 448, no infinities) of |v| <= 448, round to
// nearest even. Normal results keep 3 mantissa bits; below 2^-6 the result
// is a multiple of 2^-9 (m = 8 rounds up to the smallest normal, 0x08).
__device__ __forceinline__ uint32_t e4m3_rne(float v)
{
    uint32_t u = __float_as_uint(v);
    const uint32_t sign = (u >> 24) & 0x80U;
    u &= 0x7fffffffU;
    if (u >= 0x7f800000U) return sign | 0x7fU;  // NaN stays NaN
    if (u >= (121U << 23)) {                       // |v| >= 2^-6
        const uint32_t r = u + 0x7ffffU + ((u >> 20) & 1U);
        return sign | ((((r >> 23) - 120U) & 0xfU) << 3) | ((r >> 20) & 7U);
    }
    const uint32_t m = static_cast<uint32_t>(rintf(__uint_as_float(u) * 512.0F));
    return sign | m;
}

__device__ __forceinline__ float e4m3_value(uint32_t b)
{
    const uint32_t e = (b >> 3) & 0xfU;
    const uint32_t m = b & 7U;
    const float mag = e == 0 ? static_cast<float>(m) * 0.001953125F  // m * 2^-9
                             : __uint_as_float(((e + 120U) << 23) | (m << 20));  // (1 + m/8) 2^(e-7)
    return (b & 0x80U) ? -mag : mag;
}

__device__ __forceinline__ float block_max_f32(float v, float* scratch)
{
    for (int off = 32; off >= 1; off /= 2) v = fmaxf(v, __shfl_down(v, off));
    const int lane = threadIdx.x % 64;
    const int wave = threadIdx.x / 64;
    __syncthreads();
    if (lane == 0) scratch[wave] = v;
    __syncthreads();
    float m = 0.0F;
    for (int w = 0; w < static_cast<int>(blockDim.x / 64); ++w) m = fmaxf(m, scratch[w]);
    return m;
}

// One block of 256 threads per row: max |x| -> scale exponent -> bytes.
__global__ void __launch_bounds__(256)
to_fp8_rows_kernel(const float* __restrict__ src, uint64_t ld_src, uint32_t rows, uint32_t dim,
                   uint8_t* __restrict__ dst, uint64_t ld_dst, RowStats* __restrict__ stats)
{
    __shared__ double scratch[8];
    __shared__ float fscratch[8];
    const uint32_t row = blockIdx.x;
    if (row >= rows) return;
    const float* s = src + static_cast<uint64_t>(row) * ld_src;
    uint8_t* o = dst + static_cast<uint64_t>(row) * ld_dst;
    // 16-byte aligned rows (ld_src % 4 == 0) are read four values at a time.
    const bool vec = ld_src % 4 == 0 && (reinterpret_cast<uintptr_t>(src) & 15U) == 0;
    const uint32_t full = vec ? dim / 4 * 4 : 0;
    float mx = 0.0F;
    {
        auto qmax = [](const float4& x) { return fmaxf(fmaxf(fabsf(x.x), fabsf(x.y)), fmaxf(fabsf(x.z), fabsf(x.w))); };
        const uint32_t step = blockDim.x * 4;
        uint32_t c = threadIdx.x * 4;
        for (; c + 3 * step < full; c += 4 * step) {
            float4 x[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) x[u] = *reinterpret_cast<const float4*>(s + c + u * step);
            mx = fmaxf(mx, fmaxf(fmaxf(qmax(x[0]), qmax(x[1])), fmaxf(qmax(x[2]), qmax(x[3]))));
        }
        for (; c < full; c += step) mx = fmaxf(mx, qmax(*reinterpret_cast<const float4*>(s + c)));
    }
    for (uint32_t c = full + threadIdx.x; c < dim; c += blockDim.x) mx = fmaxf(mx, fabsf(s[c]));
    mx = block_max_f32(mx, fscratch);
    // Smallest c with 448 * 2^c >= mx: mx = f 2^e, f in [0.5, 1), 448 = 0.875 2^9.
    int c_exp = 0;
    if (mx > 0.0F) {
        int e = 0;
        const float f = frexpf(mx, &e);
        c_exp = f <= 0.875F ? e - 9 : e - 8;
        c_exp = c_exp < -126 ? -126 : c_exp;
    }
    const float scale = ldexpf(1.0F, c_exp);
    // 2^-c_exp is a normal binary32 (c_exp in [-126, 120]), so x * inv is
    // the correctly rounded x 2^-c_exp, as ldexpf gives.
    const float inv = ldexpf(1.0F, -c_exp);
    double n2 = 0.0, n4 = 0.0, h2 = 0.0, h4 = 0.0, e4 = 0.0;
    auto one = [&](float x) {
        const uint32_t b = e4m3_rne(x * inv);
        const double xd = x;
        const double hd = static_cast<double>(e4m3_value(b)) * static_cast<double>(scale);
        const double ed = xd - hd;
        n2 += xd * xd;
        n4 += (xd * xd) * (xd * xd);
        h2 += hd * hd;
        h4 += (hd * hd) * (hd * hd);
        e4 += (ed * ed) * (ed * ed);
        return b;
    };
    // ld_dst is a multiple of 128, so four-byte groups never straddle a row.
    const uint32_t step = blockDim.x * 4;
    auto quad = [&](const float4& x, uint32_t c) {
        const uint32_t packed = one(x.x) | (one(x.y) << 8) | (one(x.z) << 16) | (one(x.w) << 24);
        *reinterpret_cast<uint32_t*>(o + c) = packed;
    };
    uint32_t c4 = threadIdx.x * 4;
    for (; c4 + 3 * step < full; c4 += 4 * step) {
        float4 x[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) x[u] = *reinterpret_cast<const float4*>(s + c4 + u * step);
#pragma unroll
        for (int u = 0; u < 4; ++u) quad(x[u], c4 + u * step);
    }
    for (; c4 < full; c4 += step) quad(*reinterpret_cast<const float4*>(s + c4), c4);
    for (uint32_t c = full + threadIdx.x; c < ld_dst; c += blockDim.x) {
        o[c] = static_cast<uint8_t>(one(c < dim ? s[c] : 0.0F));
    }
    n2 = block_sum_f64(n2, scratch);
    n4 = block_sum_f64(n4, scratch);
    h2 = block_sum_f64(h2, scratch);
    h4 = block_sum_f64(h4, scratch);
    e4 = block_sum_f64(e4, scratch);
    if (threadIdx.x == 0) {
        RowStats st{};
        st.norm2 = static_cast<float>(n2);
        st.l4 = static_cast<float>(sqrt(sqrt(n4))) * 1.0001F;
        st.l4_hat = static_cast<float>(sqrt(sqrt(h4))) * 1.0001F;
        st.l2_hat = static_cast<float>(sqrt(h2)) * 1.0001F;
        st.d4 = static_cast<float>(sqrt(sqrt(e4))) * 1.0001F;
        st.scale = scale;
        stats[row] = st;
    }
}

typedef int intx8 __attribute__((ext_vector_type(8)));

// 256 x 256 outputs per block, 4 x 4 waves of 2 x 2 32x32 tiles, as the
// BF16 kernel; a K step is 128 e4m3 bytes per row (two MFMA K=64 steps).
// Operand fragments: lane holds row (lane % 32) and the 32 K values
// 32*(lane/32) .. +31 of each 64-wide step (the same assignment for both
// operands, which is what the MFMA's dot product needs).
__global__ void __launch_bounds__(1024)
gram_fp8_kernel(const uint8_t* __restrict__ Q, uint64_t ldq, uint32_t M, const uint8_t* __restrict__ V,
                uint64_t ldv, uint32_t N, uint32_t K, const RowStats* __restrict__ qs,
                const RowStats* __restrict__ vs, float* __restrict__ out, uint64_t ldo, uint32_t tiles_m,
                uint32_t tiles_n, uint32_t group_m, GramSplit sp)
{
    constexpr int WN = 4, NT = 1024, BM = 256, BN = 256;
    constexpr int CA = BM * 8 / NT;  // 16-byte chunks per thread per K step
    constexpr int CB = BN * 8 / NT;
    __shared__ __attribute__((aligned(16))) uint8_t As[2 * BM * kLdsRow8];
    __shared__ __attribute__((aligned(16))) uint8_t Bs[2 * BN * kLdsRow8];

    const GramBlock gb = gram_block(blockIdx.x, sp, K, kBK8);
    const uint2 o = tile_origin(gb.tile, tiles_m, tiles_n, group_m);
    const uint32_t m0 = o.x * BM;
    const uint32_t n0 = o.y * BN;

    const int tid = threadIdx.x;
    const int lane = tid % 64;
    const int wave = tid / 64;
    const int wm = (wave / WN) * 64;
    const int wn = (wave % WN) * 64;

    u32x4 ra[CA], rb[CB];
    const uint8_t* qbase = Q + static_cast<uint64_t>(m0 + tid / 8) * ldq + (tid % 8) * 16;
    const uint8_t* vbase = V + static_cast<uint64_t>(n0 + tid / 8) * ldv + (tid % 8) * 16;
    const uint64_t qstep = static_cast<uint64_t>(NT / 8) * ldq;
    const uint64_t vstep = static_cast<uint64_t>(NT / 8) * ldv;
    const int lds0 = (tid / 8) * kLdsRow8 + (tid % 8) * 16;
    constexpr int kLdsStep = (NT / 8) * kLdsRow8;
    const int a_off = (wm + lane % 32) * kLdsRow8 + (lane / 32) * 32;
    const int b_off = (wn + lane % 32) * kLdsRow8 + (lane / 32) * 32;

    floatx16 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int v = 0; v < 16; ++v) acc[i][j][v] = 0.0F;

    auto frag = [](const uint8_t* p) {
        const u32x4 lo = *reinterpret_cast<const u32x4*>(p);
        const u32x4 hi = *reinterpret_cast<const u32x4*>(p + 16);
        return __builtin_bit_cast(intx8, __builtin_shufflevector(lo, hi, 0, 1, 2, 3, 4, 5, 6, 7));
    };
#define UHANN_FP8_LOAD(k0)                                                                    \
    _Pragma("unroll") for (int c = 0; c < CA; ++c) ra[c] =                                    \
        *reinterpret_cast<const u32x4*>(qbase + c * qstep + (k0));                            \
    _Pragma("unroll") for (int c = 0; c < CB; ++c) rb[c] =                                    \
        *reinterpret_cast<const u32x4*>(vbase + c * vstep + (k0));                            \
    __builtin_amdgcn_sched_barrier(0);
#define UHANN_FP8_STORE(buf)                                                                  \
    _Pragma("unroll") for (int c = 0; c < CA; ++c)                                            \
        *reinterpret_cast<u32x4*>(As + (buf) * BM * kLdsRow8 + lds0 + c * kLdsStep) = ra[c];  \
    _Pragma("unroll") for (int c = 0; c < CB; ++c)                                            \
        *reinterpret_cast<u32x4*>(Bs + (buf) * BN * kLdsRow8 + lds0 + c * kLdsStep) = rb[c];
#define UHANN_FP8_COMPUTE(buf)                                                                \
    _Pragma("unroll") for (int kk = 0; kk < kBK8; kk += 64)                                   \
    {                                                                                         \
        const uint8_t* ab = As + (buf) * BM * kLdsRow8 + a_off + kk;                          \
        const uint8_t* bb = Bs + (buf) * BN * kLdsRow8 + b_off + kk;                          \
        const intx8 a0 = frag(ab);                                                            \
        const intx8 a1 = frag(ab + 32 * kLdsRow8);                                            \
        const intx8 b0 = frag(bb);                                                            \
        const intx8 b1 = frag(bb + 32 * kLdsRow8);                                            \
        acc[0][0] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(a0, b0, acc[0][0], 0, 0, 0, 127, 0, 127); \
        acc[0][1] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(a0, b1, acc[0][1], 0, 0, 0, 127, 0, 127); \
        acc[1][0] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(a1, b0, acc[1][0], 0, 0, 0, 127, 0, 127); \
        acc[1][1] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(a1, b1, acc[1][1], 0, 0, 0, 127, 0, 127); \
    }

    // Loads of the next step issue before the barrier, as in the BF16 kernel.
    const uint32_t k_last = gb.k_end - kBK8;
    UHANN_FP8_LOAD(gb.k_begin)
    UHANN_FP8_STORE(0)
    UHANN_FP8_LOAD(min(gb.k_begin + kBK8, k_last))
    __syncthreads();
    int cur = 0;
    for (uint32_t k0 = gb.k_begin + kBK8; k0 < gb.k_end; k0 += kBK8) {
        UHANN_FP8_COMPUTE(cur)
        UHANN_FP8_STORE(cur ^ 1)
        UHANN_FP8_LOAD(min(k0 + kBK8, k_last))
        __syncthreads();
        cur ^= 1;
    }
    UHANN_FP8_COMPUTE(cur)
#undef UHANN_FP8_LOAD
#undef UHANN_FP8_STORE
#undef UHANN_FP8_COMPUTE

    if (gb.part >= 0) {
        const uint64_t unit = static_cast<uint64_t>(gb.tile - sp.full) * sp.parts + gb.part;
        store_partials<NT, 2, 2>(sp.partials + unit * (BM * BN), acc, tid);
        return;
    }
    gram_epilogue<2, 2, true>(acc, m0 + wm, n0 + wn, lane, M, N, qs, vs, out, ldo);
}

// Top-k selection per query, k = a.shortlist <= kShortlist (64-bit keys
// (ordered estimate << 32 | row): per-thread sorted lists of kShortlist,
// then an 8-pass radix select of the k-th smallest key over their union),
// exact re-rank, and the certification pass.
__global__ void __launch_bounds__(kSelThreads)
select_topk_rerank_kernel(SelectArgs a)
{
    __shared__ unsigned int s_hist[256];
    __shared__ uint32_t s_wsum[kSelWaves];
    __shared__ unsigned long long s_prefix;
    __shared__ uint32_t s_need;
    __shared__ uint32_t s_count;
    __shared__ uint32_t s_cand[kShortlist];
    __shared__ float s_dist[kShortlist];
    __shared__ float s_group[8][kShortlist];
    __shared__ double s_d64[kShortlist];
    __shared__ uint32_t s_best;
    __shared__ float s_best_d;
    __shared__ int s_uncertified;

    const uint32_t q = blockIdx.x;
    if (q >= a.nq) return;
    const int t = threadIdx.x;
    const float* arow = a.approx + static_cast<uint64_t>(q) * a.ld_approx;

    unsigned long long top[kShortlist];
#pragma unroll
    for (int i = 0; i < kShortlist; ++i) top[i] = ~0ULL;
    for (uint32_t r = t; r < a.rows; r += kSelThreads) {
        const unsigned long long key = (static_cast<unsigned long long>(ordered(arow[r])) << 32) | r;
        if (key < top[kShortlist - 1]) {
            // insert into the ascending list (the last entry drops out)
            unsigned long long cur = key;
#pragma unroll
            for (int i = 0; i < kShortlist; ++i) {
                const unsigned long long lo = cur < top[i] ? cur : top[i];
                const unsigned long long hi = cur < top[i] ? top[i] : cur;
                top[i] = lo;
                cur = hi;
            }
        }
    }
    const uint32_t need0 = min(min(a.shortlist, static_cast<uint32_t>(kShortlist)), a.rows);
    if (t == 0) {
        s_prefix = 0ULL;
        s_need = need0;
        s_count = 0;
        s_uncertified = 0;
    }
    // Radix select, 8 bits per pass from the top: the bin holding the
    // need-th smallest remaining key is found by a block-wide prefix sum of
    // the histogram (thread t holds bin t).
    const int lane = t % 64;
    const int wave = t / 64;
    unsigned long long mask = 0ULL;
    for (int pass = 0; pass < 8; ++pass) {
        const int shift = 56 - 8 * pass;
        s_hist[t] = 0U;
        __syncthreads();
        const unsigned long long prefix = s_prefix;
        const uint32_t need = s_need;
#pragma unroll
        for (int i = 0; i < kShortlist; ++i) {
            if (top[i] != ~0ULL && (top[i] & mask) == prefix) {
                atomicAdd(&s_hist[(top[i] >> shift) & 255ULL], 1U);
            }
        }
        __syncthreads();
        const uint32_t h = s_hist[t];
        uint32_t incl = h;
        for (int off = 1; off < 64; off *= 2) {
            const uint32_t v = __shfl_up(incl, off);
            if (lane >= off) incl += v;
        }
        if (lane == 63) s_wsum[wave] = incl;
        __syncthreads();
        for (int w = 0; w < wave; ++w) incl += s_wsum[w];
        const uint32_t excl = incl - h;
        if (excl < need && need <= incl) {  // exactly one bin when 0 < need <= total
            s_need = need - excl;
            s_prefix = prefix | (static_cast<unsigned long long>(t) << shift);
        }
        mask |= 255ULL << shift;
        __syncthreads();
    }
    const unsigned long long threshold = s_prefix;  // the need0-th smallest key
#pragma unroll
    for (int i = 0; i < kShortlist; ++i) {
        if (top[i] <= threshold) {
            const uint32_t slot = atomicAdd(&s_count, 1U);
            if (slot < kShortlist) s_cand[slot] = static_cast<uint32_t>(top[i] & 0xffffffffULL);
        }
    }
    __syncthreads();
    const uint32_t count = min(s_count, static_cast<uint32_t>(kShortlist));

    uint32_t best;
    float best_d;
    rerank_pick<kShortlist>(a, q, s_cand, count, s_dist, s_group, s_d64, &s_best, &s_best_d, best, best_d);

    // Certified when no row outside the shortlist can be closer: its
    // estimate minus its error bound exceeds the answer's distance (widened
    // by the binary32 band).
    const RowStats qs = a.qstats[q];
    const float upper = static_cast<float>(static_cast<double>(best_d) * a.fp32_band);
    int bad = 0;
    for (uint32_t r = t; r < a.rows; r += kSelThreads) {
        const float e = arow[r];
        const unsigned long long key = (static_cast<unsigned long long>(ordered(e)) << 32) | r;
        if (key > threshold && !(e - pair_tau(qs, a.vstats[r], a.acc_coef) > upper)) bad = 1;
    }
    if (bad) atomicOr(&s_uncertified, 1);
    __syncthreads();
    if (t == 0) {
        a.out_idx[q] = best;
        a.out_dist[q] = best_d;
        if (a.stats != nullptr) {
            atomicAdd(&a.stats->candidates, count);
            if (s_uncertified == 0) atomicAdd(&a.stats->certified, 1U);
        }
    }
}


// ---------------------------------------------------------------------------
// GRAM_INT8: rows as 16-bit fixed point in two int8 planes; inner products
// summed exactly in int32 on v_mfma_i32_32x32x32_i8.
//
// x^ = s (a + b / 256), so <q^, v^> = s_q s_v (H + X / 256 + L / 65536) with
// H = <a_q, a_v>, X = <a_q, b_v> + <b_q, a_v> and L = <b_q, b_v>. H and X are
// integer sums (|H| <= 127^2 dim, |X| <= 2 127 128 dim < 2^31 for dim <=
// 65536), so the MFMAs add them without rounding; L is left out and bounded
// by |s_q s_v L / 65536| <= lo_q lo_v (Cauchy-Schwarz). With q = q^ + e_q,
// v = v^ + e_v: <q, v> - <q^, v^> = <q, e_v> + <e_q, v^>, bounded by
// ||q|| ||e_v|| + ||e_q|| ||v^||. Every norm is of the row's actual values
// (binary64 sums, rounded up), so the bound holds for every input: there is
// no model of the rounding errors.

typedef int intx4 __attribute__((ext_vector_type(4)));
typedef int intx16 __attribute__((ext_vector_type(16)));

constexpr int kBKI8 = 128;               // bytes per row and K step: 64 columns of a, then of b
constexpr int kLdsRowI8 = kBKI8 + 16;    // LDS row stride in bytes (padding)
constexpr uint32_t kI8MaxDim = 65536;    // 2 * 127 * 128 * 65536 < 2^31
constexpr int kI8Code = 32512;           // |n| <= 127 * 256

// A float no smaller than v >= 0 (binary32 rounding moves a value by at most
// 2^-24 of itself, less than the 2^-20 added first).
__device__ __forceinline__ float round_up_f(double v) { return static_cast<float>(v * (1.0 + 0x1p-20)); }

// One block of 1024 threads per row, the row held in registers: thread t
// holds the four-column groups t + 1024 i (i < NPT, 4096 NPT >= the padded
// row), so the row is read once for max |x| and then turned into codes,
// bytes and binary64 sums. Groups past the row's end are zero (padding).
template <int NPT>
__global__ void __launch_bounds__(1024)
to_i8x2_rows_kernel(const float* __restrict__ src, uint64_t ld_src, uint32_t rows, uint32_t dim,
                    uint8_t* __restrict__ dst, uint64_t ld_dst, I8Stats* __restrict__ stats)
{
    __shared__ double4 scratch4[16];
    __shared__ float fscratch[16];
    const uint32_t row = blockIdx.x;
    if (row >= rows) return;
    const float* s = src + static_cast<uint64_t>(row) * ld_src;
    uint8_t* o = dst + static_cast<uint64_t>(row) * ld_dst;
    const bool vec = ld_src % 4 == 0 && (reinterpret_cast<uintptr_t>(src) & 15U) == 0;
    const uint32_t pdim = static_cast<uint32_t>(ld_dst / 2);
    float x[NPT][4];
    float mx = 0.0F;
    int bad = 0;
#pragma unroll
    for (int i = 0; i < NPT; ++i) {
        const uint32_t c = (threadIdx.x + 1024U * i) * 4U;
        if (vec && c + 3 < dim) {
            const float4 v = *reinterpret_cast<const float4*>(s + c);
            x[i][0] = v.x;
            x[i][1] = v.y;
            x[i][2] = v.z;
            x[i][3] = v.w;
        } else {
#pragma unroll
            for (int u = 0; u < 4; ++u) x[i][u] = c + u < dim ? s[c + u] : 0.0F;
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            mx = fmaxf(mx, fabsf(x[i][u]));
            bad |= !isfinite(x[i][u]);
        }
    }
    mx = block_max_f32(mx, fscratch);
    const bool finite = __syncthreads_or(bad) == 0;
    const float scale = !finite ? __builtin_nanf("") : (mx > 0.0F ? mx / 127.0F : 1.0F);
    const float inv = (finite && mx > 0.0F) ? static_cast<float>(kI8Code) / mx : 0.0F;
    const double sd = static_cast<double>(scale) * 0.00390625;  // s / 256, exact
    if (!finite) {
        // all codes zero (the statistics below are NaN)
#pragma unroll
        for (int i = 0; i < NPT; ++i)
#pragma unroll
            for (int u = 0; u < 4; ++u) x[i][u] = 0.0F;
    }
    // n2 = sum x^2 and e2 = sum (x - x^)^2 in binary64 (fused multiply-adds,
    // one rounding per term); the code sums sum n^2 (< 2^53, so exact in
    // binary64) and sum b^2 (< 2^31 per thread) are integers, and
    // ||x^||^2 = (s/256)^2 sum n^2 is formed once at the end.
    double n2 = 0.0, e2 = 0.0, c2 = 0.0;
    uint32_t b2 = 0;
    // Four a bytes and four b bytes per group, written as two 32-bit words.
#pragma unroll
    for (int i = 0; i < NPT; ++i) {
        const uint32_t c = (threadIdx.x + 1024U * i) * 4U;
        if (c >= pdim) break;
        // m = n + 32896 lies in [384, 65408]: its byte 1 is a + 128 and its
        // byte 0 is b + 128 (a = floor((n + 128) / 256) in [-127, 127],
        // b = n - 256 a in [-128, 127]); flipping bit 7 turns each into the
        // two's-complement byte.
        uint32_t m[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            int n = static_cast<int>(rintf(x[i][u] * inv));
            n = n < -kI8Code ? -kI8Code : (n > kI8Code ? kI8Code : n);
            m[u] = static_cast<uint32_t>(n + 32896);
            const int b = static_cast<int>(m[u] & 255U) - 128;
            const double xd = x[i][u];
            const double nd = static_cast<double>(n);
            const double ed = fma(-sd, nd, xd);  // x - s n / 256, the product exact (24 x 16 bits)
            n2 = fma(xd, xd, n2);
            e2 = fma(ed, ed, e2);
            c2 = fma(nd, nd, c2);  // integers below 2^53: exact
            b2 += static_cast<uint32_t>(b * b);
        }
        // bytes 0..3 of t01 are m0.b0, m0.b1, m1.b0, m1.b1; of t23 likewise
        // for m2, m3; the selector picks bytes of {t23, t01} (t01 low)
        const uint32_t t01 = m[0] | (m[1] << 16);
        const uint32_t t23 = m[2] | (m[3] << 16);
        const uint32_t pa = __builtin_amdgcn_perm(t23, t01, 0x07050301U) ^ 0x80808080U;
        const uint32_t pb = __builtin_amdgcn_perm(t23, t01, 0x06040200U) ^ 0x80808080U;
        uint8_t* chunk = o + (c / 64U) * 128U + (c % 64U);
        *reinterpret_cast<uint32_t*>(chunk) = pa;
        *reinterpret_cast<uint32_t*>(chunk + 64) = pb;
    }
    double4 sums = block_sum4_f64(make_double4(n2, e2, c2, static_cast<double>(b2)), scratch4);
    if (threadIdx.x == 0) {
        I8Stats st{};
        if (finite) {
            // round_up_f's 2^-20 covers the one rounding of h2 and of each
            // square root
            const double h2 = (sd * sd) * sums.z;
            st.norm2 = sums.x;
            st.scale = scale;
            st.err = round_up_f(sqrt(sums.y));
            st.hat = round_up_f(sqrt(h2));
            st.lo = round_up_f(sd * sqrt(sums.w));
            st.norm = round_up_f(sqrt(sums.x));
        } else {
            const float nan = __builtin_nanf("");
            st.norm2 = static_cast<double>(nan);
            st.scale = st.err = st.hat = st.lo = st.norm = nan;
        }
        stats[row] = st;
    }
}

// est = ||q||^2 + ||v||^2 - 2 s_q s_v (H + X / 256): H + X / 256 and s_q s_v
// are exact in binary64, their product and the sums round once each, and the
// result is rounded to binary32.
__device__ __forceinline__ float i8_estimate(int h, int x, const I8Stats& q, double nv, double sv)
{
    const double t = static_cast<double>(h) + static_cast<double>(x) * 0.00390625;
    const double dot = (static_cast<double>(q.scale) * sv) * t;
    return static_cast<float>((q.norm2 + nv) - 2.0 * dot);
}

template <int TM, int TN>
__device__ __forceinline__ void i8_epilogue(const intx16 (&h)[TM][TN], const intx16 (&x)[TM][TN], uint32_t mw,
                                            uint32_t nw, int lane, uint32_t M, uint32_t N,
                                            const I8Stats* __restrict__ qs, const I8Stats* __restrict__ vs,
                                            float* __restrict__ out, uint64_t ldo)
{
#pragma unroll
    for (int i = 0; i < TM; ++i) {
#pragma unroll
        for (int j = 0; j < TN; ++j) {
            const uint32_t n = nw + j * 32 + (lane % 32);
            if (n >= N) continue;
            const double nv = vs[n].norm2;
            const double sv = vs[n].scale;
#pragma unroll
            for (int v = 0; v < 16; ++v) {
                const uint32_t m = mw + i * 32 + 8 * (v / 4) + 4 * (lane / 32) + (v % 4);
                if (m < M) out[static_cast<uint64_t>(m) * ldo + n] = i8_estimate(h[i][j][v], x[i][j][v], qs[m], nv, sv);
            }
        }
    }
}

// A split tile's two int32 accumulator sets (bit patterns in the float
// scratch): set s, chunk c = (i*TN + j)*4 + v/4 at ((s*TM*TN*4 + c)*NT + tid)*4.
template <int NT, int TM, int TN>
__device__ __forceinline__ void store_i8_partials(float* __restrict__ dst, const intx16 (&h)[TM][TN],
                                                  const intx16 (&x)[TM][TN], int tid)
{
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j)
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                const int chunk = (i * TN + j) * 4 + c;
                *reinterpret_cast<intx4*>(dst + (static_cast<uint64_t>(chunk) * NT + tid) * 4) =
                    intx4{h[i][j][4 * c], h[i][j][4 * c + 1], h[i][j][4 * c + 2], h[i][j][4 * c + 3]};
                *reinterpret_cast<intx4*>(dst + (static_cast<uint64_t>(TM * TN * 4 + chunk) * NT + tid) * 4) =
                    intx4{x[i][j][4 * c], x[i][j][4 * c + 1], x[i][j][4 * c + 2], x[i][j][4 * c + 3]};
            }
}

// Sums the K parts of one split tile (int32, exact) and applies the epilogue.
template <int WM, int WN, int TM, int TN>
__global__ void __launch_bounds__(64 * WM * WN)
gram_i8_fixup_kernel(GramSplit sp, uint32_t M, uint32_t N, const I8Stats* __restrict__ qs,
                     const I8Stats* __restrict__ vs, float* __restrict__ out, uint64_t ldo, uint32_t tiles_m,
                     uint32_t tiles_n, uint32_t group_m)
{
    constexpr int NT = 64 * WM * WN;
    constexpr int BM = 32 * TM * WM;
    constexpr int BN = 32 * TN * WN;
    const int tid = threadIdx.x;
    const int lane = tid % 64;
    const int wave = tid / 64;
    const uint2 o = tile_origin(sp.full + blockIdx.x, tiles_m, tiles_n, group_m);
    intx16 h[TM][TN], x[TM][TN];
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j)
#pragma unroll
            for (int v = 0; v < 16; ++v) h[i][j][v] = x[i][j][v] = 0;
    for (uint32_t p = 0; p < sp.parts; ++p) {
        const float* src = sp.partials + (static_cast<uint64_t>(blockIdx.x) * sp.parts + p) * (2 * BM * BN);
#pragma unroll
        for (int i = 0; i < TM; ++i)
#pragma unroll
            for (int j = 0; j < TN; ++j)
#pragma unroll
                for (int c = 0; c < 4; ++c) {
                    const int chunk = (i * TN + j) * 4 + c;
                    const intx4 hv = *reinterpret_cast<const intx4*>(src + (static_cast<uint64_t>(chunk) * NT + tid) * 4);
                    const intx4 xv = *reinterpret_cast<const intx4*>(
                        src + (static_cast<uint64_t>(TM * TN * 4 + chunk) * NT + tid) * 4);
#pragma unroll
                    for (int e = 0; e < 4; ++e) {
                        h[i][j][4 * c + e] += hv[e];
                        x[i][j][4 * c + e] += xv[e];
                    }
                }
    }
    i8_epilogue<TM, TN>(h, x, o.x * BM + (wave / WN) * 32 * TM, o.y * BN + (wave % WN) * 32 * TN, lane, M, N, qs,
                        vs, out, ldo);
}

// Blocks of WM x WN waves, each a (32 TM) x (32 TN) output tile with two
// int32 accumulator sets (H and X) of TM x TN 32x32 MFMA tiles. A K step is
// one 128-byte chunk per row (64 columns of a, then of b), double-buffered in
// LDS with the next step's loads issued before the barrier, as in the BF16
// kernel. Per 32 columns and output tile: one MFMA for H, two for X.
// Four-wave blocks are bounded to two blocks per compute unit (at most 256
// registers per lane), so one block's barrier wait overlaps the other's MFMAs.
template <int WM, int WN, int TM, int TN>
__global__ void __launch_bounds__(64 * WM * WN, WM * WN == 4 ? 2 : 1)
gram_i8x2_kernel(const uint8_t* __restrict__ Q, uint64_t ldq, uint32_t M, const uint8_t* __restrict__ V,
                 uint64_t ldv, uint32_t N, uint32_t K, const I8Stats* __restrict__ qs,
                 const I8Stats* __restrict__ vs, float* __restrict__ out, uint64_t ldo, uint32_t tiles_m,
                 uint32_t tiles_n, uint32_t group_m, GramSplit sp)
{
    constexpr int NT = 64 * WM * WN;
    constexpr int BM = 32 * TM * WM;
    constexpr int BN = 32 * TN * WN;
    static_assert((BM * 8) % NT == 0 && (BN * 8) % NT == 0, "tile rows must split evenly over the threads");
    constexpr int CA = BM * 8 / NT;  // 16-byte chunks per thread per K step
    constexpr int CB = BN * 8 / NT;
    __shared__ __attribute__((aligned(16))) uint8_t As[2 * BM * kLdsRowI8];
    __shared__ __attribute__((aligned(16))) uint8_t Bs[2 * BN * kLdsRowI8];

    const GramBlock gb = gram_block(blockIdx.x, sp, K, kBKI8);
    const uint2 o = tile_origin(gb.tile, tiles_m, tiles_n, group_m);
    const uint32_t m0 = o.x * BM;
    const uint32_t n0 = o.y * BN;

    const int tid = threadIdx.x;
    const int lane = tid % 64;
    const int wave = tid / 64;
    const int wm = (wave / WN) * 32 * TM;
    const int wn = (wave % WN) * 32 * TN;

    u32x4 ra[CA], rb[CB];
    const uint8_t* qbase = Q + static_cast<uint64_t>(m0 + tid / 8) * ldq + (tid % 8) * 16;
    const uint8_t* vbase = V + static_cast<uint64_t>(n0 + tid / 8) * ldv + (tid % 8) * 16;
    const uint64_t qstep = static_cast<uint64_t>(NT / 8) * ldq;
    const uint64_t vstep = static_cast<uint64_t>(NT / 8) * ldv;
    const int lds0 = (tid / 8) * kLdsRowI8 + (tid % 8) * 16;
    constexpr int kLdsStep = (NT / 8) * kLdsRowI8;
    // MFMA operand fragments: lane holds row (lane % 32) and 16 K values
    // 16*(lane/32) .. +15 of each 32-column step, the same assignment for
    // both operands, as the dot product needs.
    const int a_off = (wm + lane % 32) * kLdsRowI8 + (lane / 32) * 16;
    const int b_off = (wn + lane % 32) * kLdsRowI8 + (lane / 32) * 16;

    intx16 acc_h[TM][TN], acc_x[TM][TN];
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j)
#pragma unroll
            for (int v = 0; v < 16; ++v) acc_h[i][j][v] = acc_x[i][j][v] = 0;

#define UHANN_I8_LOAD(k0)                                                                      \
    _Pragma("unroll") for (int c = 0; c < CA; ++c) ra[c] =                                     \
        *reinterpret_cast<const u32x4*>(qbase + c * qstep + (k0));                             \
    _Pragma("unroll") for (int c = 0; c < CB; ++c) rb[c] =                                     \
        *reinterpret_cast<const u32x4*>(vbase + c * vstep + (k0));                             \
    __builtin_amdgcn_sched_barrier(0);
#define UHANN_I8_STORE(buf)                                                                    \
    _Pragma("unroll") for (int c = 0; c < CA; ++c)                                             \
        *reinterpret_cast<u32x4*>(As + (buf) * BM * kLdsRowI8 + lds0 + c * kLdsStep) = ra[c];  \
    _Pragma("unroll") for (int c = 0; c < CB; ++c)                                             \
        *reinterpret_cast<u32x4*>(Bs + (buf) * BN * kLdsRowI8 + lds0 + c * kLdsStep) = rb[c];
#define UHANN_I8_COMPUTE(buf)                                                                  \
    _Pragma("unroll") for (int kk = 0; kk < 64; kk += 32)                                      \
    {                                                                                          \
        const uint8_t* ab = As + (buf) * BM * kLdsRowI8 + a_off + kk;                          \
        const uint8_t* bb = Bs + (buf) * BN * kLdsRowI8 + b_off + kk;                          \
        intx4 ah[TM], al[TM], bh[TN], bl[TN];                                                  \
        _Pragma("unroll") for (int i = 0; i < TM; ++i) {                                       \
            ah[i] = *reinterpret_cast<const intx4*>(ab + i * 32 * kLdsRowI8);                  \
            al[i] = *reinterpret_cast<const intx4*>(ab + i * 32 * kLdsRowI8 + 64);             \
        }                                                                                      \
        _Pragma("unroll") for (int j = 0; j < TN; ++j) {                                       \
            bh[j] = *reinterpret_cast<const intx4*>(bb + j * 32 * kLdsRowI8);                  \
            bl[j] = *reinterpret_cast<const intx4*>(bb + j * 32 * kLdsRowI8 + 64);             \
        }                                                                                      \
        _Pragma("unroll") for (int i = 0; i < TM; ++i)                                         \
            _Pragma("unroll") for (int j = 0; j < TN; ++j) acc_h[i][j] =                       \
                __builtin_amdgcn_mfma_i32_32x32x32_i8(ah[i], bh[j], acc_h[i][j], 0, 0, 0);     \
        _Pragma("unroll") for (int i = 0; i < TM; ++i)                                         \
            _Pragma("unroll") for (int j = 0; j < TN; ++j) acc_x[i][j] =                       \
                __builtin_amdgcn_mfma_i32_32x32x32_i8(ah[i], bl[j], acc_x[i][j], 0, 0, 0);     \
        _Pragma("unroll") for (int i = 0; i < TM; ++i)                                         \
            _Pragma("unroll") for (int j = 0; j < TN; ++j) acc_x[i][j] =                       \
                __builtin_amdgcn_mfma_i32_32x32x32_i8(al[i], bh[j], acc_x[i][j], 0, 0, 0);     \
    }

    const uint32_t k_last = gb.k_end - kBKI8;
    UHANN_I8_LOAD(gb.k_begin)
    UHANN_I8_STORE(0)
    UHANN_I8_LOAD(min(gb.k_begin + kBKI8, k_last))
    __syncthreads();
    int cur = 0;
    for (uint32_t k0 = gb.k_begin + kBKI8; k0 < gb.k_end; k0 += kBKI8) {
        UHANN_I8_COMPUTE(cur)
        UHANN_I8_STORE(cur ^ 1)
        UHANN_I8_LOAD(min(k0 + kBKI8, k_last))
        __syncthreads();
        cur ^= 1;
    }
    UHANN_I8_COMPUTE(cur)
#undef UHANN_I8_LOAD
#undef UHANN_I8_STORE
#undef UHANN_I8_COMPUTE

    if (gb.part >= 0) {
        const uint64_t unit = static_cast<uint64_t>(gb.tile - sp.full) * sp.parts + gb.part;
        store_i8_partials<NT, TM, TN>(sp.partials + unit * (2 * BM * BN), acc_h, acc_x, tid);
        return;
    }
    i8_epilogue<TM, TN>(acc_h, acc_x, m0 + wm, n0 + wn, lane, M, N, qs, vs, out, ldo);
}

// s_waitcnt immediates (gfx9 encoding: vmcnt in bits 3:0 and 15:14, expcnt in
// 6:4, lgkmcnt in 11:8; a field at its maximum does not wait).
constexpr int kWaitLgkm0 = 0xC07F;  // lgkmcnt(0)
constexpr int kWaitVm6 = 0x0F76;    // vmcnt(6)
constexpr int kWaitVm0 = 0x0F70;    // vmcnt(0)
// Word 3 of a raw (stride 0) buffer resource: 32-bit data format, no swizzle.
constexpr int kRawBufferWord3 = 0x00020000;

constexpr int kPpBM = 256;
constexpr int kPpBN = 128;
constexpr int kPpStage = (kPpBM + kPpBN) * kBKI8;  // one K step of both operands: 49152 bytes

typedef __attribute__((address_space(3))) void* lds_ptr_t;

// One K step's MFMA operands of a 64 x 64 wave tile: [sub-step][tile].
struct PpFrags {
    intx4 ah[2][2], al[2][2], bh[2][2], bl[2][2];
};

// Ping-pong GRAM_INT8 block, 256 x 128 outputs, 8 waves in two groups of
// four. Wave w belongs to group w / 4 and computes the 64 x 64 outputs at rows
// 64 (w / 2), columns 64 (w % 2) (the 4 x 2 wave grid of shape 0, so the
// split-tile fix-up is shared). Waves w and w + 4 run on the same SIMD.
//
// K steps (128 bytes per row) reach LDS by direct buffer-to-LDS loads (no
// staging registers, no LDS store instructions) into a ring of three stages,
// two steps ahead. Within a stage a row's 16-byte chunk c sits in slot
// c ^ ((row / 2) % 8): LDS serves 256 bytes (64 banks) per cycle, i.e. 16
// lanes of a 16-byte read, and the 16 rows that a fragment read presents
// together then cover the 16 bank groups once (rows of one parity share a
// 128-byte half). A load instruction writes 8 consecutive rows (1 KB), its
// lanes fetching the chunks in slot order.
//
// The two groups alternate, half a step apart, between
//   memory phase (step s): read all 16 fragments of step s into registers,
//                          issue this wave's loads of step s + 2;
//   MFMA phase (step s):   the 24 MFMAs of step s from registers;
// with a block barrier between phases, so on each SIMD one wave feeds the
// matrix core while the other reads LDS. Barrier B_t separates half steps;
// group 0 reads step s in half step 2s + 1, group 1 in 2s + 2. Step s + 2 is
// loaded (by both groups, in their memory phases of step s) into the stage
// that held step s - 1, free since B_{2s+1}, and each wave waits for its own
// loads of step s + 1 (vmcnt(6): only step s + 2's six loads may remain)
// before B_{2s+3}, the barrier after which group 0 reads step s + 1. Each
// group passes 2S + 2 barriers in all (S steps). Steps past the last reload
// the last step into a stage nobody reads again, so the load counts stay
// uniform; every load has landed before the epilogue.
__global__ void __launch_bounds__(512)
gram_i8x2_pp_kernel(const uint8_t* __restrict__ Q, uint64_t ldq, uint32_t M, const uint8_t* __restrict__ V,
                    uint64_t ldv, uint32_t N, uint32_t K, const I8Stats* __restrict__ qs,
                    const I8Stats* __restrict__ vs, float* __restrict__ out, uint64_t ldo, uint32_t tiles_m,
                    uint32_t tiles_n, uint32_t group_m, GramSplit sp)
{
    __shared__ __attribute__((aligned(1024))) uint8_t st0[kPpStage];
    __shared__ __attribute__((aligned(1024))) uint8_t st1[kPpStage];
    __shared__ __attribute__((aligned(1024))) uint8_t st2[kPpStage];

    const GramBlock gb = gram_block(blockIdx.x, sp, K, kBKI8);
    const uint2 o = tile_origin(gb.tile, tiles_m, tiles_n, group_m);
    const uint32_t m0 = o.x * kPpBM;
    const uint32_t n0 = o.y * kPpBN;

    const int tid = threadIdx.x;
    const int lane = tid % 64;
    const int wave = __builtin_amdgcn_readfirstlane(tid / 64);
    const int grp = wave / 4;
    const int lw = wave % 4;

    // Loads: this wave brings query rows 128 grp + 32 lw + [0, 32) (four
    // instructions) and reference rows 64 grp + 16 lw + [0, 16) (two); lane L
    // fetches, for row R = R0 + L / 8 of an instruction's eight (R0 a multiple
    // of 8), the chunk in slot L % 8: (L % 8) ^ ((R / 2) % 8), where
    // (R / 2) % 8 = (L / 16) + 4 (R0 % 16 == 8), so it alternates between two
    // lane patterns with the instruction's parity u % 2.
    const uint32_t ldq32 = static_cast<uint32_t>(ldq);
    const uint32_t ldv32 = static_cast<uint32_t>(ldv);
    const __amdgpu_buffer_rsrc_t rq = __builtin_amdgcn_make_buffer_rsrc(
        const_cast<uint8_t*>(Q + static_cast<uint64_t>(m0) * ldq), 0, static_cast<int>(kPpBM * ldq32),
        kRawBufferWord3);
    const __amdgpu_buffer_rsrc_t rv = __builtin_amdgcn_make_buffer_rsrc(
        const_cast<uint8_t*>(V + static_cast<uint64_t>(n0) * ldv), 0, static_cast<int>(kPpBN * ldv32),
        kRawBufferWord3);
    const int qrow0 = 128 * grp + 32 * lw;
    const int vrow0 = 64 * grp + 16 * lw;
    const uint32_t chunk0 = static_cast<uint32_t>((lane % 8) ^ (lane / 16));
    const uint32_t chunk1 = chunk0 ^ 4U;
    const uint32_t qoff0 = static_cast<uint32_t>(qrow0 + lane / 8) * ldq32 + 16U * chunk0;
    const uint32_t qoff1 = static_cast<uint32_t>(qrow0 + lane / 8) * ldq32 + 16U * chunk1;
    const uint32_t voff0 = static_cast<uint32_t>(vrow0 + lane / 8) * ldv32 + 16U * chunk0;
    const uint32_t voff1 = static_cast<uint32_t>(vrow0 + lane / 8) * ldv32 + 16U * chunk1;

    // Fragments: lane l holds row l % 32 and bytes 16 (l / 32) .. + 15 of a
    // 32-byte sub-step (logical chunk 2 kk + l / 32 of the a plane, + 4 for
    // the b plane), as in gram_i8x2_kernel.
    const int lrow = lane % 32;
    const int a_base = (128 * grp + 64 * (lw / 2) + lrow) * kBKI8;
    const int b_base = (kPpBM + 64 * (lw % 2) + lrow) * kBKI8;
    int coff[2][2];
#pragma unroll
    for (int kk = 0; kk < 2; ++kk)
#pragma unroll
        for (int p = 0; p < 2; ++p) coff[kk][p] = 16 * ((2 * kk + 4 * p + lane / 32) ^ ((lrow / 2) % 8));

    intx16 acc_h[2][2], acc_x[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int v = 0; v < 16; ++v) acc_h[i][j][v] = acc_x[i][j][v] = 0;

    const uint32_t S = (gb.k_end - gb.k_begin) / kBKI8;
    const uint32_t k_last = gb.k_end - kBKI8;

    auto issue = [&](uint8_t* st, uint32_t step) {
        const uint32_t k = min(gb.k_begin + step * kBKI8, k_last);
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            __builtin_amdgcn_raw_ptr_buffer_load_lds(rq, (lds_ptr_t)(st + (qrow0 + 8 * u) * kBKI8), 16,
                                                     u % 2 ? qoff1 : qoff0, k + 8U * u * ldq32, 0, 0);
        }
#pragma unroll
        for (int u = 0; u < 2; ++u) {
            __builtin_amdgcn_raw_ptr_buffer_load_lds(rv, (lds_ptr_t)(st + (kPpBM + vrow0 + 8 * u) * kBKI8), 16,
                                                     u % 2 ? voff1 : voff0, k + 8U * u * ldv32, 0, 0);
        }
    };
    auto read = [&](const uint8_t* st, PpFrags& f) {
#pragma unroll
        for (int kk = 0; kk < 2; ++kk)
#pragma unroll
            for (int i = 0; i < 2; ++i) {
                f.ah[kk][i] = *reinterpret_cast<const intx4*>(st + a_base + i * 32 * kBKI8 + coff[kk][0]);
                f.al[kk][i] = *reinterpret_cast<const intx4*>(st + a_base + i * 32 * kBKI8 + coff[kk][1]);
                f.bh[kk][i] = *reinterpret_cast<const intx4*>(st + b_base + i * 32 * kBKI8 + coff[kk][0]);
                f.bl[kk][i] = *reinterpret_cast<const intx4*>(st + b_base + i * 32 * kBKI8 + coff[kk][1]);
            }
    };
    auto mfma = [&](const PpFrags& f) {
#pragma unroll
        for (int kk = 0; kk < 2; ++kk) {
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 2; ++j)
                    acc_h[i][j] = __builtin_amdgcn_mfma_i32_32x32x32_i8(f.ah[kk][i], f.bh[kk][j], acc_h[i][j], 0, 0, 0);
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 2; ++j)
                    acc_x[i][j] = __builtin_amdgcn_mfma_i32_32x32x32_i8(f.ah[kk][i], f.bl[kk][j], acc_x[i][j], 0, 0, 0);
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 2; ++j)
                    acc_x[i][j] = __builtin_amdgcn_mfma_i32_32x32x32_i8(f.al[kk][i], f.bh[kk][j], acc_x[i][j], 0, 0, 0);
        }
    };
    // A block barrier that nothing is scheduled across (no fence: the loads'
    // completion is awaited explicitly with s_waitcnt).
#define UHANN_PP_SYNC()                       \
    __builtin_amdgcn_sched_barrier(0);        \
    asm volatile("" ::: "memory");            \
    __builtin_amdgcn_s_barrier();             \
    asm volatile("" ::: "memory");            \
    __builtin_amdgcn_sched_barrier(0);
#define UHANN_PP_G0(CUR, NXT, s)                  \
    read(CUR, f);                                 \
    issue(NXT, (s) + 2);                          \
    __builtin_amdgcn_s_waitcnt(kWaitLgkm0);       \
    UHANN_PP_SYNC()                               \
    mfma(f);                                      \
    __builtin_amdgcn_s_waitcnt(kWaitVm6);         \
    UHANN_PP_SYNC()
#define UHANN_PP_G1(CUR, NXT, s)                  \
    read(CUR, f);                                 \
    issue(NXT, (s) + 2);                          \
    __builtin_amdgcn_s_waitcnt(kWaitLgkm0);       \
    __builtin_amdgcn_s_waitcnt(kWaitVm6);         \
    UHANN_PP_SYNC()                               \
    mfma(f);                                      \
    UHANN_PP_SYNC()

    PpFrags f;
    // Prologue: steps 0 and 1; step 0 visible to all at B_1.
    issue(st0, 0);
    issue(st1, 1);
    __builtin_amdgcn_s_waitcnt(kWaitVm6);
    UHANN_PP_SYNC()
    uint32_t s = 0;
    if (grp == 0) {
        for (; s + 3 <= S; s += 3) {
            UHANN_PP_G0(st0, st2, s)
            UHANN_PP_G0(st1, st0, s + 1)
            UHANN_PP_G0(st2, st1, s + 2)
        }
        if (s < S) {
            UHANN_PP_G0(st0, st2, s)
            ++s;
        }
        if (s < S) {
            UHANN_PP_G0(st1, st0, s)
            ++s;
        }
        UHANN_PP_SYNC()  // B_{2S+2}, group 1's last MFMA phase
    } else {
        UHANN_PP_SYNC()  // B_2: group 0 reads step 0 meanwhile
        for (; s + 3 <= S; s += 3) {
            UHANN_PP_G1(st0, st2, s)
            UHANN_PP_G1(st1, st0, s + 1)
            UHANN_PP_G1(st2, st1, s + 2)
        }
        if (s < S) {
            UHANN_PP_G1(st0, st2, s)
            ++s;
        }
        if (s < S) {
            UHANN_PP_G1(st1, st0, s)
            ++s;
        }
    }
#undef UHANN_PP_G0
#undef UHANN_PP_G1
#undef UHANN_PP_SYNC
    __builtin_amdgcn_s_waitcnt(kWaitVm0);

    if (gb.part >= 0) {
        const uint64_t unit = static_cast<uint64_t>(gb.tile - sp.full) * sp.parts + gb.part;
        store_i8_partials<512, 2, 2>(sp.partials + unit * (2 * kPpBM * kPpBN), acc_h, acc_x, tid);
        return;
    }
    i8_epilogue<2, 2>(acc_h, acc_x, m0 + 64 * (wave / 2), n0 + 64 * (wave % 2), lane, M, N, qs, vs, out, ldo);
}

// Lower and upper ends of ||q - v||^2 from a GRAM_INT8 estimate e: the bound
// of the file comment above, computed in binary64 from the rounded-up row
// norms, plus 2^-23 |e| for the binary32 rounding of the estimate and
// 1e-10 (||q||^2 + ||v||^2) for the binary64 sums and roundings behind it;
// the ends are rounded outwards to binary32. A non-finite estimate gives NaN
// ends (never a candidate; the query escalates when nothing else is).
__device__ __forceinline__ void i8_ends(const I8Stats& q, const I8Stats& v, float e, float& lo, float& up)
{
    const double beta = static_cast<double>(q.norm) * v.err + static_cast<double>(q.err) * v.hat +
                        static_cast<double>(q.lo) * v.lo;
    const double tau = 2.0 * beta + 0x1p-23 * fabs(static_cast<double>(e)) + 1e-10 * (q.norm2 + v.norm2);
    lo = __double2float_rd(static_cast<double>(e) - tau);
    up = __double2float_ru(static_cast<double>(e) + tau);
}

// One block per query. U = min_r (upper end); every row whose binary32
// distance could be the binary32 minimum m, or lie in its near-tie band
// (<= m (1+g)/(1-g)), has ||q - v||^2 <= U ((1+g)/(1-g))^2 and is kept. The
// candidates are re-ranked exactly as in the BF16 strategy; more than
// kMaxCandidates (or none: a non-finite query) escalates to the direct scan.
__global__ void __launch_bounds__(kSelThreads)
select_rerank_i8_kernel(SelectArgs a)
{
    __shared__ uint32_t s_min;
    __shared__ uint32_t s_count;
    __shared__ uint32_t s_cand[kMaxCandidates];
    __shared__ float s_dist[kMaxCandidates];
    __shared__ float s_group[8][kRerankRows];
    __shared__ double s_d64[kMaxCandidates];
    __shared__ uint32_t s_best;
    __shared__ float s_best_d;

    const uint32_t q = blockIdx.x;
    if (q >= a.nq) return;
    const int t = threadIdx.x;
    const I8Stats qs = a.qstats8[q];
    const float* arow = a.approx + static_cast<uint64_t>(q) * a.ld_approx;

    if (t == 0) {
        s_min = 0xffffffffU;
        s_count = 0;
    }
    __syncthreads();
    uint32_t local = 0xffffffffU;
    float lo_keep[kKeepLo];
#pragma unroll
    for (int i = 0; i < kKeepLo; ++i) {
        const uint32_t r = t + static_cast<uint32_t>(i) * kSelThreads;
        lo_keep[i] = __builtin_inff();
        if (r < a.rows) {
            float lo, up;
            i8_ends(qs, a.vstats8[r], arow[r], lo, up);
            local = min(local, ordered(up));
            lo_keep[i] = lo;
        }
    }
    for (uint32_t r = t + kKeepLo * kSelThreads; r < a.rows; r += kSelThreads) {
        float lo, up;
        i8_ends(qs, a.vstats8[r], arow[r], lo, up);
        local = min(local, ordered(up));
    }
    for (int off = 32; off >= 1; off /= 2) local = min(local, __shfl_down(local, off));
    if (t % 64 == 0) atomicMin(&s_min, local);
    __syncthreads();
    // Rounded up: the product's rounding is far below the 2^-40 added.
    const double threshold = static_cast<double>(unordered(s_min)) * a.fp32_band * a.fp32_band * (1.0 + 0x1p-40);
#pragma unroll
    for (int i = 0; i < kKeepLo; ++i) {
        if (static_cast<double>(lo_keep[i]) <= threshold) {
            const uint32_t slot = atomicAdd(&s_count, 1U);
            if (slot < kMaxCandidates) s_cand[slot] = t + static_cast<uint32_t>(i) * kSelThreads;
        }
    }
    for (uint32_t r = t + kKeepLo * kSelThreads; r < a.rows; r += kSelThreads) {
        float lo, up;
        i8_ends(qs, a.vstats8[r], arow[r], lo, up);
        if (static_cast<double>(lo) <= threshold) {
            const uint32_t slot = atomicAdd(&s_count, 1U);
            if (slot < kMaxCandidates) s_cand[slot] = r;
        }
    }
    __syncthreads();
    const uint32_t count = s_count;
    if (count > kMaxCandidates || count == 0) {
        if (t == 0) {
            const uint32_t slot = atomicAdd(a.escalate_count, 1U);
            a.escalate_list[slot] = q;
            a.out_idx[q] = 0;
            a.out_dist[q] = 0.0F;
            if (a.stats != nullptr) atomicAdd(&a.stats->escalated, 1U);
        }
        return;
    }

    uint32_t best;
    float best_d;
    rerank_pick<kRerankRows>(a, q, s_cand, count, s_dist, s_group, s_d64, &s_best, &s_best_d, best, best_d);
    if (t == 0) {
        a.out_idx[q] = best;
        a.out_dist[q] = best_d;
        if (a.stats != nullptr) {
            atomicAdd(&a.stats->candidates, count);
            atomicAdd(&a.stats->certified, 1U);
        }
    }
}

}  // namespace

void to_bf16_rows(const float* src, uint64_t ld_src, uint32_t rows, uint32_t dim, __bf16* dst,
                  uint64_t ld_dst, RowStats* stats, hipStream_t stream)
{
    if (rows == 0) return;
    to_bf16_rows_kernel<<<rows, 256, 0, stream>>>(src, ld_src, rows, dim, dst, ld_dst, stats);
    check_launch("bf16 conversion kernel");
}

namespace {

constexpr uint32_t kMaxParts = 16;
constexpr uint32_t kMinPartSteps = 8;

uint32_t current_device_cus()
{
    int dev = 0;
    int cus = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
    return cus > 0 ? static_cast<uint32_t>(cus) : 0U;
}

// The plan for this launch if the scratch holds it, else every tile whole.
GramSplit split_for(uint32_t tiles, uint32_t ksteps, uint64_t block_outputs, float* partials,
                    uint64_t partial_floats)
{
    GramSplit none;
    none.full = tiles;
    if (partials == nullptr) return none;
    GramSplit sp = gram_split_plan(tiles, ksteps, current_device_cus());
    if (sp.tail == 0 || static_cast<uint64_t>(sp.tail) * sp.parts * block_outputs > partial_floats) return none;
    sp.partials = partials;
    return sp;
}

}  // namespace

GramSplit gram_split_plan(uint32_t tiles, uint32_t ksteps, uint32_t cus)
{
    GramSplit sp;
    sp.full = tiles;
    if (cus == 0) return sp;
    const uint32_t full = tiles / cus * cus;
    const uint32_t tail = tiles - full;
    if (tail == 0) return sp;
    // Rounds the tail costs with p parts (in whole-tile times): ceil(tail p / cus) / p.
    uint32_t best_p = 1;
    uint64_t best_num = 1, best_den = 1;
    const uint32_t max_p = std::min(kMaxParts, ksteps / kMinPartSteps);
    for (uint32_t p = 2; p <= max_p; ++p) {
        const uint64_t units = static_cast<uint64_t>(tail) * p;
        if (units > 3ULL * cus) break;
        const uint64_t rounds = (units + cus - 1) / cus;
        if (rounds * best_den < best_num * p) {
            best_num = rounds;
            best_den = p;
            best_p = p;
        }
    }
    if (best_p == 1) return sp;
    sp.full = full;
    sp.tail = tail;
    sp.part_steps = (ksteps + best_p - 1) / best_p;
    sp.parts = (ksteps + sp.part_steps - 1) / sp.part_steps;
    return sp;
}

uint64_t gram_split_floats(uint32_t max_m, uint32_t n, uint32_t k_steps, uint32_t cus)
{
    const uint32_t tiles_n = (n + 255U) / 256U;
    uint64_t need = 0;
    for (uint32_t tm = 1; tm <= (max_m + 255U) / 256U; ++tm) {
        const GramSplit sp = gram_split_plan(tm * tiles_n, k_steps, cus);
        need = std::max(need, static_cast<uint64_t>(sp.tail) * sp.parts * gram_block_outputs());
    }
    return need;
}

uint32_t gram_block_outputs() { return 256U * 256U; }

void gram_bf16(const __bf16* q, uint64_t ldq, uint32_t m, const __bf16* v, uint64_t ldv, uint32_t n,
               uint32_t k, const RowStats* qs, const RowStats* vs, float* out, uint64_t ldo,
               int tile, float* partials, uint64_t partial_floats, hipStream_t stream)
{
    if (m == 0 || n == 0) return;
    if (k % kBK != 0) throw std::invalid_argument("Gram K must be a multiple of 64");
    // tile: waves WM x WN, MFMA tiles per wave TM x TN, block (32 TM WM) x (32 TN WN):
    //   0: 2x2 waves, 2x2 (128x128)   1: 4x2, 2x2 (256x128)   2: 2x4, 2x2 (128x256)
    //   3: 4x4, 2x2 (256x256)         4: 2x2, 4x4 (256x256)   5: 2x4, 4x2 (256x256)
    //   6: 4x2, 2x4 (256x256)         7: 2x2, 4x2 (256x128)
    struct Shape {
        int wm, wn, tm, tn;
    };
    static const Shape shapes[8] = {{2, 2, 2, 2}, {4, 2, 2, 2}, {2, 4, 2, 2}, {4, 4, 2, 2},
                                    {2, 2, 4, 4}, {2, 4, 4, 2}, {4, 2, 2, 4}, {2, 2, 4, 2}};
    if (tile < 0 || tile >= gram_tile_count()) throw std::invalid_argument("unknown Gram tile");
    const Shape sh = shapes[tile];
    const uint32_t bm = 32U * sh.tm * sh.wm;
    const uint32_t bn = 32U * sh.tn * sh.wn;
    const uint32_t tiles_m = (m + bm - 1) / bm;
    const uint32_t tiles_n = (n + bn - 1) / bn;
    const uint32_t group_m = 8;
    const GramSplit sp = split_for(tiles_m * tiles_n, k / kBK, static_cast<uint64_t>(bm) * bn, partials,
                                   partial_floats);
    const dim3 grid(sp.full + sp.tail * sp.parts);
    const dim3 block(64 * sh.wm * sh.wn);
#define UHANN_GRAM_CASE(T, A, B, C, D)                                                                   \
    case T:                                                                                            \
        gram_bf16_kernel<A, B, C, D><<<grid, block, 0, stream>>>(q, ldq, m, v, ldv, n, k, qs, vs, out, ldo, \
                                                                tiles_m, tiles_n, group_m, sp);         \
        if (sp.tail > 0) {                                                                             \
            gram_fixup_kernel<A, B, C, D, false><<<sp.tail, block, 0, stream>>>(sp, m, n, qs, vs, out, ldo, \
                                                                               tiles_m, tiles_n, group_m); \
        }                                                                                              \
        break;
    switch (tile) {
        UHANN_GRAM_CASE(0, 2, 2, 2, 2)
        UHANN_GRAM_CASE(1, 4, 2, 2, 2)
        UHANN_GRAM_CASE(2, 2, 4, 2, 2)
        UHANN_GRAM_CASE(3, 4, 4, 2, 2)
        UHANN_GRAM_CASE(4, 2, 2, 4, 4)
        UHANN_GRAM_CASE(5, 2, 4, 4, 2)
        UHANN_GRAM_CASE(6, 4, 2, 2, 4)
        UHANN_GRAM_CASE(7, 2, 2, 4, 2)
    default:
        break;
    }
#undef UHANN_GRAM_CASE
    check_launch("bf16 Gram kernel");
}

void select_rerank(const SelectArgs& args, hipStream_t stream)
{
    if (args.nq == 0) return;
    select_rerank_kernel<<<args.nq, kSelThreads, 0, stream>>>(args);
    check_launch("candidate selection / re-rank kernel");
}

void to_fp8_rows(const float* src, uint64_t ld_src, uint32_t rows, uint32_t dim, uint8_t* dst, uint64_t ld_dst,
                 RowStats* stats, hipStream_t stream)
{
    if (rows == 0) return;
    to_fp8_rows_kernel<<<rows, 256, 0, stream>>>(src, ld_src, rows, dim, dst, ld_dst, stats);
    check_launch("fp8 conversion kernel");
}

void gram_fp8(const uint8_t* q, uint64_t ldq, uint32_t m, const uint8_t* v, uint64_t ldv, uint32_t n, uint32_t k,
              const RowStats* qs, const RowStats* vs, float* out, uint64_t ldo, float* partials,
              uint64_t partial_floats, hipStream_t stream)
{
    if (m == 0 || n == 0) return;
    if (k % kBK8 != 0) throw std::invalid_argument("FP8 Gram K must be a multiple of 128");
    if (ldq % 16 != 0 || ldv % 16 != 0) throw std::invalid_argument("FP8 Gram rows must be 16-byte aligned");
    const uint32_t tiles_m = (m + 255U) / 256U;
    const uint32_t tiles_n = (n + 255U) / 256U;
    const GramSplit sp = split_for(tiles_m * tiles_n, k / kBK8, gram_block_outputs(), partials, partial_floats);
    gram_fp8_kernel<<<sp.full + sp.tail * sp.parts, 1024, 0, stream>>>(q, ldq, m, v, ldv, n, k, qs, vs, out, ldo,
                                                                      tiles_m, tiles_n, 8U, sp);
    if (sp.tail > 0) {
        gram_fixup_kernel<4, 4, 2, 2, true><<<sp.tail, 1024, 0, stream>>>(sp, m, n, qs, vs, out, ldo, tiles_m,
                                                                          tiles_n, 8U);
    }
    check_launch("fp8 Gram kernel");
}

void select_topk_rerank(const SelectArgs& args, hipStream_t stream)
{
    if (args.nq == 0) return;
    select_topk_rerank_kernel<<<args.nq, kSelThreads, 0, stream>>>(args);
    check_launch("shortlist re-rank kernel");
}

uint32_t gram_fp8_k_padding() { return kBK8; }

uint32_t gram_k_padding() { return kBK; }
int gram_tile_count() { return 8; }
uint32_t gram_m_tile() { return 256; }

void to_i8x2_rows(const float* src, uint64_t ld_src, uint32_t rows, uint32_t dim, uint8_t* dst, uint64_t ld_dst,
                  I8Stats* stats, hipStream_t stream)
{
    if (rows == 0) return;
    if (ld_dst % kBKI8 != 0 || ld_dst / 2 < dim) throw std::invalid_argument("int8 rows need 128 ceil(dim / 64) bytes");
    const uint64_t pdim = ld_dst / 2;
    if (pdim > kI8MaxDim) throw std::length_error("int8 rows are limited to 65536 columns");
    // the fewest four-column groups per thread (of 1024) that cover the padded row
    if (pdim <= 4096) {
        to_i8x2_rows_kernel<1><<<rows, 1024, 0, stream>>>(src, ld_src, rows, dim, dst, ld_dst, stats);
    } else if (pdim <= 8192) {
        to_i8x2_rows_kernel<2><<<rows, 1024, 0, stream>>>(src, ld_src, rows, dim, dst, ld_dst, stats);
    } else if (pdim <= 16384) {
        to_i8x2_rows_kernel<4><<<rows, 1024, 0, stream>>>(src, ld_src, rows, dim, dst, ld_dst, stats);
    } else if (pdim <= 32768) {
        to_i8x2_rows_kernel<8><<<rows, 1024, 0, stream>>>(src, ld_src, rows, dim, dst, ld_dst, stats);
    } else {
        to_i8x2_rows_kernel<16><<<rows, 1024, 0, stream>>>(src, ld_src, rows, dim, dst, ld_dst, stats);
    }
    check_launch("int8 conversion kernel");
}

namespace {

// GRAM_INT8 block shapes: waves WM x WN, MFMA tiles per wave TM x TN.
//   0: 4x2 waves, 2x2 tiles (256 x 128, 8 waves: 128 accumulator registers per lane)
//   1: 4x4 waves, 2x1 tiles (256 x 128, 16 waves)
//   2: 2x4 waves, 2x2 tiles (128 x 256, 8 waves)
//   3: 4x4 waves, 1x2 tiles (128 x 256, 16 waves)
//   4: 2x2 waves, 2x2 tiles (128 x 128, 4 waves: two blocks per CU)
//   5: ping-pong 256 x 128 (gram_i8x2_pp_kernel; the 4x2 wave grid of shape 0)
//   6: 1x4 waves, 1x1 tiles (32 x 128, small batches: up to 32 queries)
//   7: 1x4 waves, 2x1 tiles (64 x 128, up to 64 queries)
// Small batches read every reference row once whatever the tile height, so
// a 256-row tile would spend its MFMAs on padding rows; the 32- and 64-row
// shapes leave the scan bound by the int8 planes' HBM traffic, with the K
// split spreading their few tiles over every CU.
// group_m: tile rows swept together (tile_origin). Each XCD runs 32
// consecutive tiles at once; four rows of 256 x 128 tiles touch
// 4 x 256 + 8 x 128 operand rows per K step, eight rows 8 x 256 + 4 x 128.
struct I8Shape {
    int wm, wn, tm, tn;
    uint32_t group_m;
};
constexpr I8Shape kI8Shapes[8] = {{4, 2, 2, 2, 4}, {4, 4, 2, 1, 8}, {2, 4, 2, 2, 8}, {4, 4, 1, 2, 8},
                                   {2, 2, 2, 2, 8}, {4, 2, 2, 2, 4}, {1, 4, 1, 1, 1}, {1, 4, 2, 1, 1}};
constexpr int kI8DefaultTile = 5;

// The default shape for a batch of m queries.
int i8_default_tile(uint32_t m) { return m <= 32 ? 6 : m <= 64 ? 7 : kI8DefaultTile; }

}  // namespace

uint64_t gram_i8_split_floats(uint32_t max_m, uint32_t n, uint32_t k_steps, uint32_t cus)
{
    // every batch size up to max_m with its default shape
    uint64_t need = 0;
    for (uint32_t m = 1; m <= max_m;) {
        const I8Shape sh = kI8Shapes[i8_default_tile(m)];
        const uint32_t bm = 32U * sh.tm * sh.wm;
        const uint32_t bn = 32U * sh.tn * sh.wn;
        const uint32_t tiles_n = (n + bn - 1) / bn;
        const GramSplit sp = gram_split_plan(((m + bm - 1) / bm) * tiles_n, k_steps, cus);
        need = std::max<uint64_t>(need, static_cast<uint64_t>(sp.tail) * sp.parts * 2U * bm * bn);
        m = m <= 32 ? 33 : m <= 64 ? 65 : m + bm;  // next tile count (or shape)
    }
    return need;
}

void gram_i8x2(const uint8_t* q, uint64_t ldq, uint32_t m, const uint8_t* v, uint64_t ldv, uint32_t n, uint32_t k,
               const I8Stats* qs, const I8Stats* vs, float* out, uint64_t ldo, int tile, float* partials,
               uint64_t partial_floats, hipStream_t stream)
{
    if (m == 0 || n == 0) return;
    if (k % kBKI8 != 0) throw std::invalid_argument("int8 Gram K must be a multiple of 128 bytes");
    if (k / 2 > kI8MaxDim) throw std::length_error("int8 Gram rows are limited to 65536 columns");
    if (ldq % 16 != 0 || ldv % 16 != 0) throw std::invalid_argument("int8 Gram rows must be 16-byte aligned");
    if (tile < -1 || tile >= gram_i8_tile_count()) throw std::invalid_argument("unknown int8 Gram tile");
    const int shape = tile < 0 ? i8_default_tile(m) : tile;
    const I8Shape sh = kI8Shapes[shape];
    const uint32_t bm = 32U * sh.tm * sh.wm;
    const uint32_t bn = 32U * sh.tn * sh.wn;
    const uint32_t tiles_m = (m + bm - 1) / bm;
    const uint32_t tiles_n = (n + bn - 1) / bn;
    const uint32_t group_m = sh.group_m;
    GramSplit sp;
    sp.full = tiles_m * tiles_n;
    if (tile < 0) sp = split_for(tiles_m * tiles_n, k / kBKI8, static_cast<uint64_t>(2U * bm * bn), partials, partial_floats);
    const dim3 grid(sp.full + sp.tail * sp.parts);
    const dim3 block(64 * sh.wm * sh.wn);
#define UHANN_I8_CASE(T, A, B, C, D)                                                                       \
    case T:                                                                                                \
        gram_i8x2_kernel<A, B, C, D><<<grid, block, 0, stream>>>(q, ldq, m, v, ldv, n, k, qs, vs, out, ldo,  \
                                                                 tiles_m, tiles_n, group_m, sp);           \
        if (sp.tail > 0) {                                                                                 \
            gram_i8_fixup_kernel<A, B, C, D><<<sp.tail, block, 0, stream>>>(sp, m, n, qs, vs, out, ldo,       \
                                                                            tiles_m, tiles_n, group_m);    \
        }                                                                                                  \
        break;
    switch (shape) {
        UHANN_I8_CASE(0, 4, 2, 2, 2)
        UHANN_I8_CASE(1, 4, 4, 2, 1)
        UHANN_I8_CASE(2, 2, 4, 2, 2)
        UHANN_I8_CASE(3, 4, 4, 1, 2)
        UHANN_I8_CASE(4, 2, 2, 2, 2)
        UHANN_I8_CASE(6, 1, 4, 1, 1)
        UHANN_I8_CASE(7, 1, 4, 2, 1)
    case 5:
        gram_i8x2_pp_kernel<<<grid, block, 0, stream>>>(q, ldq, m, v, ldv, n, k, qs, vs, out, ldo, tiles_m, tiles_n,
                                                        group_m, sp);
        if (sp.tail > 0) {
            gram_i8_fixup_kernel<4, 2, 2, 2><<<sp.tail, block, 0, stream>>>(sp, m, n, qs, vs, out, ldo, tiles_m,
                                                                            tiles_n, group_m);
        }
        break;
    default:
        break;
    }
#undef UHANN_I8_CASE
    check_launch("int8 Gram kernel");
}

void select_rerank_i8(const SelectArgs& args, hipStream_t stream)
{
    if (args.nq == 0) return;
    if (args.qstats8 == nullptr || args.vstats8 == nullptr) throw std::invalid_argument("int8 selection needs row statistics");
    select_rerank_i8_kernel<<<args.nq, kSelThreads, 0, stream>>>(args);
    check_launch("int8 candidate selection / re-rank kernel");
}

uint32_t gram_i8_k_padding() { return kBKI8; }
uint32_t gram_i8_max_dim() { return kI8MaxDim; }
int gram_i8_tile_count() { return 8; }

}  // namespace uhann::kern
