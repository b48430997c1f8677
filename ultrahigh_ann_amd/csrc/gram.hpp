// Launch interfaces of the mixed-precision kernels in gram.hip: BF16 and
// per-row-scaled FP8 Gram estimates, the certified BF16 candidate re-rank and
// the FP8 top-k shortlist re-rank.
#pragma once

#include <hip/hip_runtime_api.h>

#include <cstdint>

namespace uhann::kern {

// Per-row statistics used by the Gram estimate and its error bound; x^ is
// the row as the Gram kernel sees it (bf16, or scale * e4m3).
struct RowStats {
    float norm2;   // ||x||^2 of the binary32 row (binary64 sum, rounded)
    float l4;      // ||x||_4
    float l4_hat;  // ||x^||_4
    float l2_hat;  // ||x^||_2
    float d4;      // ||x - x^||_4, the row's actual rounding error
    float scale;   // FP8: power-of-two row scale (x^ = scale * e4m3); BF16: 1
    float pad[2];
};

// GRAM_INT8 per-row statistics. A row x is held as 16-bit fixed point
// x^ = s (a + b / 256) with int8 planes a (in [-127, 127]) and b (in
// [-128, 127]) and s = max|x| / 127; every norm is of the actual values and
// rounded up, so the error bound built from them is deterministic.
struct I8Stats {
    double norm2;  // ||x||^2 of the binary32 row (binary64 sum)
    float scale;   // s (1 for a zero row, NaN for a non-finite row)
    float err;     // ||x - x^||_2
    float hat;     // ||x^||_2
    float lo;      // s ||b||_2 / 256
    float norm;    // ||x||_2
    float pad;
};

struct MixedStats {
    unsigned int candidates;     // sum over queries of the candidate-set size
    unsigned int fp64_reranked;  // queries decided in binary64
    unsigned int escalated;      // queries sent to the exact direct scan
    unsigned int certified;      // FP8 shortlist: queries whose answer the error bound proves
};

// Split of a Gram grid's last, partial round of tiles along K. A grid of
// `tiles` output tiles on C compute units (one block each) runs in
// ceil(tiles / C) rounds; the tail tiles = tiles mod C of the last round
// instead run as `parts` K ranges each, so the units fill the machine, and a
// fix-up kernel sums each tile's parts in part order. Whole tiles keep the
// sequential K accumulation; split tiles add their parts' sums.
struct GramSplit {
    uint32_t full = 0;          // whole tiles: the first `full` tiles of the grouped order
    uint32_t tail = 0;          // tiles split along K
    uint32_t parts = 1;         // K parts per tail tile
    uint32_t part_steps = 0;    // K steps per part (the last part may have fewer)
    float* partials = nullptr;  // tail * parts block tiles of binary32 partial sums
};

constexpr int kMaxCandidates = 64;
constexpr int kShortlist = 16;  // FP8 strategy: candidates re-ranked per query

struct SelectArgs {
    uint32_t nq = 0;
    uint32_t rows = 0;
    uint32_t dim = 0;
    const float* approx = nullptr;  // nq x rows, leading dimension ld_approx
    uint64_t ld_approx = 0;
    const RowStats* qstats = nullptr;
    const RowStats* vstats = nullptr;
    float acc_coef = 0.0F;          // accumulation term: acc_coef u_32 ||q^||_2 ||v^||_2
    const float* queries = nullptr; // binary32 queries, row stride ld_q
    uint64_t ld_q = 0;
    const float* refs = nullptr;    // binary32 references, row stride ld_ref
    uint64_t ld_ref = 0;
    double fp32_band = 1.0;         // (1+gamma)/(1-gamma) of the binary32 scan
    const I8Stats* qstats8 = nullptr;  // GRAM_INT8 row statistics (instead of qstats / vstats)
    const I8Stats* vstats8 = nullptr;
    uint64_t* out_idx = nullptr;
    float* out_dist = nullptr;
    unsigned int* escalate_count = nullptr;
    unsigned int* escalate_list = nullptr;  // capacity nq
    MixedStats* stats = nullptr;
    uint32_t shortlist = kShortlist;  // FP8 strategy: rows re-ranked per query (1 .. kShortlist)
};

void to_bf16_rows(const float* src, uint64_t ld_src, uint32_t rows, uint32_t dim, __bf16* dst,
                  uint64_t ld_dst, RowStats* stats, hipStream_t stream);
// tile selects the block shape (0 .. gram_tile_count() - 1; block tiles of
// at most gram_m_tile() x gram_m_tile() outputs). q must hold
// ceil(m / gram_m_tile()) * gram_m_tile() readable rows and v likewise for n:
// the kernel reads whole tiles, rows past m / n only feed outputs that are
// not written.
// partials (or null: every tile whole) is scratch for the tail split of
// up to partial_floats floats; a split that needs more is not used.
void gram_bf16(const __bf16* q, uint64_t ldq, uint32_t m, const __bf16* v, uint64_t ldv, uint32_t n,
               uint32_t k, const RowStats* qs, const RowStats* vs, float* out, uint64_t ldo,
               int tile, float* partials, uint64_t partial_floats, hipStream_t stream);
void select_rerank(const SelectArgs& args, hipStream_t stream);

// binary32 rows -> e4m3 bytes (round to nearest even) with a power-of-two
// row scale 2^ceil(log2(max|x| / 448)) (1 for a zero row), columns padded
// with zeros to ld_dst bytes, plus row statistics of scale * e4m3.
void to_fp8_rows(const float* src, uint64_t ld_src, uint32_t rows, uint32_t dim, uint8_t* dst,
                 uint64_t ld_dst, RowStats* stats, hipStream_t stream);
// out[m][n] = ||q||^2 + ||v||^2 - 2 s_q s_v <q8, v8> with the inner product
// on v_mfma_scale_f32_32x32x64_f8f6f4 (binary32 accumulation); k a multiple
// of gram_fp8_k_padding(), q and v padded to whole tiles like gram_bf16.
void gram_fp8(const uint8_t* q, uint64_t ldq, uint32_t m, const uint8_t* v, uint64_t ldv, uint32_t n, uint32_t k,
              const RowStats* qs, const RowStats* vs, float* out, uint64_t ldo, float* partials,
              uint64_t partial_floats, hipStream_t stream);
// The K-split plan for `tiles` block tiles of ksteps K steps on `cus`
// compute units (full == tiles: no split).
GramSplit gram_split_plan(uint32_t tiles, uint32_t ksteps, uint32_t cus);
// Scratch floats that the split of any batch of up to max_m queries against
// n references needs with the 256 x 256 block tiles (k_steps K steps).
uint64_t gram_split_floats(uint32_t max_m, uint32_t n, uint32_t k_steps, uint32_t cus);
uint32_t gram_block_outputs();  // outputs per 256 x 256 block tile
// Per query: the kShortlist rows with the smallest estimates (ties to the
// lower row), re-ranked exactly in the binary32 reference order, binary64
// for near ties; stats->certified counts queries where every other row's
// estimate minus its error bound exceeds the answer's distance.
void select_topk_rerank(const SelectArgs& args, hipStream_t stream);
uint32_t gram_fp8_k_padding();
uint32_t gram_k_padding();

// binary32 rows -> 16-bit fixed point in two int8 planes, interleaved per 64
// columns: bytes [128 c, 128 c + 64) of a row hold a of columns 64 c .. 64 c
// + 63 and the next 64 bytes their b; ld_dst = 128 ceil(dim / 64) bytes,
// columns past dim are zero. n = rint(x (32512 / max|x|)) (binary32
// division, then binary32 product), clamped to [-32512, 32512]; a = floor((n
// + 128) / 256), b = n - 256 a; s = max|x| / 127 (binary32).
void to_i8x2_rows(const float* src, uint64_t ld_src, uint32_t rows, uint32_t dim, uint8_t* dst, uint64_t ld_dst,
                  I8Stats* stats, hipStream_t stream);
// out[m][n] = binary32 of ||q||^2 + ||v||^2 - 2 s_q s_v (H + X / 256) with
// H = <a_q, a_v> and X = <a_q, b_v> + <b_q, a_v> summed exactly in int32 on
// v_mfma_i32_32x32x32_i8, combined in binary64. k (bytes) a multiple of 128,
// dim <= gram_i8_max_dim(); q and v padded to whole tiles (gram_m_tile() rows).
// tile -1: the default shape with the K split of a partial last round;
// 0 .. gram_i8_tile_count() - 1 force a shape with every tile whole.
void gram_i8x2(const uint8_t* q, uint64_t ldq, uint32_t m, const uint8_t* v, uint64_t ldv, uint32_t n, uint32_t k,
               const I8Stats* qs, const I8Stats* vs, float* out, uint64_t ldo, int tile, float* partials,
               uint64_t partial_floats, hipStream_t stream);
// Scratch floats the GRAM_INT8 K split needs for batches up to max_m.
uint64_t gram_i8_split_floats(uint32_t max_m, uint32_t n, uint32_t k_steps, uint32_t cus);
uint32_t gram_i8_k_padding();  // bytes per K step (64 columns, both planes)
uint32_t gram_i8_max_dim();    // largest dim whose int32 sums cannot overflow
int gram_i8_tile_count();
// Certified candidate selection for GRAM_INT8: the bound
// 2 (||q|| err_v + err_q ||v^|| + lo_q lo_v) + 2^-23 |est| + 1e-10 (||q||^2 +
// ||v||^2) on |est - ||q - v||^2| holds for every input; rows whose lower
// end exceeds U (1+g)^2/(1-g)^2 (U = min upper end, g the binary32 scan's
// relative error) cannot be the binary32 answer or in its near-tie band.
void select_rerank_i8(const SelectArgs& args, hipStream_t stream);
int gram_tile_count();
uint32_t gram_m_tile();

// Zero fill of the strategy counters (gather.hip).
void zero_u32(unsigned int* p, uint32_t n, hipStream_t stream);

}  // namespace uhann::kern
