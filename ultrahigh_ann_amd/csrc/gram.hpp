// Launch interfaces of the mixed-precision (BF16 Gram + exact re-rank)
// kernels in gram.hip.
#pragma once

#include <hip/hip_runtime_api.h>

#include <cstdint>

namespace uhann::kern {

// Per-row statistics used by the Gram estimate and its error bound.
struct RowStats {
    float norm2;   // ||x||^2 of the binary32 row (binary64 sum, rounded)
    float l4;      // ||x||_4
    float l4_hat;  // ||bf16(x)||_4
    float l2_hat;  // ||bf16(x)||_2
    float d4;      // ||x - bf16(x)||_4, the row's actual rounding error
    float pad[3];
};

struct MixedStats {
    unsigned int candidates;     // sum over queries of the candidate-set size
    unsigned int fp64_reranked;  // queries decided in binary64
    unsigned int escalated;      // queries sent to the exact direct scan
    unsigned int pad;
};

constexpr int kMaxCandidates = 64;

struct SelectArgs {
    uint32_t nq = 0;
    uint32_t rows = 0;
    uint32_t dim = 0;
    const float* approx = nullptr;  // nq x rows, leading dimension ld_approx
    uint64_t ld_approx = 0;
    const RowStats* qstats = nullptr;
    const RowStats* vstats = nullptr;
    float acc_steps_sqrt = 0.0F;    // sqrt(ceil(K/16))
    const float* queries = nullptr; // binary32 queries, row stride ld_q
    uint64_t ld_q = 0;
    const float* refs = nullptr;    // binary32 references, row stride ld_ref
    uint64_t ld_ref = 0;
    double fp32_band = 1.0;         // (1+gamma)/(1-gamma) of the binary32 scan
    uint64_t* out_idx = nullptr;
    float* out_dist = nullptr;
    unsigned int* escalate_count = nullptr;
    unsigned int* escalate_list = nullptr;  // capacity nq
    MixedStats* stats = nullptr;
};

void to_bf16_rows(const float* src, uint64_t ld_src, uint32_t rows, uint32_t dim, __bf16* dst,
                  uint64_t ld_dst, RowStats* stats, hipStream_t stream);
// tile selects the block shape (0 .. gram_tile_count() - 1; block tiles of
// at most gram_m_tile() x gram_m_tile() outputs). q must hold
// ceil(m / gram_m_tile()) * gram_m_tile() readable rows and v likewise for n:
// the kernel reads whole tiles, rows past m / n only feed outputs that are
// not written.
void gram_bf16(const __bf16* q, uint64_t ldq, uint32_t m, const __bf16* v, uint64_t ldv, uint32_t n,
               uint32_t k, const RowStats* qs, const RowStats* vs, float* out, uint64_t ldo,
               int tile, hipStream_t stream);
void select_rerank(const SelectArgs& args, hipStream_t stream);
uint32_t gram_k_padding();
int gram_tile_count();
uint32_t gram_m_tile();

// Small helpers for the escalation path (scan.hip / gather.hip).
void zero_u32(unsigned int* p, uint32_t n, hipStream_t stream);
void gather_rows(const float* src, uint64_t ld_src, const unsigned int* rows, uint32_t n, uint32_t dim,
                 float* dst, uint64_t ld_dst, hipStream_t stream);
void scatter_results(const uint64_t* idx, const float* dist, const unsigned int* rows, uint32_t n,
                     uint64_t* out_idx, float* out_dist, hipStream_t stream);

}  // namespace uhann::kern
