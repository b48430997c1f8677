// uhann_oracle.cpp -- TEST INFRASTRUCTURE, NOT PRODUCT CODE.
//
// A plain CPU restatement of the reference's query-path arithmetic
// (martingherold/ultrahigh_ann). It is the checker for the HIP engine: only
// the tests, the smoke check and the CPU-baseline leg of the benchmark load
// it. Nothing in ultrahigh_ann_amd/ links, loads or calls it.
//
// Every function names the reference lines it restates. Two families:
//   * FP64 "CPU index" arithmetic: the reference's own CPU classes, which
//     accumulate in binary64 and keep the first minimum (strict <).
//   * FP32 emulators of the reference's CUDA kernels. Each CUDA thread t of a
//     256-thread block sums columns t, t+256, ... sequentially, then a 32-lane
//     shfl_down tree, then warp 0 reduces the 8 warp sums with the same tree
//     (lanes 8..31 contribute +0.0f). IEEE binary32 fmaf/add are
//     deterministic, so replaying that order reproduces the CUDA values bit
//     for bit without a GPU.
//
// Build: oracle/Makefile (g++ -O3 -ffp-contract=off -fopenmp). Every entry
// point is extern "C" with plain pointers so ctypes can call it.

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <cstring>
#include <limits>
#include <random>
#include <vector>

#include <omp.h>

namespace {

constexpr int kThreads = 256;  // cuda_dense_l2_scan_index.cu:22
constexpr int kWarp = 32;      // cuda_dense_l2_scan_index.cu:23
constexpr int kWarps = kThreads / kWarp;

// cuda_dense_l2_scan_index.cu:196-202 warp_sum: shfl_down offsets 16..1.
// Lane i < off becomes v[i] + v[i+off]; higher lanes never reach lane 0.
float warp_tree(float* v)
{
    for (int off = kWarp / 2; off != 0; off /= 2) {
        for (int i = 0; i < off; ++i) {
            v[i] = v[i] + v[i + off];
        }
    }
    return v[0];
}

// cuda_dense_l2_scan_index.cu:204-219 block_sum over 256 per-thread values.
float block_tree(const float* chains)
{
    float warp_sums[kWarps];
    for (int w = 0; w < kWarps; ++w) {
        float v[kWarp];
        std::memcpy(v, chains + w * kWarp, sizeof(v));
        warp_sums[w] = warp_tree(v);
    }
    float v[kWarp];
    for (int i = 0; i < kWarp; ++i) {
        v[i] = i < kWarps ? warp_sums[i] : 0.0F;  // :216
    }
    return warp_tree(v);
}

// K2 compute_dense_l2_distances, cuda_dense_l2_scan_index.cu:295-324.
float k2_pair(const float* ref, const float* q, std::size_t dim)
{
    float chains[kThreads];
    for (int t = 0; t < kThreads; ++t) {
        float acc = 0.0F;
        for (std::size_t c = static_cast<std::size_t>(t); c < dim; c += kThreads) {
            const float diff = ref[c] - q[c];
            acc = std::fmaf(diff, diff, acc);
        }
        chains[t] = acc;
    }
    return block_tree(chains);
}

// K5 compute_dense_l1_distances, cuda_dense_l1_scan_index.cu:218-247.
// weight_mode 0: w*|d| + acc contracted to fmaf (nvcc --fmad=true default);
// weight_mode 1: rounded product then add. Unweighted is exact either way.
float k5_pair(const float* ref, const float* q, const float* w, std::size_t dim,
              int weight_mode)
{
    float chains[kThreads];
    for (int t = 0; t < kThreads; ++t) {
        float acc = 0.0F;
        for (std::size_t c = static_cast<std::size_t>(t); c < dim; c += kThreads) {
            const float diff = std::fabs(ref[c] - q[c]);
            if (w == nullptr) {
                acc = acc + diff;
            } else if (weight_mode == 0) {
                acc = std::fmaf(w[c], diff, acc);
            } else {
                const float prod = w[c] * diff;
                acc = acc + prod;
            }
        }
        chains[t] = acc;
    }
    return block_tree(chains);
}

// l2_distance.cpp:9-25 squared_l2_dist (FP32 inputs, binary64 sequential).
double sq_l2_fp64(const float* a, const float* b, std::size_t n)
{
    double s = 0.0;
    for (std::size_t i = 0; i < n; ++i) {
        const double d = static_cast<double>(a[i]) - static_cast<double>(b[i]);
        s += d * d;
    }
    return s;
}

// Same loop over binary64 operands (hierarchical_l2_ann_index.cpp:169-183
// calls squared_l2_dist on projected double rows).
double sq_l2_fp64d(const double* a, const double* b, std::size_t n)
{
    double s = 0.0;
    for (std::size_t i = 0; i < n; ++i) {
        const double d = a[i] - b[i];
        s += d * d;
    }
    return s;
}

// l1_distance.cpp:12-28 uses std::transform_reduce(first1,last1,first2,0.0,
// plus, |a-b|). libstdc++ (GCC 11, <numeric>:372-395) groups random-access
// ranges by four: init += ((t0+t1)+(t2+t3)), then the tail sequentially.
double l1_fp64(const float* a, const float* b, std::size_t n)
{
    auto t = [&](std::size_t i) {
        return std::abs(static_cast<double>(a[i]) - static_cast<double>(b[i]));
    };
    double init = 0.0;
    std::size_t i = 0;
    for (; n - i >= 4; i += 4) {
        const double v1 = t(i) + t(i + 1);
        const double v2 = t(i + 2) + t(i + 3);
        const double v3 = v1 + v2;
        init = init + v3;
    }
    for (; i < n; ++i) {
        init = init + t(i);
    }
    return init;
}

// hierarchical_l1_ann_index.cpp:65-99 median_absolute_difference (binary64).
double median_abs_diff_fp64(const double* a, const double* b, std::size_t n,
                            std::vector<double>& scratch)
{
    scratch.resize(n);
    for (std::size_t i = 0; i < n; ++i) {
        scratch[i] = std::abs(a[i] - b[i]);
    }
    auto upper = scratch.begin() + static_cast<std::ptrdiff_t>(n / 2);
    std::nth_element(scratch.begin(), upper, scratch.end());
    if (n % 2 != 0) {
        return *upper;
    }
    const auto lower = std::max_element(scratch.begin(), upper);
    return *lower + (*upper - *lower) * 0.5;
}

}  // namespace

extern "C" {

// ---------------------------------------------------------------- FP64 scans

// dense_l2_scan_index.cpp:25-41 nearest_representative (strict <: first
// minimum wins) over l2_distance.cpp:9-25. Queries are independent, so
// splitting them over OpenMP threads does not change any result.
void orc_l2_fp64_scan(const float* refs, uint64_t r, uint64_t dim,
                      const float* queries, uint64_t b, int threads,
                      uint64_t* out_idx, double* out_dist)
{
#pragma omp parallel for schedule(dynamic, 1) num_threads(threads > 0 ? threads : 1)
    for (int64_t qi = 0; qi < static_cast<int64_t>(b); ++qi) {
        const float* q = queries + static_cast<std::size_t>(qi) * dim;
        std::size_t best = 0;
        double best_d = sq_l2_fp64(q, refs, dim);
        for (std::size_t row = 1; row < r; ++row) {
            const double d = sq_l2_fp64(q, refs + row * dim, dim);
            if (d < best_d) {
                best_d = d;
                best = row;
            }
        }
        out_idx[qi] = best;
        if (out_dist != nullptr) out_dist[qi] = best_d;
    }
}

// All B x r binary64 squared distances (used for agreement/excess metrics,
// benchmark_metrics.cpp:112-300 builds the same table).
void orc_l2_fp64_table(const float* refs, uint64_t r, uint64_t dim,
                       const float* queries, uint64_t b, int threads,
                       double* out)
{
#pragma omp parallel for schedule(dynamic, 1) num_threads(threads > 0 ? threads : 1)
    for (int64_t qi = 0; qi < static_cast<int64_t>(b); ++qi) {
        const float* q = queries + static_cast<std::size_t>(qi) * dim;
        for (std::size_t row = 0; row < r; ++row) {
            out[static_cast<std::size_t>(qi) * r + row] =
                sq_l2_fp64(q, refs + row * dim, dim);
        }
    }
}

// exact_l1_index.cpp:26-57 ExactL1Index::query over l1_distance.cpp:12-28.
void orc_l1_fp64_scan(const float* refs, uint64_t r, uint64_t dim,
                      const float* queries, uint64_t b, int threads,
                      uint64_t* out_idx, double* out_dist)
{
#pragma omp parallel for schedule(dynamic, 1) num_threads(threads > 0 ? threads : 1)
    for (int64_t qi = 0; qi < static_cast<int64_t>(b); ++qi) {
        const float* q = queries + static_cast<std::size_t>(qi) * dim;
        std::size_t best = 0;
        double best_d = l1_fp64(q, refs, dim);
        for (std::size_t row = 1; row < r; ++row) {
            const double d = l1_fp64(q, refs + row * dim, dim);
            if (d < best_d) {
                best_d = d;
                best = row;
            }
        }
        out_idx[qi] = best;
        if (out_dist != nullptr) out_dist[qi] = best_d;
    }
}

void orc_l1_fp64_table(const float* refs, uint64_t r, uint64_t dim,
                       const float* queries, uint64_t b, int threads,
                       double* out)
{
#pragma omp parallel for schedule(dynamic, 1) num_threads(threads > 0 ? threads : 1)
    for (int64_t qi = 0; qi < static_cast<int64_t>(b); ++qi) {
        const float* q = queries + static_cast<std::size_t>(qi) * dim;
        for (std::size_t row = 0; row < r; ++row) {
            out[static_cast<std::size_t>(qi) * r + row] =
                l1_fp64(q, refs + row * dim, dim);
        }
    }
}

// sampled_coordinate_l1_ann_index.cpp:63-83 weighted_l1_distance and
// :119-154 query: the query is read at source_column, the weight is
// (double)m * inverse_probability, binary64 sequential, r == 1 -> 0.
void orc_weighted_l1_fp64_scan(const float* sampled_refs, uint64_t r,
                               uint64_t k, const float* queries, uint64_t dim,
                               const uint64_t* cols, const uint64_t* mult,
                               const double* invp, uint64_t b, int threads,
                               uint64_t* out_idx, double* out_dist)
{
#pragma omp parallel for schedule(dynamic, 1) num_threads(threads > 0 ? threads : 1)
    for (int64_t qi = 0; qi < static_cast<int64_t>(b); ++qi) {
        const float* q = queries + static_cast<std::size_t>(qi) * dim;
        auto dist = [&](std::size_t row) {
            const float* s = sampled_refs + row * k;
            double acc = 0.0;
            for (std::size_t i = 0; i < k; ++i) {
                const double diff = std::abs(static_cast<double>(q[cols[i]]) -
                                             static_cast<double>(s[i]));
                acc += static_cast<double>(mult[i]) * invp[i] * diff;
            }
            return acc;
        };
        std::size_t best = 0;
        double best_d = dist(0);
        if (r > 1) {
            for (std::size_t row = 1; row < r; ++row) {
                const double d = dist(row);
                if (d < best_d) {
                    best_d = d;
                    best = row;
                }
            }
        }
        out_idx[qi] = best;
        if (out_dist != nullptr) out_dist[qi] = best_d;
    }
}

// ------------------------------------------------------- FP32 CUDA emulators

// K2 for every (query, reference) pair; out is B x r row-major, as the
// reference's distance workspace (cuda_dense_l2_scan_index.cu:300-323).
void orc_k2_l2_distances(const float* refs, uint64_t r, uint64_t dim,
                         const float* queries, uint64_t b, int threads,
                         float* out)
{
#pragma omp parallel for collapse(2) schedule(dynamic, 16) num_threads(threads > 0 ? threads : 1)
    for (int64_t qi = 0; qi < static_cast<int64_t>(b); ++qi) {
        for (int64_t row = 0; row < static_cast<int64_t>(r); ++row) {
            out[static_cast<std::size_t>(qi) * r + static_cast<std::size_t>(row)] =
                k2_pair(refs + static_cast<std::size_t>(row) * dim,
                        queries + static_cast<std::size_t>(qi) * dim, dim);
        }
    }
}

// K5 (L1, optional per-coordinate weights).
void orc_k5_l1_distances(const float* refs, uint64_t r, uint64_t dim,
                         const float* queries, uint64_t b,
                         const float* weights, int weight_mode, int threads,
                         float* out)
{
#pragma omp parallel for collapse(2) schedule(dynamic, 16) num_threads(threads > 0 ? threads : 1)
    for (int64_t qi = 0; qi < static_cast<int64_t>(b); ++qi) {
        for (int64_t row = 0; row < static_cast<int64_t>(r); ++row) {
            out[static_cast<std::size_t>(qi) * r + static_cast<std::size_t>(row)] =
                k5_pair(refs + static_cast<std::size_t>(row) * dim,
                        queries + static_cast<std::size_t>(qi) * dim, weights,
                        dim, weight_mode);
        }
    }
}

// K3/K6/K10 argmin: candidate_is_better (cuda_dense_l2_scan_index.cu:221-229)
// keeps the smaller distance and, on equality, the lower index; NaN never
// wins. The reduction order cannot change that result.
void orc_fp32_argmin(const float* dist, uint64_t b, uint64_t r,
                     uint64_t* out_idx, float* out_dist)
{
    for (std::size_t qi = 0; qi < b; ++qi) {
        float best = std::numeric_limits<float>::infinity();
        uint64_t best_i = std::numeric_limits<uint64_t>::max();
        for (std::size_t row = 0; row < r; ++row) {
            const float d = dist[qi * r + row];
            if (d < best || (d == best && row < best_i)) {
                best = d;
                best_i = row;
            }
        }
        out_idx[qi] = best_i;
        if (out_dist != nullptr) out_dist[qi] = best;
    }
}

// K7 project_queries_direct, cuda_float_projection.cu:234-263:
// y[q][k] = block_sum of fmaf(x[q][c], P[k][c], acc) chains.
void orc_k7_project(const float* x, uint64_t b, uint64_t kdim, const float* proj,
                    uint64_t p, float* out, int* invalid)
{
    int bad = 0;
    for (std::size_t qi = 0; qi < b; ++qi) {
        for (std::size_t k = 0; k < p; ++k) {
            float chains[kThreads];
            for (int t = 0; t < kThreads; ++t) {
                float acc = 0.0F;
                for (std::size_t c = static_cast<std::size_t>(t); c < kdim; c += kThreads) {
                    acc = std::fmaf(x[qi * kdim + c], proj[k * kdim + c], acc);
                }
                chains[t] = acc;
            }
            const float y = block_tree(chains);
            if (!std::isfinite(y)) bad = 1;
            out[qi * p + k] = y;
        }
    }
    if (invalid != nullptr) *invalid = bad;
}

// K9 compute_median_absolute_differences,
// cuda_median_l1_scan_index.cu:202-267: |r-q| in binary32 (non-finite ->
// flag, +inf), sorted ascending, median = s[p/2] or
// lower + (upper - lower) * 0.5f. The sorted multiset does not depend on the
// sorting network, so std::sort reproduces the bitonic result.
void orc_k9_median_l1(const float* refs, uint64_t r, uint64_t p,
                      const float* queries, uint64_t b, float* out,
                      int* invalid)
{
    int bad = 0;
    std::vector<float> s(p);
    for (std::size_t qi = 0; qi < b; ++qi) {
        for (std::size_t row = 0; row < r; ++row) {
            for (std::size_t k = 0; k < p; ++k) {
                float diff = std::fabs(refs[row * p + k] - queries[qi * p + k]);
                if (!std::isfinite(diff)) {
                    bad = 1;
                    diff = std::numeric_limits<float>::infinity();
                }
                s[k] = diff;
            }
            std::sort(s.begin(), s.end());
            const std::size_t mid = p / 2;
            float med = s[mid];
            if (p % 2 == 0) {
                const float lower = s[mid - 1];
                med = lower + (med - lower) * 0.5F;
            }
            if (!std::isfinite(med)) bad = 1;
            out[qi * r + row] = med;
        }
    }
    if (invalid != nullptr) *invalid = bad;
}

// ----------------------------------------------------- index construction

// coordinate_sampling.cpp:14-45 build_coordinate_sample. One fresh
// std::binomial_distribution<size_t>(T, p_j) draw per column with p_j > 0;
// p_j == 0 or T == 0 consumes nothing. Returns the number of kept columns,
// or -1 for an invalid probability. The engine is mt19937_64(seed) unless
// engine_state (text from operator<<) is given; the final state is written
// back to engine_state_out when non-null (buffer of engine_state_cap bytes).
static int64_t sample_with_engine(const double* probs, uint64_t dim, uint64_t reps,
                                  std::mt19937_64& eng, uint64_t* cols,
                                  uint64_t* mult, double* invp)
{
    int64_t n = 0;
    for (std::size_t c = 0; c < dim; ++c) {
        const double p = probs[c];
        if (!std::isfinite(p) || p < 0.0 || p > 1.0) return -1;
        if (p == 0.0 || reps == 0) continue;
        std::binomial_distribution<std::size_t> dist(reps, p);
        const std::size_t m = dist(eng);
        if (m != 0) {
            cols[n] = c;
            mult[n] = m;
            invp[n] = 1.0 / p;
            ++n;
        }
    }
    return n;
}

int64_t orc_build_coordinate_sample(const double* probs, uint64_t dim,
                                    uint64_t reps, uint64_t seed,
                                    uint64_t* cols, uint64_t* mult,
                                    double* invp)
{
    std::mt19937_64 eng(seed);
    return sample_with_engine(probs, dim, reps, eng, cols, mult, invp);
}

// l2_coordinate_embedding.cpp:29-88: scale_j = (float)sqrt((double)m_j *
// (1/p_j)); l1_coordinate_embedding.cpp:29-79: weight_j = (float)(m_j/p_j).
void orc_embedding_factors(const uint64_t* mult, const double* invp, uint64_t k,
                           float* l2_scales, float* l1_weights)
{
    for (std::size_t i = 0; i < k; ++i) {
        const double w = static_cast<double>(mult[i]) * invp[i];
        if (l2_scales != nullptr) l2_scales[i] = static_cast<float>(std::sqrt(w));
        if (l1_weights != nullptr) l1_weights[i] = static_cast<float>(w);
    }
}

// Hierarchical build, reference order: sample (above), then the projection
// matrix from the same engine, row-major (row outer, sampled column inner).
//   L2 (l2_hierarchical_projection.cpp:53-77): fresh
//     binomial_distribution<size_t>(m_j, 0.5) per entry,
//     coef = (2*pos - m_j) * sqrt(1/p_j) * (1/sqrt(p)).
//   L1 (l1_hierarchical_projection.cpp:54-77): one cauchy_distribution<double>
//     (0,1), coef = c * (1/p_j) * m_j.
// Outputs: sample arrays (capacity dim), P64 (p x k), P32 = (float)P64,
// projected references in binary64 (DenseMatrix::multiply_right_transposed,
// dense_matrix.hpp:61-121) and in binary32 (multiply_right_transposed_fp32,
// fp32_matrix_operations.hpp:29-67: float product then float add).
// Returns k, or -1 (bad probability), -2 (empty sample with r > 1,
// l2_hierarchical_projection.cpp:98-100), -3 (non-finite coefficient).
int64_t orc_build_hierarchical(int metric_l1, const float* refs, uint64_t r,
                               uint64_t dim, const double* probs, uint64_t reps,
                               uint64_t p, uint64_t seed, uint64_t* cols,
                               uint64_t* mult, double* invp, double* p64,
                               float* p32, double* proj_refs64,
                               float* proj_refs32)
{
    std::mt19937_64 eng(seed);
    const int64_t k = sample_with_engine(probs, dim, reps, eng, cols, mult, invp);
    if (k < 0) return -1;
    if (k == 0 && r > 1) return -2;
    const std::size_t kk = static_cast<std::size_t>(k);
    if (metric_l1 == 0) {
        const double row_scaling = 1.0 / std::sqrt(static_cast<double>(p));
        for (std::size_t row = 0; row < p; ++row) {
            for (std::size_t j = 0; j < kk; ++j) {
                std::binomial_distribution<std::size_t> signs(mult[j], 0.5);
                const std::size_t pos = signs(eng);
                const double signed_sum = 2.0 * static_cast<double>(pos) -
                                          static_cast<double>(mult[j]);
                const double coef = signed_sum * std::sqrt(invp[j]) * row_scaling;
                if (!std::isfinite(coef)) return -3;
                p64[row * kk + j] = coef;
            }
        }
    } else {
        std::cauchy_distribution<double> cauchy(0.0, 1.0);
        for (std::size_t row = 0; row < p; ++row) {
            for (std::size_t j = 0; j < kk; ++j) {
                const double coef =
                    cauchy(eng) * invp[j] * static_cast<double>(mult[j]);
                if (!std::isfinite(coef)) return -3;
                p64[row * kk + j] = coef;
            }
        }
    }
    for (std::size_t i = 0; i < p * kk; ++i) p32[i] = static_cast<float>(p64[i]);
    std::vector<float> sampled(kk);
    for (std::size_t row = 0; row < r; ++row) {
        for (std::size_t j = 0; j < kk; ++j) sampled[j] = refs[row * dim + cols[j]];
        for (std::size_t o = 0; o < p; ++o) {
            double s64 = 0.0;
            float s32 = 0.0F;
            for (std::size_t j = 0; j < kk; ++j) {
                s64 += static_cast<double>(sampled[j]) * p64[o * kk + j];
                const float prod = sampled[j] * p32[o * kk + j];
                s32 = s32 + prod;
            }
            proj_refs64[row * p + o] = s64;
            proj_refs32[row * p + o] = s32;
        }
    }
    return k;
}

// project_l2_hierarchical_query (l2_hierarchical_projection.cpp:164-200) and
// its L1 twin (l1_hierarchical_projection.cpp:171-210): binary64 sum of
// (double)x_j * P64[o][j], then the binary64 scan of
// hierarchical_l2_ann_index.cpp:152-188 (squared L2) or
// hierarchical_l1_ann_index.cpp:189-236 (median |.|). r == 1 -> 0.
void orc_hierarchical_query_fp64(int metric_l1, const double* proj_refs64,
                                 uint64_t r, uint64_t p, const double* p64,
                                 uint64_t k, const uint64_t* cols,
                                 const float* queries, uint64_t dim, uint64_t b,
                                 int threads, uint64_t* out_idx,
                                 double* out_dist)
{
#pragma omp parallel for schedule(dynamic, 1) num_threads(threads > 0 ? threads : 1)
    for (int64_t qi = 0; qi < static_cast<int64_t>(b); ++qi) {
        const float* q = queries + static_cast<std::size_t>(qi) * dim;
        std::vector<double> y(p);
        std::vector<double> scratch;
        for (std::size_t o = 0; o < p; ++o) {
            double s = 0.0;
            for (std::size_t j = 0; j < k; ++j) {
                s += static_cast<double>(q[cols[j]]) * p64[o * k + j];
            }
            y[o] = s;
        }
        if (r == 1) {
            out_idx[qi] = 0;
            if (out_dist != nullptr) out_dist[qi] = 0.0;
            continue;
        }
        auto dist = [&](std::size_t row) {
            return metric_l1 ? median_abs_diff_fp64(proj_refs64 + row * p, y.data(), p, scratch)
                             : sq_l2_fp64d(proj_refs64 + row * p, y.data(), p);
        };
        std::size_t best = 0;
        double best_d = dist(0);
        for (std::size_t row = 1; row < r; ++row) {
            const double d = dist(row);
            if (d < best_d) {
                best_d = d;
                best = row;
            }
        }
        out_idx[qi] = best;
        if (out_dist != nullptr) out_dist[qi] = best_d;
    }
}

// Importance probabilities (sequential policy).
// L2, importance_sampling.cpp:18-48: p_j = max over pairs a<b with
// ||a-b||^2 != 0 of min(1, (a_j-b_j)^2 * (1/||a-b||^2)).
// L1, l1/importance_sampling.cpp:19-47: p_j = max of |a_j-b_j| * (1/||a-b||_1).
void orc_importance_probabilities(int metric_l1, const float* refs, uint64_t r,
                                  uint64_t dim, double* out)
{
    for (std::size_t c = 0; c < dim; ++c) out[c] = 0.0;
    for (std::size_t a = 0; a < r; ++a) {
        const float* ra = refs + a * dim;
        for (std::size_t bb = a + 1; bb < r; ++bb) {
            const float* rb = refs + bb * dim;
            const double dist = metric_l1 ? l1_fp64(ra, rb, dim) : sq_l2_fp64(ra, rb, dim);
            if (dist == 0.0) continue;
            const double inv = 1.0 / dist;
            for (std::size_t c = 0; c < dim; ++c) {
                const double diff = static_cast<double>(ra[c]) - static_cast<double>(rb[c]);
                const double prob = metric_l1 ? std::abs(diff) * inv
                                              : std::min(1.0, diff * diff * inv);
                out[c] = std::max(out[c], prob);
            }
        }
    }
}

int orc_max_threads(void)
{
    return omp_get_max_threads();
}

}  // extern "C"
