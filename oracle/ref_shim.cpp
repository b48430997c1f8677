// ref_shim.cpp -- TEST INFRASTRUCTURE, NOT PRODUCT CODE.
//
// extern "C" entry points over the reference's own CPU classes, compiled
// together with the reference's unmodified C++ sources into
// oracle/_ref/libuhann_ref.so (recipe: oracle/Makefile, target `ref`). It lets
// the tests pin the CPU restatement in uhann_oracle.cpp against the real
// reference arithmetic, and lets the benchmark time the reference CPU path.
// This file contains no reference code; it only calls the reference's public
// API and two of its build-time helpers (detail::build_*_hierarchical_projection*).

#include "ultrahigh_ann/coordinate_sampling/coordinate_sampling.hpp"
#include "ultrahigh_ann/coordinate_sampling/uniform_l1_ann_index.hpp"
#include "ultrahigh_ann/coordinate_sampling/uniform_l2_ann_index.hpp"
#include "ultrahigh_ann/coordinate_sampling/uniform_probabilities.hpp"
#include "ultrahigh_ann/core/dense_l2_scan_index.hpp"
#include "ultrahigh_ann/core/dense_matrix.hpp"
#include "ultrahigh_ann/exact/l1/exact_l1_index.hpp"
#include "ultrahigh_ann/exact/l2/exact_l2_index.hpp"
#include "ultrahigh_ann/hnsvw25/l1/flat/l1_ann_index.hpp"
#include "ultrahigh_ann/hnsvw25/l1/hierarchical/hierarchical_l1_ann_index.hpp"
#include "ultrahigh_ann/hnsvw25/l1/importance_sampling.hpp"
#include "ultrahigh_ann/hnsvw25/l2/flat/l2_ann_index.hpp"
#include "ultrahigh_ann/hnsvw25/l2/hierarchical/hierarchical_l2_ann_index.hpp"
#include "ultrahigh_ann/hnsvw25/l2/importance_sampling.hpp"
#include "ultrahigh_ann/io/importance_probability_io.hpp"
#include "ultrahigh_ann/io/sha256.hpp"

#include "benchmarks/benchmark_metrics.hpp"
#include "hnsvw25/l1/hierarchical/l1_hierarchical_projection.hpp"
#include "hnsvw25/l2/hierarchical/l2_hierarchical_projection.hpp"

#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <exception>
#include <chrono>
#include <random>
#include <span>
#include <vector>

namespace uh = ultrahigh_ann;

namespace {

uh::DenseMatrix make_matrix(const float* values, uint64_t rows, uint64_t cols)
{
    std::vector<float> data(values, values + rows * cols);
    return uh::DenseMatrix(std::move(data), rows, cols);
}

template <class F> int guarded(F&& body)
{
    try {
        body();
        return 0;
    } catch (const std::exception&) {
        return 1;
    }
}

namespace bm = ultrahigh_ann::benchmark;

constexpr double kNone = std::numeric_limits<double>::quiet_NaN();

double* put_summary(double* o, const bm::DistributionSummary& s)
{
    *o++ = static_cast<double>(s.count);
    *o++ = s.mean.value_or(kNone);
    *o++ = s.median.value_or(kNone);
    *o++ = s.percentile_95.value_or(kNone);
    *o++ = s.percentile_99.value_or(kNone);
    *o++ = s.maximum.value_or(kNone);
    return o;
}

// Five margin buckets of 6 values (lower, upper, queries, non-optimal,
// ratio-eligible, mean ratio); NaN where a bucket list is empty.
double* put_buckets(double* o, const std::vector<bm::MarginBucketMetrics>& buckets)
{
    for (std::size_t i = 0; i < 5; ++i) {
        if (i < buckets.size()) {
            const auto& b = buckets[i];
            *o++ = b.lower_inclusive;
            *o++ = b.upper_exclusive.value_or(kNone);
            *o++ = static_cast<double>(b.query_count);
            *o++ = static_cast<double>(b.non_optimal_count);
            *o++ = static_cast<double>(b.ratio_eligible_count);
            *o++ = b.mean_distance_ratio.value_or(kNone);
        } else {
            for (int k = 0; k < 6; ++k) *o++ = kNone;
        }
    }
    return o;
}

// 63 values: 9 counts / means, 6 x (epsilon, violations), the distance-ratio
// and rank summaries, the buckets.
void put_metrics(double* o, const bm::ApproximationMetrics& m)
{
    *o++ = static_cast<double>(m.query_count);
    *o++ = static_cast<double>(m.optimal_representative_count);
    *o++ = static_cast<double>(m.non_optimal_count);
    *o++ = static_cast<double>(m.reference_improvement_count);
    *o++ = static_cast<double>(m.zero_optimum_query_count);
    *o++ = static_cast<double>(m.zero_optimum_non_optimal_count);
    *o++ = static_cast<double>(m.conditional_non_optimal_ratio_count);
    *o++ = m.conditional_mean_distance_ratio.value_or(kNone);
    *o++ = m.conditional_mean_relative_excess.value_or(kNone);
    for (const auto& f : m.approximation_failures) {
        *o++ = f.epsilon;
        *o++ = static_cast<double>(f.violation_count);
    }
    o = put_summary(o, m.distance_ratio);
    o = put_summary(o, m.returned_representative_rank);
    put_buckets(o, m.margin_buckets);
}

uh::CpuDenseL2QueryStrategy cpu_strategy(int code)
{
    switch (code) {
    case 0: return uh::CpuDenseL2QueryStrategy::sequential;
    case 1: return uh::CpuDenseL2QueryStrategy::parallel_queries;
    case 2: return uh::CpuDenseL2QueryStrategy::parallel_representatives;
    default: return uh::CpuDenseL2QueryStrategy::automatic;
    }
}

// The reference benchmark's timed scope for a CPU L2 index
// (nearest_neighbor_benchmark.cpp:386-413 around run_cpu_l2_batched,
// :454-466): query_batch over batches of `batch` queries with `strategy`
// (0 sequential, 1 parallel_queries, 2 parallel_representatives, 3
// automatic), steady_clock around the whole query set. Index build is not
// timed. *seconds gets the wall time.
template <class Index>
void timed_batches(const Index& index, const float* queries, uint64_t d, uint64_t b, uint64_t batch, int strategy,
                   uint64_t* out, double* seconds)
{
    std::vector<std::size_t> res(b);
    const auto t0 = std::chrono::steady_clock::now();
    for (uint64_t begin = 0; begin < b; begin += batch) {
        const uint64_t count = std::min(batch, b - begin);
        index.query_batch(std::span<const float>(queries + begin * d, count * d), count,
                          std::span<std::size_t>(res).subspan(begin, count), cpu_strategy(strategy));
    }
    const auto t1 = std::chrono::steady_clock::now();
    *seconds = std::chrono::duration<double>(t1 - t0).count();
    for (uint64_t i = 0; i < b; ++i) out[i] = res[i];
}

}  // namespace

extern "C" {

// benchmark::ExactDistanceTable(metric, refs, queries, b): query_geometry()
// into geo[40] (4 counts, the finite-margin summary, the buckets),
// evaluate(pred) into ev[63] and evaluate_selected_distances(ref_pred, pred)
// into sel[63].
int ref_distance_diagnostics(int metric_l1, const float* refs, uint64_t r, uint64_t d, const float* queries,
                             uint64_t b, const uint64_t* pred, const uint64_t* ref_pred, double* geo, double* ev,
                             double* sel)
{
    return guarded([&] {
        const auto m = make_matrix(refs, r, d);
        const auto q = make_matrix(queries, b, d);
        const auto metric = metric_l1 ? bm::DistanceMetric::l1 : bm::DistanceMetric::l2;
        const bm::ExactDistanceTable table(metric, m, q, b);
        const auto& g = table.query_geometry();
        double* o = geo;
        *o++ = static_cast<double>(g.query_count);
        *o++ = static_cast<double>(g.zero_optimum_query_count);
        *o++ = static_cast<double>(g.non_unique_optimum_query_count);
        *o++ = static_cast<double>(g.infinite_multiplicative_margin_count);
        o = put_summary(o, g.finite_multiplicative_margin);
        put_buckets(o, g.margin_buckets);
        std::vector<std::size_t> p(pred, pred + b), rp(ref_pred, ref_pred + b);
        put_metrics(ev, table.evaluate(p));
        put_metrics(sel, bm::evaluate_selected_distances(metric, m, q, b, rp, p));
    });
}


// ExactL2Index::query_batch with CpuDenseL2QueryStrategy: 0 sequential,
// 2 parallel_queries (dense_l2_scan_index.hpp enum order checked in tests).
int ref_exact_l2(const float* refs, uint64_t r, uint64_t d, const float* queries,
                 uint64_t b, int parallel, uint64_t* out)
{
    return guarded([&] {
        const auto m = make_matrix(refs, r, d);
        const uh::ExactL2Index index(m);
        std::vector<std::size_t> res(b);
        index.query_batch(std::span<const float>(queries, b * d), b, res,
                          parallel ? uh::CpuDenseL2QueryStrategy::parallel_queries
                                   : uh::CpuDenseL2QueryStrategy::sequential);
        for (uint64_t i = 0; i < b; ++i) out[i] = res[i];
    });
}

int ref_exact_l2_timed(const float* refs, uint64_t r, uint64_t d, const float* queries, uint64_t b,
                       uint64_t batch, int strategy, uint64_t* out, double* seconds)
{
    return guarded([&] {
        const auto m = make_matrix(refs, r, d);
        const uh::ExactL2Index index(m);
        timed_batches(index, queries, d, b, batch == 0 ? b : batch, strategy, out, seconds);
    });
}

// FlatL2AnnIndex(input, probabilities, t, mt19937_64(seed)) built untimed,
// then the timed query set as above.
int ref_flat_l2_timed(const float* refs, uint64_t r, uint64_t d, const double* probs, uint64_t t, uint64_t seed,
                      const float* queries, uint64_t b, uint64_t batch, int strategy, uint64_t* out,
                      double* seconds)
{
    return guarded([&] {
        const auto m = make_matrix(refs, r, d);
        std::mt19937_64 eng(seed);
        const uh::FlatL2AnnIndex index(m, std::span<const double>(probs, d), t, eng);
        timed_batches(index, queries, d, b, batch == 0 ? b : batch, strategy, out, seconds);
    });
}

int ref_exact_l1(const float* refs, uint64_t r, uint64_t d, const float* queries,
                 uint64_t b, uint64_t* out)
{
    return guarded([&] {
        const auto m = make_matrix(refs, r, d);
        const uh::ExactL1Index index(m);
        for (uint64_t i = 0; i < b; ++i) {
            out[i] = index.query(std::span<const float>(queries + i * d, d));
        }
    });
}

int64_t ref_build_coordinate_sample(const double* probs, uint64_t d, uint64_t t,
                                    uint64_t seed, uint64_t* cols, uint64_t* mult,
                                    double* invp)
{
    try {
        std::mt19937_64 eng(seed);
        const auto s = uh::build_coordinate_sample(std::span<const double>(probs, d),
                                                   t, eng);
        for (std::size_t i = 0; i < s.columns.size(); ++i) {
            cols[i] = s.columns[i].source_column;
            mult[i] = s.columns[i].multiplicity;
            invp[i] = s.columns[i].inverse_probability;
        }
        return static_cast<int64_t>(s.columns.size());
    } catch (const std::exception&) {
        return -1;
    }
}

int ref_flat(int metric_l1, const float* refs, uint64_t r, uint64_t d,
             const double* probs, uint64_t t, uint64_t seed, const float* queries,
             uint64_t b, uint64_t* out)
{
    return guarded([&] {
        const auto m = make_matrix(refs, r, d);
        std::mt19937_64 eng(seed);
        const std::span<const double> p(probs, d);
        if (metric_l1) {
            const uh::FlatL1AnnIndex index(m, p, t, eng);
            for (uint64_t i = 0; i < b; ++i) {
                out[i] = index.query(std::span<const float>(queries + i * d, d));
            }
        } else {
            const uh::FlatL2AnnIndex index(m, p, t, eng);
            std::vector<std::size_t> res(b);
            index.query_batch(std::span<const float>(queries, b * d), b, res);
            for (uint64_t i = 0; i < b; ++i) out[i] = res[i];
        }
    });
}

// compute_sampling_mass; 1 on an exception (domain_error for bad values).
int ref_sampling_mass(const double* probs, uint64_t d, double* mass)
{
    return guarded([&] { *mass = uh::compute_sampling_mass(std::span<const double>(probs, d)); });
}

// make_mass_matched_uniform_probabilities into out (d doubles).
int ref_uniform_probabilities(uint64_t d, double mass, double* out)
{
    return guarded([&] {
        const auto p = uh::make_mass_matched_uniform_probabilities(d, mass);
        for (std::size_t i = 0; i < p.size(); ++i) out[i] = p[i];
    });
}

// UniformL{1,2}AnnIndex(input, mass, t, mt19937_64(seed)).query per query;
// usage[0..1] = unique query coordinates, sampled multiplicity.
int ref_uniform(int metric_l1, const float* refs, uint64_t r, uint64_t d, double mass, uint64_t t,
                uint64_t seed, const float* queries, uint64_t b, uint64_t* out, uint64_t* usage)
{
    return guarded([&] {
        const auto m = make_matrix(refs, r, d);
        std::mt19937_64 eng(seed);
        auto run = [&](const auto& index) {
            for (uint64_t i = 0; i < b; ++i) out[i] = index.query(std::span<const float>(queries + i * d, d));
            usage[0] = index.space_usage().unique_query_coordinates;
            usage[1] = index.space_usage().sampled_multiplicity;
        };
        if (metric_l1) {
            run(uh::UniformL1AnnIndex(m, mass, t, eng));
        } else {
            run(uh::UniformL2AnnIndex(m, mass, t, eng));
        }
    });
}

// io::sha256_file into digest[32].
int ref_sha256_file(const char* path, uint8_t* digest)
{
    return guarded([&] {
        const auto d = uh::io::sha256_file(path);
        std::memcpy(digest, d.data(), d.size());
    });
}

// io::save_importance_probabilities (kind_l1: 1 = l1, 0 = l2).
int ref_save_probabilities(const char* path, int kind_l1, uint64_t rows, const uint8_t* digest,
                           const double* probs, uint64_t d)
{
    return guarded([&] {
        uh::io::Sha256Digest dg{};
        std::memcpy(dg.data(), digest, dg.size());
        uh::io::save_importance_probabilities(
            path, kind_l1 ? uh::io::ImportanceProbabilityKind::l1 : uh::io::ImportanceProbabilityKind::l2, rows, dg,
            std::span<const double>(probs, d));
    });
}

// io::load_importance_probabilities; out holds capacity doubles.
int ref_load_probabilities(const char* path, int* kind_l1, uint64_t* rows, uint8_t* digest, double* out,
                           uint64_t capacity, uint64_t* d)
{
    return guarded([&] {
        const auto f = uh::io::load_importance_probabilities(path);
        *kind_l1 = f.kind == uh::io::ImportanceProbabilityKind::l1 ? 1 : 0;
        *rows = f.representative_count;
        std::memcpy(digest, f.representatives_sha256.data(), 32);
        *d = f.probabilities.size();
        for (std::size_t i = 0; i < f.probabilities.size() && i < capacity; ++i) out[i] = f.probabilities[i];
    });
}

int ref_hierarchical(int metric_l1, const float* refs, uint64_t r, uint64_t d,
                     const double* probs, uint64_t t, uint64_t p, uint64_t seed,
                     const float* queries, uint64_t b, uint64_t* out)
{
    return guarded([&] {
        const auto m = make_matrix(refs, r, d);
        std::mt19937_64 eng(seed);
        const std::span<const double> pr(probs, d);
        if (metric_l1) {
            const uh::HierarchicalL1AnnIndex index(m, pr, t, p, eng);
            for (uint64_t i = 0; i < b; ++i) {
                out[i] = index.query(std::span<const float>(queries + i * d, d));
            }
        } else {
            const uh::HierarchicalL2AnnIndex index(m, pr, t, p, eng);
            for (uint64_t i = 0; i < b; ++i) {
                out[i] = index.query(std::span<const float>(queries + i * d, d));
            }
        }
    });
}

// The FP32 device-side construction of the CUDA hierarchies
// (build_l{1,2}_hierarchical_projection_fp32): sample, projection matrix and
// projected references, with the engine seeded like the index constructors.
// Buffers: cols/mult/invp of capacity d, proj32 p x k, refs32 r x p.
int64_t ref_hierarchical_fp32_build(int metric_l1, const float* refs, uint64_t r,
                                    uint64_t d, const double* probs, uint64_t t,
                                    uint64_t p, uint64_t seed, uint64_t* cols,
                                    uint64_t* mult, double* invp, float* proj32,
                                    float* refs32)
{
    try {
        const auto m = make_matrix(refs, r, d);
        std::mt19937_64 eng(seed);
        auto sample = uh::build_coordinate_sample(std::span<const double>(probs, d),
                                                  t, eng);
        auto copy_out = [&](const auto& built) {
            const auto& cs = built.importance_sample.columns;
            for (std::size_t i = 0; i < cs.size(); ++i) {
                cols[i] = cs[i].source_column;
                mult[i] = cs[i].multiplicity;
                invp[i] = cs[i].inverse_probability;
            }
            return static_cast<int64_t>(cs.size());
        };
        if (metric_l1) {
            auto built = uh::detail::build_l1_hierarchical_projection_fp32(
                m, std::move(sample), p, eng);
            const int64_t k = copy_out(built);
            std::memcpy(proj32, built.cauchy_matrix.values().data(),
                        built.cauchy_matrix.values().size_bytes());
            std::memcpy(refs32, built.projected_representatives.values().data(),
                        built.projected_representatives.values().size_bytes());
            return k;
        }
        auto built = uh::detail::build_l2_hierarchical_projection_fp32(
            m, std::move(sample), p, eng);
        const int64_t k = copy_out(built);
        std::memcpy(proj32, built.jl_matrix.values().data(),
                    built.jl_matrix.values().size_bytes());
        std::memcpy(refs32, built.projected_representatives.values().data(),
                    built.projected_representatives.values().size_bytes());
        return k;
    } catch (const std::exception&) {
        return -1;
    }
}

int ref_importance_probabilities(int metric_l1, const float* refs, uint64_t r,
                                 uint64_t d, double* out)
{
    return guarded([&] {
        const auto m = make_matrix(refs, r, d);
        const auto p = metric_l1
                           ? uh::compute_l1_importance_probabilities(
                                 m, uh::ExecutionPolicy::sequential)
                           : uh::compute_l2_importance_probabilities(
                                 m, uh::ExecutionPolicy::sequential);
        for (uint64_t i = 0; i < d; ++i) out[i] = p[i];
    });
}

}  // extern "C"
