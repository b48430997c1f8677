// Host-side index construction in the reference's exact order, so an index
// built here is identical to the reference's for the same engine state.
//
// Header-only C++20, shared by the engine library (C ABI entry points that
// own an engine) and by the C++ facades (which draw from the caller's
// std::mt19937_64&). The draws use libstdc++'s <random> distributions, whose
// algorithms are implementation-defined: identity with the reference holds
// for the same standard library (checked against the compiled reference in
// the tests).
#pragma once

#include <cmath>
#include <cstddef>
#include <cstdint>
#include <limits>
#include <random>
#include <span>
#include <stdexcept>
#include <string>
#include <vector>

namespace ultrahigh_ann_amd {

// One coordinate kept by T Bernoulli repetitions (reference SampledColumn,
// include/ultrahigh_ann/coordinate_sampling/coordinate_sample.hpp:10-18).
struct SampledColumn {
    std::size_t source_column;
    std::size_t multiplicity;
    double inverse_probability;
};

struct CoordinateSample {
    std::vector<SampledColumn> columns;
};

// build_coordinate_sample (coordinate_sampling.cpp:14-45): a fresh
// binomial_distribution<size_t>(T, p_j) per column with p_j > 0 and T > 0;
// columns with p_j == 0 consume no randomness. Non-finite p_j ->
// domain_error, p_j outside [0, 1] -> domain_error.
inline CoordinateSample build_coordinate_sample(std::span<const double> probabilities,
                                                std::size_t repetitions,
                                                std::mt19937_64& engine)
{
    CoordinateSample sample;
    for (std::size_t column = 0; column < probabilities.size(); ++column) {
        const double p = probabilities[column];
        if (!std::isfinite(p)) throw std::domain_error("sampling probability must be finite");
        if (p < 0.0 || p > 1.0) throw std::domain_error("invalid sampling probability");
        if (p == 0.0 || repetitions == 0) continue;
        std::binomial_distribution<std::size_t> draw(repetitions, p);
        const std::size_t m = draw(engine);
        if (m != 0) sample.columns.push_back(SampledColumn{column, m, 1.0 / p});
    }
    return sample;
}

// compute_sampling_mass (coordinate_sampling/uniform_probabilities.cpp:12-26):
// the probabilities summed in order; a non-finite value or one outside
// [0, 1] -> domain_error.
inline double sampling_mass(std::span<const double> probabilities)
{
    double mass = 0.0;
    for (const double p : probabilities) {
        if (!std::isfinite(p)) throw std::domain_error("sampling probability must be finite");
        if (p < 0.0 || p > 1.0) throw std::domain_error("invalid sampling probability");
        mass += p;
    }
    return mass;
}

// make_mass_matched_uniform_probabilities (uniform_probabilities.cpp:28-45):
// `dimension` copies of mass / dimension, the uniform sampler with the same
// expected number of kept columns. A non-finite mass, or one outside
// [0, dimension] -> domain_error.
inline std::vector<double> uniform_probabilities(std::size_t dimension, double mass)
{
    if (!std::isfinite(mass)) throw std::domain_error("sampling mass must be finite");
    if (mass < 0.0 || mass > static_cast<double>(dimension)) {
        throw std::domain_error("sampling mass must be between zero and the dimension");
    }
    if (dimension == 0) return {};
    return std::vector<double>(dimension, mass / static_cast<double>(dimension));
}

// Sample checks shared by the embeddings and the hierarchies
// (l2_coordinate_embedding.cpp:41-60, l2_hierarchical_projection.cpp:29-51).
inline std::size_t validate_sample(const CoordinateSample& sample, std::size_t dim)
{
    std::size_t total = 0;
    for (const SampledColumn& c : sample.columns) {
        if (c.source_column >= dim) throw std::out_of_range("sampled column index out of range");
        if (c.multiplicity == 0 || !std::isfinite(c.inverse_probability) ||
            c.inverse_probability <= 0.0) {
            throw std::invalid_argument("sampled columns require positive finite weights");
        }
        if (total > std::numeric_limits<std::size_t>::max() - c.multiplicity) {
            throw std::length_error("sampled multiplicity overflows");
        }
        total += c.multiplicity;
    }
    return total;
}

// Per-column factors of the flat embeddings:
// L2 scale_j = (float)sqrt((double)m_j * (1/p_j)) (l2_coordinate_embedding.cpp:62-72),
// L1 weight_j = (float)((double)m_j * (1/p_j)) (l1_coordinate_embedding.cpp:52-62).
inline std::vector<float> flat_factors(const CoordinateSample& sample, bool metric_l1)
{
    std::vector<float> out;
    out.reserve(sample.columns.size());
    for (const SampledColumn& c : sample.columns) {
        const double w = static_cast<double>(c.multiplicity) * c.inverse_probability;
        const float f = metric_l1 ? static_cast<float>(w) : static_cast<float>(std::sqrt(w));
        if (!std::isfinite(w) || w <= 0.0 || !std::isfinite(f) || f <= 0.0F) {
            throw std::invalid_argument("sampled coordinate weight is not representable in float32");
        }
        out.push_back(f);
    }
    return out;
}

// Projection matrix of the hierarchical indexes, p rows x |I| columns,
// binary64, drawn row-major (row outer, sampled column inner) from the engine
// state that follows the coordinate sample.
//   L2, build_jl_matrix (l2_hierarchical_projection.cpp:53-77): fresh
//     binomial_distribution<size_t>(m_j, 0.5) per entry,
//     coef = (2 pos - m_j) * sqrt(1/p_j) * (1/sqrt(p)).
//   L1, build_cauchy_matrix (l1_hierarchical_projection.cpp:54-77): one
//     cauchy_distribution<double>(0, 1), coef = c * (1/p_j) * m_j.
// A non-finite coefficient -> overflow_error.
inline std::vector<double> build_projection(const CoordinateSample& sample, std::size_t rows,
                                            bool metric_l1, std::mt19937_64& engine)
{
    const std::size_t cols = sample.columns.size();
    if (cols != 0 && rows > std::numeric_limits<std::size_t>::max() / cols) {
        throw std::length_error("projection matrix size overflows");
    }
    std::vector<double> data;
    data.reserve(rows * cols);
    if (!metric_l1) {
        const double row_scaling = 1.0 / std::sqrt(static_cast<double>(rows));
        for (std::size_t row = 0; row < rows; ++row) {
            for (const SampledColumn& c : sample.columns) {
                std::binomial_distribution<std::size_t> signs(c.multiplicity, 0.5);
                const std::size_t positives = signs(engine);
                const double signed_sum =
                    2.0 * static_cast<double>(positives) - static_cast<double>(c.multiplicity);
                const double coef = signed_sum * std::sqrt(c.inverse_probability) * row_scaling;
                if (!std::isfinite(coef)) throw std::overflow_error("JL projection coefficient is not finite");
                data.push_back(coef);
            }
        }
    } else {
        std::cauchy_distribution<double> cauchy(0.0, 1.0);
        for (std::size_t row = 0; row < rows; ++row) {
            for (const SampledColumn& c : sample.columns) {
                const double coef =
                    cauchy(engine) * c.inverse_probability * static_cast<double>(c.multiplicity);
                if (!std::isfinite(coef)) {
                    throw std::overflow_error("Cauchy projection coefficient is not finite");
                }
                data.push_back(coef);
            }
        }
    }
    return data;
}

// FP32 device-side form of a hierarchy (build_l2_hierarchical_projection_fp32,
// l2_hierarchical_projection.cpp:142-162, and its L1 twin): the projection
// narrowed to binary32 (narrow_matrix_to_fp32, fp32_matrix_operations.hpp:
// 14-25) and the sampled references multiplied by it with binary32 products
// and a sequential binary32 sum (multiply_right_transposed_fp32, :29-67).
// refs is rows x dim row-major; outputs proj32 (p x k) and projected refs
// (rows x p). Non-finite values -> overflow_error. The product and the sum
// are separate roundings (no fused multiply-add), as in the reference build.
#if defined(__GNUC__) && !defined(__clang__)
__attribute__((optimize("fp-contract=off")))
#endif
inline void fp32_hierarchy(std::span<const float> refs, std::size_t rows, std::size_t dim,
                           const CoordinateSample& sample, std::span<const double> proj64,
                           std::size_t p, std::vector<float>& proj32,
                           std::vector<float>& projected_refs)
{
    const std::size_t k = sample.columns.size();
    proj32.resize(p * k);
    for (std::size_t i = 0; i < p * k; ++i) {
        const float v = static_cast<float>(proj64[i]);
        if (!std::isfinite(v)) throw std::overflow_error("FP32 projection coefficient is not finite");
        proj32[i] = v;
    }
    projected_refs.assign(rows * p, 0.0F);
    std::vector<float> sampled(k);
    for (std::size_t r = 0; r < rows; ++r) {
        for (std::size_t j = 0; j < k; ++j) sampled[j] = refs[r * dim + sample.columns[j].source_column];
        for (std::size_t o = 0; o < p; ++o) {
            float sum = 0.0F;
            for (std::size_t j = 0; j < k; ++j) {
                const float prod = sampled[j] * proj32[o * k + j];
                sum = sum + prod;
            }
            if (!std::isfinite(sum)) throw std::overflow_error("FP32 projected representative is not finite");
            projected_refs[r * p + o] = sum;
        }
    }
}

}  // namespace ultrahigh_ann_amd
