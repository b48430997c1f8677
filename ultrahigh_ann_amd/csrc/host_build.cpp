// C ABI for host-side index construction: a seeded std::mt19937_64 handle,
// the coordinate sample, the hierarchical projection draws and the
// importance probabilities, all in the reference's order (see
// include/ultrahigh_ann_amd/build.hpp for the shared implementation).

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <random>
#include <span>
#include <thread>
#include <vector>

#include "common.hpp"
#include "engine.hpp"
#include "probabilities.hpp"
#include "ultrahigh_ann_amd/build.hpp"
#include "ultrahigh_ann_amd/probability_file.hpp"

namespace uab = ultrahigh_ann_amd;
using uhann::guarded;

struct uhann_rng {
    std::mt19937_64 engine;
};

namespace {

uab::CoordinateSample to_sample(const uhann_sampled_column* cols, uint64_t count)
{
    if (count != 0 && cols == nullptr) throw std::invalid_argument("sample must not be null");
    uab::CoordinateSample s;
    s.columns.reserve(count);
    for (uint64_t i = 0; i < count; ++i) {
        s.columns.push_back({cols[i].source_column, cols[i].multiplicity, cols[i].inverse_probability});
    }
    return s;
}

// Binary64 squared L2 / L1 distance of two binary32 rows in the reference's
// order (l2_distance.cpp:9-25; l1_distance.cpp:12-28 via libstdc++
// transform_reduce, which sums groups of four pairwise).
double sq_l2(const float* a, const float* b, std::size_t n)
{
    double s = 0.0;
    for (std::size_t i = 0; i < n; ++i) {
        const double d = static_cast<double>(a[i]) - static_cast<double>(b[i]);
        s += d * d;
    }
    return s;
}

double l1(const float* a, const float* b, std::size_t n)
{
    auto t = [&](std::size_t i) { return std::abs(static_cast<double>(a[i]) - static_cast<double>(b[i])); };
    double init = 0.0;
    std::size_t i = 0;
    for (; n - i >= 4; i += 4) init = init + ((t(i) + t(i + 1)) + (t(i + 2) + t(i + 3)));
    for (; i < n; ++i) init = init + t(i);
    return init;
}

}  // namespace

extern "C" {

int uhann_rng_create(uint64_t seed, uhann_rng** out)
{
    return guarded([&] {
        if (out == nullptr) throw std::invalid_argument("output handle must not be null");
        *out = new uhann_rng{std::mt19937_64(seed)};
    });
}

void uhann_rng_destroy(uhann_rng* rng) { delete rng; }

int uhann_rng_equal(const uhann_rng* a, const uhann_rng* b, int* equal)
{
    return guarded([&] {
        if (a == nullptr || b == nullptr || equal == nullptr) throw std::invalid_argument("null argument");
        *equal = a->engine == b->engine ? 1 : 0;
    });
}

int uhann_rng_next(uhann_rng* rng, uint64_t* value)
{
    return guarded([&] {
        if (rng == nullptr || value == nullptr) throw std::invalid_argument("null argument");
        *value = rng->engine();
    });
}

int uhann_build_coordinate_sample(uhann_rng* rng, const double* probabilities, uint64_t dim,
                                  uint64_t repetitions, uhann_sampled_column* out, uint64_t capacity,
                                  uint64_t* count)
{
    return guarded([&] {
        if (rng == nullptr || count == nullptr) throw std::invalid_argument("null argument");
        if (dim != 0 && probabilities == nullptr) throw std::invalid_argument("probabilities must not be null");
        const auto s = uab::build_coordinate_sample(std::span<const double>(probabilities, dim),
                                                    repetitions, rng->engine);
        *count = s.columns.size();
        if (s.columns.size() > capacity) throw std::length_error("sample output capacity is too small");
        for (std::size_t i = 0; i < s.columns.size(); ++i) {
            out[i] = {s.columns[i].source_column, s.columns[i].multiplicity, s.columns[i].inverse_probability};
        }
    });
}

int uhann_sampling_mass(const double* probabilities, uint64_t dim, double* mass)
{
    return guarded([&] {
        if (mass == nullptr || (dim != 0 && probabilities == nullptr)) throw std::invalid_argument("null argument");
        *mass = uab::sampling_mass(std::span<const double>(probabilities, dim));
    });
}

int uhann_uniform_probabilities(uint64_t dim, double mass, double* out)
{
    return guarded([&] {
        if (dim != 0 && out == nullptr) throw std::invalid_argument("output must not be null");
        const auto p = uab::uniform_probabilities(dim, mass);
        std::copy(p.begin(), p.end(), out);
    });
}

int uhann_sha256_file(const char* path, uint8_t* digest)
{
    return guarded([&] {
        if (path == nullptr || digest == nullptr) throw std::invalid_argument("null argument");
        const auto d = uab::sha256_file(path);
        std::memcpy(digest, d.data(), d.size());
    });
}

int uhann_save_probabilities(const char* path, int metric, uint64_t representative_count,
                             const uint8_t* digest, const double* probabilities, uint64_t dim)
{
    return guarded([&] {
        if (path == nullptr || digest == nullptr || (dim != 0 && probabilities == nullptr)) {
            throw std::invalid_argument("null argument");
        }
        if (metric != UHANN_METRIC_L2 && metric != UHANN_METRIC_L1) {
            throw std::invalid_argument("unknown importance-probability kind");
        }
        uab::Sha256Digest d{};
        std::memcpy(d.data(), digest, d.size());
        uab::save_probabilities(path, metric == UHANN_METRIC_L1 ? uab::ProbabilityKind::l1 : uab::ProbabilityKind::l2,
                                representative_count, d, std::span<const double>(probabilities, dim));
    });
}

int uhann_load_probabilities(const char* path, int* metric, uint64_t* representative_count, uint8_t* digest,
                             double* out, uint64_t capacity, uint64_t* dim)
{
    return guarded([&] {
        if (path == nullptr || metric == nullptr || representative_count == nullptr || digest == nullptr ||
            dim == nullptr) {
            throw std::invalid_argument("null argument");
        }
        const auto f = uab::load_probabilities(path);
        *metric = f.kind == uab::ProbabilityKind::l1 ? UHANN_METRIC_L1 : UHANN_METRIC_L2;
        *representative_count = f.representative_count;
        std::memcpy(digest, f.representatives_sha256.data(), f.representatives_sha256.size());
        *dim = f.probabilities.size();
        if (out == nullptr) return;
        if (capacity < f.probabilities.size()) throw std::length_error("probability output capacity is too small");
        std::copy(f.probabilities.begin(), f.probabilities.end(), out);
    });
}

int uhann_build_projection(uhann_rng* rng, int metric, const uhann_sampled_column* sample,
                           uint64_t count, uint64_t projection_dim, double* out)
{
    return guarded([&] {
        if (rng == nullptr) throw std::invalid_argument("null argument");
        if (metric != UHANN_METRIC_L2 && metric != UHANN_METRIC_L1) throw std::invalid_argument("unknown metric");
        if (projection_dim == 0) throw std::invalid_argument("projection_dimension has to be a positive integer");
        const auto s = to_sample(sample, count);
        const auto p = uab::build_projection(s, projection_dim, metric == UHANN_METRIC_L1, rng->engine);
        if (!p.empty()) {
            if (out == nullptr) throw std::invalid_argument("output must not be null");
            std::memcpy(out, p.data(), p.size() * sizeof(double));
        }
    });
}

// importance_sampling.cpp:50-107 (L2) and l1/importance_sampling.cpp:49-104
// (L1): per-thread column maxima over the row pairs a thread takes, then the
// maximum over threads. max is exact and order-free, so the result equals the
// sequential policy bit for bit.
int uhann_importance_probabilities(int metric, const float* refs, uint64_t rows, uint64_t dim,
                                   double* out)
{
    return guarded([&] {
        if (metric != UHANN_METRIC_L2 && metric != UHANN_METRIC_L1) throw std::invalid_argument("unknown metric");
        if (out == nullptr || (rows * dim != 0 && refs == nullptr)) throw std::invalid_argument("null argument");
        uhann::validate_finite_host(refs, rows * dim, "representatives");
        std::fill(out, out + dim, 0.0);
        if (rows < 2 || dim == 0) return;
        const unsigned hw = std::max(1U, std::thread::hardware_concurrency());
        const unsigned workers = static_cast<unsigned>(std::min<uint64_t>(rows - 1, std::min(hw, 16U)));
        std::vector<std::vector<double>> local(workers, std::vector<double>(dim, 0.0));
        std::atomic<uint64_t> next{0};
        const bool l1m = metric == UHANN_METRIC_L1;
        auto work = [&](unsigned w) {
            auto& mine = local[w];
            for (uint64_t a = next.fetch_add(1); a + 1 < rows; a = next.fetch_add(1)) {
                const float* ra = refs + a * dim;
                for (uint64_t b = a + 1; b < rows; ++b) {
                    const float* rb = refs + b * dim;
                    const double dist = l1m ? l1(ra, rb, dim) : sq_l2(ra, rb, dim);
                    if (dist == 0.0) continue;
                    const double inv = 1.0 / dist;
                    for (uint64_t c = 0; c < dim; ++c) {
                        const double diff = static_cast<double>(ra[c]) - static_cast<double>(rb[c]);
                        const double prob = l1m ? std::abs(diff) * inv : std::min(1.0, diff * diff * inv);
                        mine[c] = std::max(mine[c], prob);
                    }
                }
            }
        };
        std::vector<std::thread> pool;
        for (unsigned w = 1; w < workers; ++w) pool.emplace_back(work, w);
        work(0);
        for (auto& t : pool) t.join();
        for (const auto& l : local) {
            for (uint64_t c = 0; c < dim; ++c) out[c] = std::max(out[c], l[c]);
        }
    });
}

int uhann_importance_probabilities_device(int metric, const float* host_refs, uint64_t rows,
                                          uint64_t dim, int device, double* out)
{
    return guarded([&] {
        if (metric != UHANN_METRIC_L2 && metric != UHANN_METRIC_L1) throw std::invalid_argument("unknown metric");
        if (out == nullptr || (rows * dim != 0 && host_refs == nullptr)) throw std::invalid_argument("null argument");
        if (rows > 0xffffffffULL || dim > 0xffffffffULL) throw std::length_error("matrix exceeds the 32-bit range");
        uhann::validate_finite_host(host_refs, rows * dim, "representatives");
        if (dim == 0) return;
        uhann::select_device(device);
        uhann::DeviceBuffer<float> refs(device, uhann::checked_mul(rows, dim, "representatives"));
        uhann::DeviceBuffer<double> inv(device, rows >= 2 ? uhann::checked_mul(rows, rows, "pair table") : 0);
        uhann::DeviceBuffer<unsigned long long> bits(device, dim);
        uhann::check_hip(hipMemcpy(refs.data(), host_refs, refs.bytes(), hipMemcpyHostToDevice),
                         "cannot copy representatives to the HIP device");
        uhann::kern::importance_probabilities(metric == UHANN_METRIC_L1, refs.data(), dim,
                                              static_cast<uint32_t>(rows), static_cast<uint32_t>(dim),
                                              inv.data(), bits.data(), nullptr);
        uhann::check_hip(hipDeviceSynchronize(), "importance probability kernels failed");
        uhann::check_hip(hipMemcpy(out, bits.data(), dim * sizeof(double), hipMemcpyDeviceToHost),
                         "cannot copy probabilities from the HIP device");
    });
}

}  // extern "C"
