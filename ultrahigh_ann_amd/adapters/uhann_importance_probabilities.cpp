// The GPU importance-probability entry points of the reference
// (compute_l{1,2}_importance_probabilities_gpu, include/ultrahigh_ann/
// hnsvw25/l{1,2}/importance_sampling.hpp) on this engine: replaces
// src/hnsvw25/l{1,2}/cuda_importance_sampling.cu. The engine computes the
// probabilities in binary64 with the CPU policy's operation order
// (uhann_importance_probabilities_device), so the result equals the
// reference's sequential CPU policy bit for bit -- where the reference's
// binary32 GPU kernels approximate it (its tests allow 1e-6). Both L2
// distance backends run that same computation: there is no separate Gram
// backend, so the working set reported for `cublas` is the direct one.

#include "ultrahigh_ann/hnsvw25/l1/importance_sampling.hpp"
#include "ultrahigh_ann/hnsvw25/l2/importance_sampling.hpp"

#include "uhann_adapter_support.hpp"

#include <chrono>

namespace ultrahigh_ann {

namespace ua = uhann_adapter;

namespace {

template <class Result>
Result compute(int metric, const DenseMatrix& reps, int device, std::size_t pair_chunks)
{
    const auto t0 = std::chrono::steady_clock::now();
    if (pair_chunks == 0) throw std::invalid_argument("GPU pair chunk count must be positive");
    if (device < 0) throw std::invalid_argument("CUDA device index must be nonnegative");
    const std::size_t r = reps.rows();
    const std::size_t d = reps.cols();
    Result res{};
    res.probabilities.assign(d, 0.0);
    ua::check(uhann_importance_probabilities_device(metric, reps.values().data(), r, d, device,
                                                    res.probabilities.data()));
    res.device_name = ua::device_name(device);
    res.device = device;
    res.pair_count = r < 2 ? 0 : r * (r - 1) / 2;
    res.pair_chunks = pair_chunks;
    // rows x dim binary32 references and rows x rows binary64 inverse distances
    res.device_working_set_bytes = ua::product(r, d, "representative matrix") * sizeof(float) +
                                   ua::product(r, r, "inverse distances") * sizeof(double);
    res.timings.total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return res;
}

}  // namespace

bool l2_gpu_backend_available() noexcept { return uhann_device_available() != 0; }
bool l1_gpu_backend_available() noexcept { return uhann_device_available() != 0; }

L2GpuProbabilityResult compute_l2_importance_probabilities_gpu(const DenseMatrix& representatives, int device,
                                                               std::size_t pair_chunks,
                                                               L2GpuDistanceBackend distance_backend)
{
    auto res = compute<L2GpuProbabilityResult>(UHANN_METRIC_L2, representatives, device, pair_chunks);
    res.distance_backend = distance_backend;
    return res;
}

L1GpuProbabilityResult compute_l1_importance_probabilities_gpu(const DenseMatrix& representatives, int device,
                                                               std::size_t pair_chunks,
                                                               L1GpuDistanceBackend distance_backend)
{
    auto res = compute<L1GpuProbabilityResult>(UHANN_METRIC_L1, representatives, device, pair_chunks);
    res.distance_backend = distance_backend;
    return res;
}

}  // namespace ultrahigh_ann
