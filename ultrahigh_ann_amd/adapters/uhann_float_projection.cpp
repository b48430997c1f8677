// detail::CudaFloatProjection (src/core/cuda_float_projection.hpp) on this
// engine: replaces src/core/cuda_float_projection.cu. The projection and
// its workspaces are uhann_projection / uhann_workspace handles. Both
// reference strategies run the binary32 K7 order (the reference's `direct`
// arithmetic, cuda_float_projection.cu:234-263); its `gemm` strategy is a
// cuBLAS SGEMM whose summation order cannot be reproduced, and the direct
// order it is replaced by has the same binary32 precision.

#include "core/cuda_float_projection.hpp"

#include "uhann_adapter_support.hpp"

#include <utility>

namespace ultrahigh_ann::detail {

namespace ua = uhann_adapter;

struct CudaFloatProjection::Impl {
    Impl(const DenseMatrix& matrix, int dev) : device(dev), p(matrix.rows()), k(matrix.cols())
    {
        ua::check(uhann_projection_create(matrix.values().data(), p, k, device, &h));
        payload = matrix.values().size() * sizeof(float);
    }
    Impl(const Impl&) = delete;
    Impl& operator=(const Impl&) = delete;
    ~Impl() { uhann_projection_destroy(h); }

    uhann_projection* h = nullptr;
    int device = 0;
    std::size_t p = 0;
    std::size_t k = 0;
    std::size_t payload = 0;
};

struct CudaFloatProjection::Workspace::Impl {
    Impl(const CudaFloatProjection::Impl* o, std::size_t max_batch) : owner(o), capacity(max_batch)
    {
        if (max_batch == 0) throw std::invalid_argument("CUDA projection workspace batch size must be positive");
        ua::check(uhann_projection_workspace_create(owner->h, max_batch, &ws));
        // sampled queries, projected queries (binary32) and the validation flag
        payload = (ua::product(max_batch, owner->k, "CUDA projection workspace") +
                   ua::product(max_batch, owner->p, "CUDA projection workspace")) *
                      sizeof(float) +
                  sizeof(int);
    }
    Impl(const Impl&) = delete;
    Impl& operator=(const Impl&) = delete;
    ~Impl() { uhann_workspace_destroy(ws); }

    const CudaFloatProjection::Impl* owner = nullptr;
    uhann_workspace* ws = nullptr;
    std::size_t capacity = 0;
    std::size_t payload = 0;
};

CudaFloatProjection::Workspace::Workspace(std::unique_ptr<Impl> implementation)
    : implementation_(std::move(implementation))
{
}
CudaFloatProjection::Workspace::Workspace(Workspace&&) noexcept = default;
CudaFloatProjection::Workspace& CudaFloatProjection::Workspace::operator=(Workspace&&) noexcept = default;
CudaFloatProjection::Workspace::~Workspace() = default;

std::size_t CudaFloatProjection::Workspace::maximum_batch_size() const noexcept
{
    return implementation_ ? implementation_->capacity : 0;
}

std::size_t CudaFloatProjection::Workspace::payload_bytes() const noexcept
{
    return implementation_ ? implementation_->payload : 0;
}

CudaFloatProjection::CudaFloatProjection(const DenseMatrix& projection_matrix, int device)
    : implementation_(std::make_unique<Impl>(projection_matrix, device))
{
}
CudaFloatProjection::CudaFloatProjection(CudaFloatProjection&&) noexcept = default;
CudaFloatProjection& CudaFloatProjection::operator=(CudaFloatProjection&&) noexcept = default;
CudaFloatProjection::~CudaFloatProjection() = default;

CudaFloatProjection::Workspace CudaFloatProjection::make_workspace(std::size_t maximum_batch_size) const
{
    return Workspace(std::make_unique<Workspace::Impl>(implementation_.get(), maximum_batch_size));
}

const float* CudaFloatProjection::project_batch_float(std::span<const float> sampled_queries, std::size_t query_count,
                                                      Workspace& workspace, CudaFloatProjectionStrategy strategy) const
{
    const auto* w = workspace.implementation_.get();
    if (w == nullptr || w->owner != implementation_.get()) {
        throw std::invalid_argument("CUDA projection workspace belongs to a different index");
    }
    if (query_count > w->capacity) throw std::invalid_argument("query count exceeds CUDA projection workspace capacity");
    if (sampled_queries.size() != ua::product(query_count, implementation_->k, "CUDA sampled-query projection batch")) {
        throw std::invalid_argument("sampled query batch dimensions do not match projection matrix");
    }
    if (strategy != CudaFloatProjectionStrategy::direct && strategy != CudaFloatProjectionStrategy::gemm) {
        throw std::invalid_argument("unknown CUDA projection strategy");
    }
    const float* out = nullptr;
    ua::check(uhann_projection_project(implementation_->h, w->ws, sampled_queries.data(), query_count, &out));
    return out;
}

std::size_t CudaFloatProjection::sampled_dimension() const noexcept { return implementation_->k; }
std::size_t CudaFloatProjection::projection_dimension() const noexcept { return implementation_->p; }
std::size_t CudaFloatProjection::payload_bytes() const noexcept { return implementation_->payload; }
int CudaFloatProjection::device() const noexcept { return implementation_->device; }

}  // namespace ultrahigh_ann::detail
