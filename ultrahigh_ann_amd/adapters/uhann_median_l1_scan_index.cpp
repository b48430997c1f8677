// detail::CudaMedianL1ScanIndex (src/hnsvw25/l1/hierarchical/
// cuda_median_l1_scan_index.hpp) on this engine: replaces
// src/hnsvw25/l1/hierarchical/cuda_median_l1_scan_index.cu. The scan and its
// workspaces are uhann_median / uhann_workspace handles; the median of the p
// absolute differences is the reference's value (it depends only on the
// multiset), ties go to the lowest row.

#include "hnsvw25/l1/hierarchical/cuda_median_l1_scan_index.hpp"

#include "uhann_adapter_support.hpp"

#include <utility>

namespace ultrahigh_ann::detail {

namespace ua = uhann_adapter;

struct CudaMedianL1ScanIndex::Impl {
    Impl(const DenseMatrix& projected, int dev) : device(dev), rows(projected.rows()), p(projected.cols())
    {
        ua::check(uhann_median_create(projected.values().data(), rows, p, device, &h));
        name = ua::device_name(device);
        payload = projected.values().size() * sizeof(float);
    }
    Impl(const Impl&) = delete;
    Impl& operator=(const Impl&) = delete;
    ~Impl() { uhann_median_destroy(h); }

    uhann_median* h = nullptr;
    int device = 0;
    std::size_t rows = 0;
    std::size_t p = 0;
    std::size_t payload = 0;
    std::string name;
};

struct CudaMedianL1ScanIndex::Workspace::Impl {
    Impl(const CudaMedianL1ScanIndex::Impl* o, std::size_t max_batch) : owner(o), capacity(max_batch)
    {
        if (max_batch == 0) throw std::invalid_argument("CUDA median L1 workspace batch size must be positive");
        ua::check(uhann_median_workspace_create(owner->h, max_batch, &ws));
        // medians (binary32), results (size_t) and the validation flag
        payload = ua::product(max_batch, owner->rows, "CUDA median L1 workspace") * sizeof(float) +
                  max_batch * sizeof(std::size_t) + sizeof(int);
    }
    Impl(const Impl&) = delete;
    Impl& operator=(const Impl&) = delete;
    ~Impl() { uhann_workspace_destroy(ws); }

    const CudaMedianL1ScanIndex::Impl* owner = nullptr;
    uhann_workspace* ws = nullptr;
    std::size_t capacity = 0;
    std::size_t payload = 0;
};

CudaMedianL1ScanIndex::Workspace::Workspace(std::unique_ptr<Impl> implementation)
    : implementation_(std::move(implementation))
{
}
CudaMedianL1ScanIndex::Workspace::Workspace(Workspace&&) noexcept = default;
CudaMedianL1ScanIndex::Workspace& CudaMedianL1ScanIndex::Workspace::operator=(Workspace&&) noexcept = default;
CudaMedianL1ScanIndex::Workspace::~Workspace() = default;

std::size_t CudaMedianL1ScanIndex::Workspace::maximum_batch_size() const noexcept
{
    return implementation_ ? implementation_->capacity : 0;
}

std::size_t CudaMedianL1ScanIndex::Workspace::payload_bytes() const noexcept
{
    return implementation_ ? implementation_->payload : 0;
}

CudaMedianL1ScanIndex::CudaMedianL1ScanIndex(const DenseMatrix& projected_representatives, int device)
    : implementation_(std::make_unique<Impl>(projected_representatives, device))
{
}
CudaMedianL1ScanIndex::CudaMedianL1ScanIndex(CudaMedianL1ScanIndex&&) noexcept = default;
CudaMedianL1ScanIndex& CudaMedianL1ScanIndex::operator=(CudaMedianL1ScanIndex&&) noexcept = default;
CudaMedianL1ScanIndex::~CudaMedianL1ScanIndex() = default;

CudaMedianL1ScanIndex::Workspace CudaMedianL1ScanIndex::make_workspace(std::size_t maximum_batch_size) const
{
    return Workspace(std::make_unique<Workspace::Impl>(implementation_.get(), maximum_batch_size));
}

void CudaMedianL1ScanIndex::query_device_batch(const float* projected_queries, std::size_t query_count,
                                               std::span<std::size_t> output, Workspace& workspace) const
{
    const auto* w = workspace.implementation_.get();
    if (w == nullptr || w->owner != implementation_.get()) {
        throw std::invalid_argument("CUDA median L1 workspace belongs to a different index");
    }
    if (query_count > w->capacity) throw std::invalid_argument("query count exceeds CUDA median L1 workspace capacity");
    if (output.size() != query_count) throw std::invalid_argument("CUDA median L1 output size does not match query count");
    if (query_count == 0) return;
    ua::check(uhann_median_query_device_batch(implementation_->h, w->ws, projected_queries, query_count,
                                              ua::as_u64(output.data())));
}

std::size_t CudaMedianL1ScanIndex::payload_bytes() const noexcept { return implementation_->payload; }
std::size_t CudaMedianL1ScanIndex::representative_count() const noexcept { return implementation_->rows; }
std::size_t CudaMedianL1ScanIndex::projection_dimension() const noexcept { return implementation_->p; }
int CudaMedianL1ScanIndex::device() const noexcept { return implementation_->device; }
const std::string& CudaMedianL1ScanIndex::device_name() const noexcept { return implementation_->name; }

}  // namespace ultrahigh_ann::detail
