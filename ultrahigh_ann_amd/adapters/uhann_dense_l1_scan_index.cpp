// CudaDenseL1ScanIndex (include/ultrahigh_ann/core/cuda_dense_l1_scan_index.hpp:
// 19-72) on this engine: replaces src/core/cuda_dense_l1_scan_index.cu. The
// index and its workspaces are uhann_scan / uhann_workspace handles with the
// L1 metric; the optional coordinate weights scale |r - q| as in the
// reference's kernel (binary32, weight and difference fused as fmaf).
//
// space_usage() and payload_bytes() report the reference's logical buffers
// (representatives and weights; queries, distances and results);
// uhann_scan_info and uhann_workspace_info give this engine's device bytes.

#include "ultrahigh_ann/core/cuda_dense_l1_scan_index.hpp"

#include "uhann_adapter_support.hpp"

#include <array>
#include <cmath>
#include <utility>

namespace ultrahigh_ann {

namespace ua = uhann_adapter;

bool cuda_dense_l1_scan_available() noexcept { return uhann_device_available() != 0; }

struct CudaDenseL1ScanIndex::Impl {
    Impl(const DenseMatrix& reps, std::span<const float> weights, int dev)
        : device(dev), rows(reps.rows()), dim(reps.cols())
    {
        if (!weights.empty() && weights.size() != dim) {
            throw std::invalid_argument("CUDA dense L1 coordinate weights do not match dimension");
        }
        for (const float w : weights) {
            if (!std::isfinite(w) || !(w > 0.0F)) {
                throw std::invalid_argument("CUDA dense L1 coordinate weights must be positive and finite");
            }
        }
        ua::check(uhann_scan_create(UHANN_METRIC_L1, reps.values().data(), rows, dim,
                                    weights.empty() ? nullptr : weights.data(), device, &scan));
        name = ua::device_name(device);
        payload = (ua::product(rows, dim, "CUDA dense L1 index") + weights.size()) * sizeof(float);
    }
    Impl(const Impl&) = delete;
    Impl& operator=(const Impl&) = delete;
    ~Impl() { uhann_scan_destroy(scan); }

    uhann_scan* scan = nullptr;
    int device = 0;
    std::size_t rows = 0;
    std::size_t dim = 0;
    std::size_t payload = 0;
    std::string name;
};

struct CudaDenseL1ScanIndex::QueryWorkspace::Impl {
    Impl(const CudaDenseL1ScanIndex::Impl* o, std::size_t max_batch) : owner(o), capacity(max_batch)
    {
        if (max_batch == 0) throw std::invalid_argument("CUDA dense L1 workspace batch size must be positive");
        ua::check(uhann_scan_workspace_create(owner->scan, max_batch, &ws));
        payload = (ua::product(max_batch, owner->dim, "CUDA dense L1 query workspace") +
                   ua::product(max_batch, owner->rows, "CUDA dense L1 distance workspace")) *
                      sizeof(float) +
                  max_batch * sizeof(std::size_t);
    }
    Impl(const Impl&) = delete;
    Impl& operator=(const Impl&) = delete;
    ~Impl() { uhann_workspace_destroy(ws); }

    const CudaDenseL1ScanIndex::Impl* owner = nullptr;
    uhann_workspace* ws = nullptr;
    std::size_t capacity = 0;
    std::size_t payload = 0;
};

CudaDenseL1ScanIndex::QueryWorkspace::QueryWorkspace(std::unique_ptr<Impl> implementation)
    : implementation_(std::move(implementation))
{
}
CudaDenseL1ScanIndex::QueryWorkspace::QueryWorkspace(QueryWorkspace&&) noexcept = default;
CudaDenseL1ScanIndex::QueryWorkspace& CudaDenseL1ScanIndex::QueryWorkspace::operator=(QueryWorkspace&&) noexcept =
    default;
CudaDenseL1ScanIndex::QueryWorkspace::~QueryWorkspace() = default;

std::size_t CudaDenseL1ScanIndex::QueryWorkspace::maximum_batch_size() const noexcept
{
    return implementation_ ? implementation_->capacity : 0;
}

std::size_t CudaDenseL1ScanIndex::QueryWorkspace::payload_bytes() const noexcept
{
    return implementation_ ? implementation_->payload : 0;
}

CudaDenseL1ScanIndex::CudaDenseL1ScanIndex(const DenseMatrix& representatives, int device)
    : implementation_(std::make_unique<Impl>(representatives, std::span<const float>{}, device))
{
}

CudaDenseL1ScanIndex::CudaDenseL1ScanIndex(const DenseMatrix& representatives,
                                           std::span<const float> coordinate_weights, int device)
    : implementation_(std::make_unique<Impl>(representatives, coordinate_weights, device))
{
}

CudaDenseL1ScanIndex::CudaDenseL1ScanIndex(CudaDenseL1ScanIndex&&) noexcept = default;
CudaDenseL1ScanIndex& CudaDenseL1ScanIndex::operator=(CudaDenseL1ScanIndex&&) noexcept = default;
CudaDenseL1ScanIndex::~CudaDenseL1ScanIndex() = default;

CudaDenseL1ScanIndex::QueryWorkspace CudaDenseL1ScanIndex::make_query_workspace(std::size_t maximum_batch_size) const
{
    return QueryWorkspace(std::make_unique<QueryWorkspace::Impl>(implementation_.get(), maximum_batch_size));
}

std::size_t CudaDenseL1ScanIndex::query(std::span<const float> query, QueryWorkspace& workspace) const
{
    std::array<std::size_t, 1> out{};
    query_batch(query, 1, out, workspace);
    return out[0];
}

// The reference's checks in its order (cuda_dense_l1_scan_index.cu:459-485),
// then the engine validates the queries and runs the batch.
void CudaDenseL1ScanIndex::query_batch(std::span<const float> queries, std::size_t query_count,
                                       std::span<std::size_t> output, QueryWorkspace& workspace) const
{
    const auto* w = workspace.implementation_.get();
    if (w == nullptr || w->owner != implementation_.get()) {
        throw std::invalid_argument("CUDA dense L1 workspace belongs to a different index");
    }
    if (query_count > w->capacity) throw std::invalid_argument("query count exceeds CUDA dense L1 workspace capacity");
    if (queries.size() != ua::product(query_count, implementation_->dim, "CUDA dense L1 query batch")) {
        throw std::invalid_argument("query batch dimensions do not match index dimension");
    }
    if (output.size() != query_count) {
        throw std::invalid_argument("CUDA dense L1 output size does not match query count");
    }
    ua::check(uhann_scan_query_batch(implementation_->scan, w->ws, queries.data(), query_count,
                                     ua::as_u64(output.data()), nullptr, UHANN_STRATEGY_DIRECT));
}

IndexSpaceUsage CudaDenseL1ScanIndex::space_usage() const noexcept
{
    IndexSpaceUsage u{};
    u.index_payload_bytes = implementation_->payload;
    u.query_workspace_payload_bytes = 0;
    u.unique_query_coordinates = implementation_->dim;
    u.sampled_multiplicity = implementation_->dim;
    return u;
}

int CudaDenseL1ScanIndex::device() const noexcept { return implementation_->device; }

const std::string& CudaDenseL1ScanIndex::device_name() const noexcept { return implementation_->name; }

}  // namespace ultrahigh_ann
