// CudaDenseL2ScanIndex (include/ultrahigh_ann/core/cuda_dense_l2_scan_index.hpp:
// 26-89) on this engine: replaces src/core/cuda_dense_l2_scan_index.cu. The
// index and its workspaces are uhann_scan / uhann_workspace handles.
//
// Strategies: `direct` is UHANN_STRATEGY_DIRECT, the reference's binary32
// kernel arithmetic bit for bit. `gemm` is UHANN_STRATEGY_GRAM_INT8: a Gram
// product on MFMA like the reference's SGEMM strategy, but certified -- it
// returns the direct path's row (or the binary64 winner of a binary32 near
// tie), where the reference's unrefined SGEMM may return another row under
// cancellation (its own test asserts that, cuda_exact_l2_index_test.cpp:
// 152-174). The int8 reference copy is built on the first `gemm` call.
//
// space_usage() and payload_bytes() report the reference's logical buffers
// (representatives and squared norms; queries, distances, norms and
// results), so reports compare across backends; uhann_scan_info and
// uhann_workspace_info give this engine's device bytes.

#include "ultrahigh_ann/core/cuda_dense_l2_scan_index.hpp"

#include "uhann_adapter_support.hpp"

#include <array>
#include <mutex>
#include <utility>

namespace ultrahigh_ann {

namespace ua = uhann_adapter;

bool cuda_dense_l2_scan_available() noexcept { return uhann_device_available() != 0; }

struct CudaDenseL2ScanIndex::Impl {
    Impl(const DenseMatrix& reps, int dev) : device(dev), rows(reps.rows()), dim(reps.cols())
    {
        ua::check(uhann_scan_create(UHANN_METRIC_L2, reps.values().data(), rows, dim, nullptr, device, &scan));
        name = ua::device_name(device);
        payload = ua::product(rows, dim + 1, "CUDA dense L2 index") * sizeof(float);
    }
    Impl(const Impl&) = delete;
    Impl& operator=(const Impl&) = delete;
    ~Impl() { uhann_scan_destroy(scan); }

    void enable_gram()
    {
        std::lock_guard<std::mutex> lock(mu);
        if (!gram) {
            ua::check(uhann_scan_enable_mixed(scan, UHANN_STRATEGY_GRAM_INT8));
            gram = true;
        }
    }

    uhann_scan* scan = nullptr;
    int device = 0;
    std::size_t rows = 0;
    std::size_t dim = 0;
    std::size_t payload = 0;
    std::string name;
    std::mutex mu;
    bool gram = false;
};

struct CudaDenseL2ScanIndex::QueryWorkspace::Impl {
    Impl(const CudaDenseL2ScanIndex::Impl* o, std::size_t max_batch) : owner(o), capacity(max_batch)
    {
        if (max_batch == 0) throw std::invalid_argument("CUDA dense L2 workspace batch size must be positive");
        ua::check(uhann_scan_workspace_create(owner->scan, max_batch, &ws));
        // queries, distances and query norms (binary32), results (size_t)
        payload = (ua::product(max_batch, owner->dim, "CUDA dense L2 query workspace") +
                   ua::product(max_batch, owner->rows, "CUDA dense L2 distance workspace") + max_batch) *
                      sizeof(float) +
                  max_batch * sizeof(std::size_t);
    }
    Impl(const Impl&) = delete;
    Impl& operator=(const Impl&) = delete;
    ~Impl() { uhann_workspace_destroy(ws); }

    const CudaDenseL2ScanIndex::Impl* owner = nullptr;
    uhann_workspace* ws = nullptr;
    std::size_t capacity = 0;
    std::size_t payload = 0;
};

CudaDenseL2ScanIndex::QueryWorkspace::QueryWorkspace(std::unique_ptr<Impl> implementation)
    : implementation_(std::move(implementation))
{
}
CudaDenseL2ScanIndex::QueryWorkspace::QueryWorkspace(QueryWorkspace&&) noexcept = default;
CudaDenseL2ScanIndex::QueryWorkspace& CudaDenseL2ScanIndex::QueryWorkspace::operator=(QueryWorkspace&&) noexcept =
    default;
CudaDenseL2ScanIndex::QueryWorkspace::~QueryWorkspace() = default;

std::size_t CudaDenseL2ScanIndex::QueryWorkspace::maximum_batch_size() const noexcept
{
    return implementation_ ? implementation_->capacity : 0;
}

std::size_t CudaDenseL2ScanIndex::QueryWorkspace::payload_bytes() const noexcept
{
    return implementation_ ? implementation_->payload : 0;
}

CudaDenseL2ScanIndex::CudaDenseL2ScanIndex(const DenseMatrix& representatives, int device)
    : implementation_(std::make_unique<Impl>(representatives, device))
{
}
CudaDenseL2ScanIndex::CudaDenseL2ScanIndex(CudaDenseL2ScanIndex&&) noexcept = default;
CudaDenseL2ScanIndex& CudaDenseL2ScanIndex::operator=(CudaDenseL2ScanIndex&&) noexcept = default;
CudaDenseL2ScanIndex::~CudaDenseL2ScanIndex() = default;

CudaDenseL2ScanIndex::QueryWorkspace CudaDenseL2ScanIndex::make_query_workspace(std::size_t maximum_batch_size) const
{
    return QueryWorkspace(std::make_unique<QueryWorkspace::Impl>(implementation_.get(), maximum_batch_size));
}

std::size_t CudaDenseL2ScanIndex::query(std::span<const float> query, QueryWorkspace& workspace) const
{
    std::array<std::size_t, 1> out{};
    query_batch(query, 1, out, workspace);
    return out[0];
}

namespace {

// Workspace owner, capacity, output size: the reference's checks, in its
// order (cuda_dense_l2_scan_index.cu:598-620, :644-660).
void check_call(bool owned, std::size_t capacity, std::size_t count, std::size_t outputs)
{
    if (!owned) throw std::invalid_argument("CUDA dense L2 workspace belongs to a different index");
    if (count > capacity) throw std::invalid_argument("query count exceeds CUDA dense L2 workspace capacity");
    if (outputs != count) throw std::invalid_argument("CUDA dense L2 output size does not match query count");
}

int engine_strategy(CudaDenseL2QueryStrategy s)
{
    switch (s) {
    case CudaDenseL2QueryStrategy::direct: return UHANN_STRATEGY_DIRECT;
    case CudaDenseL2QueryStrategy::gemm: return UHANN_STRATEGY_GRAM_INT8;
    }
    throw std::invalid_argument("unknown CUDA dense L2 query strategy");
}

}  // namespace

void CudaDenseL2ScanIndex::query_batch(std::span<const float> queries, std::size_t query_count,
                                       std::span<std::size_t> output, QueryWorkspace& workspace,
                                       CudaDenseL2QueryStrategy strategy) const
{
    const auto* w = workspace.implementation_.get();
    const bool owned = w != nullptr && w->owner == implementation_.get();
    check_call(owned, owned ? w->capacity : 0, query_count, query_count);
    if (queries.size() != ua::product(query_count, implementation_->dim, "CUDA dense L2 query batch")) {
        throw std::invalid_argument("query batch dimensions do not match index dimension");
    }
    check_call(true, w->capacity, query_count, output.size());
    const int s = engine_strategy(strategy);
    if (s == UHANN_STRATEGY_GRAM_INT8) implementation_->enable_gram();
    ua::check(uhann_scan_query_batch(implementation_->scan, w->ws, queries.data(), query_count,
                                     ua::as_u64(output.data()), nullptr, s));
}

void CudaDenseL2ScanIndex::query_device_batch(const float* device_queries, std::size_t query_count,
                                              std::span<std::size_t> output, QueryWorkspace& workspace,
                                              CudaDenseL2QueryStrategy strategy) const
{
    const auto* w = workspace.implementation_.get();
    const bool owned = w != nullptr && w->owner == implementation_.get();
    check_call(owned, owned ? w->capacity : 0, query_count, output.size());
    if (query_count == 0) return;
    if (implementation_->dim != 0 && device_queries == nullptr) {
        throw std::invalid_argument("CUDA dense L2 device query pointer must not be null");
    }
    const int s = engine_strategy(strategy);
    if (s == UHANN_STRATEGY_GRAM_INT8) implementation_->enable_gram();
    ua::check(uhann_scan_query_device_batch(implementation_->scan, w->ws, device_queries, query_count,
                                            ua::as_u64(output.data()), nullptr, s));
}

IndexSpaceUsage CudaDenseL2ScanIndex::space_usage() const noexcept
{
    IndexSpaceUsage u{};
    u.index_payload_bytes = implementation_->payload;
    u.query_workspace_payload_bytes = 0;
    u.unique_query_coordinates = implementation_->dim;
    u.sampled_multiplicity = implementation_->dim;
    return u;
}

int CudaDenseL2ScanIndex::device() const noexcept { return implementation_->device; }

const std::string& CudaDenseL2ScanIndex::device_name() const noexcept { return implementation_->name; }

}  // namespace ultrahigh_ann
