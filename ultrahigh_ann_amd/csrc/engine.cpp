// Host side of the dense scan engine behind include/uhann.h: device
// selection, reference upload, workspaces, validation in the reference's
// order, and the launch sequence fill-keys -> direct scan -> finalize.
//
// Reference behaviour mirrored here: cuda_dense_l2_scan_index.cu:407-751 and
// cuda_dense_l1_scan_index.cu:290-535 (constructor checks, workspace
// ownership and capacity checks, finite-query validation before any device
// work, count == 0 returns early, dim == 0 selects row 0).

#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "common.hpp"
#include "engine.hpp"
#include "kernels.hpp"

namespace uhann {

namespace {
thread_local std::string g_last_error;
}

void set_last_error(std::string message) noexcept
{
    try {
        g_last_error = std::move(message);
    } catch (...) {
    }
}

const char* last_error() noexcept { return g_last_error.c_str(); }

uint64_t padded_ld(uint64_t dim) { return (dim + 3U) / 4U * 4U; }

void validate_finite_host(const float* values, std::size_t n, const char* what)
{
    for (std::size_t i = 0; i < n; ++i) {
        if (!std::isfinite(values[i])) {
            throw std::invalid_argument(std::string(what) + " must be finite");
        }
    }
}

Scan::Scan(int metric_value, const float* host_refs, uint64_t rows_value, uint64_t dim_value,
           const float* host_weights, int device_value)
    : metric(metric_value), device(device_value), rows(rows_value), dim(dim_value),
      ld(padded_ld(dim_value))
{
    if (metric != UHANN_METRIC_L2 && metric != UHANN_METRIC_L1) {
        throw std::invalid_argument("unknown metric");
    }
    select_device(device);
    hipDeviceProp_t prop{};
    check_hip(hipGetDeviceProperties(&prop, device), "cannot query HIP device properties");
    device_name = prop.name;
    if (rows == 0) {
        throw std::invalid_argument("dense scan expects at least one representative");
    }
    if (rows > 0xffffffffULL || dim > 0x3fffffffULL) {
        throw std::length_error("dense scan dimensions exceed the 32-bit index range");
    }
    if (host_refs == nullptr && dim != 0) {
        throw std::invalid_argument("representatives must not be null");
    }
    if (host_weights != nullptr && metric != UHANN_METRIC_L1) {
        throw std::invalid_argument("coordinate weights apply to the L1 metric only");
    }
    const std::size_t count = checked_mul(rows, dim, "dense scan representatives");
    validate_finite_host(host_refs, count, "representatives");
    if (host_weights != nullptr) {
        for (uint64_t c = 0; c < dim; ++c) {
            if (!std::isfinite(host_weights[c]) || host_weights[c] <= 0.0F) {
                throw std::invalid_argument("L1 coordinate weights must be positive and finite");
            }
        }
    }
    refs = DeviceBuffer<float>(device, checked_mul(rows, ld, "dense scan representatives"));
    if (refs.bytes() != 0) {
        // Row-major with a padded leading dimension (multiple of 4 floats) so
        // the kernels can use 16-byte loads; pad columns are zero and never read.
        std::vector<float> staged(static_cast<std::size_t>(rows * ld), 0.0F);
        for (uint64_t r = 0; r < rows; ++r) {
            std::memcpy(staged.data() + r * ld, host_refs + r * dim, dim * sizeof(float));
        }
        check_hip(hipMemcpy(refs.data(), staged.data(), refs.bytes(), hipMemcpyHostToDevice),
                  "cannot copy representatives to the HIP device");
    }
    if (host_weights != nullptr && dim != 0) {
        weights = DeviceBuffer<float>(device, ld);
        std::vector<float> staged(static_cast<std::size_t>(ld), 0.0F);
        std::memcpy(staged.data(), host_weights, dim * sizeof(float));
        check_hip(hipMemcpy(weights.data(), staged.data(), weights.bytes(), hipMemcpyHostToDevice),
                  "cannot copy L1 coordinate weights to the HIP device");
        weighted = true;
    }
    // The reference keeps rows*dim floats (+ rows norms for L2, + dim weights
    // for L1); report the same logical payload.
    payload_bytes = count * sizeof(float) + (metric == UHANN_METRIC_L2 ? rows * sizeof(float) : 0) +
                    (weighted ? dim * sizeof(float) : 0);
}

Workspace::Workspace(const void* owner_value, int device_value, uint64_t max_batch_value,
                     uint64_t query_ld, uint64_t extra_floats_per_query)
    : owner(owner_value), device(device_value), max_batch(max_batch_value)
{
    if (max_batch == 0) {
        throw std::invalid_argument("workspace batch size must be positive");
    }
    if (max_batch > 0xffffffffULL) {
        throw std::length_error("workspace batch size exceeds the 32-bit range");
    }
    select_device(device);
    queries = DeviceBuffer<float>(device, checked_mul(max_batch, query_ld, "query workspace"));
    extra = DeviceBuffer<float>(device, checked_mul(max_batch, extra_floats_per_query,
                                                    "query workspace"));
    keys = DeviceBuffer<unsigned long long>(device, max_batch);
    out_idx = DeviceBuffer<uint64_t>(device, max_batch);
    out_dist = DeviceBuffer<float>(device, max_batch);
    invalid = DeviceBuffer<int>(device, 1);
    check_hip(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking),
              "cannot create a HIP stream");
    payload_bytes = queries.bytes() + extra.bytes() + keys.bytes() + out_idx.bytes() +
                    out_dist.bytes() + invalid.bytes();
}

Workspace::~Workspace()
{
    if (stream != nullptr) {
        static_cast<void>(hipSetDevice(device));
        static_cast<void>(hipStreamSynchronize(stream));
        static_cast<void>(hipStreamDestroy(stream));
    }
}

kern::ScanTile choose_tile(uint64_t nq, bool aligned4)
{
    kern::ScanTile t;
    if (nq <= 1) {
        t = aligned4 ? kern::ScanTile{4, 1, 8, 1} : kern::ScanTile{1, 1, 16, 1};
    } else if (nq <= 2) {
        t = aligned4 ? kern::ScanTile{4, 2, 8, 1} : kern::ScanTile{1, 2, 16, 1};
    } else if (nq <= 4) {
        t = aligned4 ? kern::ScanTile{4, 4, 8, 1} : kern::ScanTile{1, 4, 16, 1};
    } else if (nq <= 16) {
        t = kern::ScanTile{1, 8, 16, 2};
    } else {
        t = kern::ScanTile{1, 12, 12, 8};
    }
    return t;
}

void scan_launch(const Scan& scan, const float* queries, uint64_t ld_q, uint64_t count,
                 unsigned long long* keys, uint64_t* out_idx, float* out_dist,
                 hipStream_t stream, const kern::ScanTile* forced_tile)
{
    kern::ScanArgs a;
    a.refs = scan.refs.data();
    a.ld_ref = scan.ld;
    a.rows = static_cast<uint32_t>(scan.rows);
    a.queries = queries;
    a.ld_q = ld_q;
    a.nq = static_cast<uint32_t>(count);
    a.weights = scan.weighted ? scan.weights.data() : nullptr;
    a.dim = static_cast<uint32_t>(scan.dim);
    a.keys = keys;
    const bool aligned4 = (ld_q % 4U == 0) && (scan.ld % 4U == 0) &&
                          (reinterpret_cast<uintptr_t>(queries) % 16U == 0);
    kern::ScanTile tile = forced_tile != nullptr ? *forced_tile : choose_tile(count, aligned4);
    if (tile.rpt > 1 && !aligned4) {
        throw std::invalid_argument("vector scan tile needs 16-byte aligned query rows");
    }
    if (a.dim == 0) {
        // No coordinates: every distance is zero and the reference returns row 0.
        a.queries = scan.refs.data();
        a.ld_q = 0;
    }
    kern::fill_keys(keys, a.nq, stream);
    kern::scan_direct(scan.metric == UHANN_METRIC_L2 ? kern::Metric::l2 : kern::Metric::l1,
                      /*weight_fma=*/true, a, tile, stream);
    kern::finalize_keys(keys, a.nq, out_idx, out_dist, stream);
}

}  // namespace uhann

using namespace uhann;

struct uhann_scan : Scan {
    using Scan::Scan;
};

struct uhann_workspace : Workspace {
    using Workspace::Workspace;
};

namespace {

uhann_workspace* checked_ws(const void* owner, uhann_workspace* ws, uint64_t count)
{
    if (ws == nullptr || ws->owner != owner) {
        throw std::invalid_argument("workspace belongs to a different index");
    }
    if (count > ws->max_batch) {
        throw std::invalid_argument("query count exceeds workspace capacity");
    }
    return ws;
}

void copy_results(Workspace& ws, uint64_t count, uint64_t* out_idx, float* out_dist)
{
    if (out_idx == nullptr) throw std::invalid_argument("output indices must not be null");
    check_hip(hipMemcpyAsync(out_idx, ws.out_idx.data(), count * sizeof(uint64_t),
                             hipMemcpyDeviceToHost, ws.stream),
              "cannot copy query results from the HIP device");
    if (out_dist != nullptr) {
        check_hip(hipMemcpyAsync(out_dist, ws.out_dist.data(), count * sizeof(float),
                                 hipMemcpyDeviceToHost, ws.stream),
                  "cannot copy query distances from the HIP device");
    }
    check_hip(hipStreamSynchronize(ws.stream), "HIP query failed");
}

}  // namespace

extern "C" {

const char* uhann_last_error(void) { return uhann::last_error(); }

int uhann_device_available(void)
{
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess) {
        static_cast<void>(hipGetLastError());
        return 0;
    }
    return count > 0 ? 1 : 0;
}

int uhann_device_name(int device, char* buf, size_t cap)
{
    return guarded([&] {
        select_device(device);
        hipDeviceProp_t prop{};
        check_hip(hipGetDeviceProperties(&prop, device), "cannot query HIP device properties");
        if (buf == nullptr || cap == 0) throw std::invalid_argument("name buffer is empty");
        std::strncpy(buf, prop.name, cap - 1);
        buf[cap - 1] = '\0';
    });
}

int uhann_scan_create(int metric, const float* host_refs, uint64_t rows, uint64_t dim,
                      const float* weights, int device, uhann_scan** out)
{
    return guarded([&] {
        if (out == nullptr) throw std::invalid_argument("output handle must not be null");
        *out = nullptr;
        *out = new uhann_scan(metric, host_refs, rows, dim, weights, device);
    });
}

void uhann_scan_destroy(uhann_scan* scan) { delete scan; }

int uhann_scan_info(const uhann_scan* scan, uint64_t* rows, uint64_t* dim,
                    uint64_t* payload_bytes, int* device)
{
    return guarded([&] {
        if (scan == nullptr) throw std::invalid_argument("scan handle must not be null");
        if (rows) *rows = scan->rows;
        if (dim) *dim = scan->dim;
        if (payload_bytes) *payload_bytes = scan->payload_bytes;
        if (device) *device = scan->device;
    });
}

int uhann_scan_workspace_create(const uhann_scan* scan, uint64_t max_batch, uhann_workspace** out)
{
    return guarded([&] {
        if (scan == nullptr || out == nullptr) throw std::invalid_argument("null handle");
        *out = nullptr;
        *out = new uhann_workspace(scan, scan->device, max_batch, scan->ld, 0);
    });
}

void uhann_workspace_destroy(uhann_workspace* ws) { delete ws; }

int uhann_workspace_info(const uhann_workspace* ws, uint64_t* max_batch, uint64_t* payload_bytes)
{
    return guarded([&] {
        if (ws == nullptr) throw std::invalid_argument("workspace handle must not be null");
        if (max_batch) *max_batch = ws->max_batch;
        if (payload_bytes) *payload_bytes = ws->payload_bytes;
    });
}

int uhann_scan_query_batch(const uhann_scan* scan, uhann_workspace* ws, const float* host_queries,
                           uint64_t count, uint64_t* out_idx, float* out_dist, int strategy)
{
    return guarded([&] {
        if (scan == nullptr) throw std::invalid_argument("scan handle must not be null");
        checked_ws(scan, ws, count);
        if (strategy != UHANN_STRATEGY_DIRECT) throw std::invalid_argument("unknown strategy");
        const std::size_t n = checked_mul(count, scan->dim, "query batch");
        if (n != 0 && host_queries == nullptr) throw std::invalid_argument("queries must not be null");
        validate_finite_host(host_queries, n, "queries");
        if (count == 0) return;
        select_device(scan->device);
        if (n != 0) {
            check_hip(hipMemcpy2DAsync(ws->queries.data(), scan->ld * sizeof(float), host_queries,
                                       scan->dim * sizeof(float), scan->dim * sizeof(float), count,
                                       hipMemcpyHostToDevice, ws->stream),
                      "cannot copy queries to the HIP device");
        }
        scan_launch(*scan, ws->queries.data(), scan->ld, count, ws->keys.data(), ws->out_idx.data(),
                    ws->out_dist.data(), ws->stream, nullptr);
        copy_results(*ws, count, out_idx, out_dist);
    });
}

int uhann_scan_query_device_batch(const uhann_scan* scan, uhann_workspace* ws,
                                  const float* device_queries, uint64_t count, uint64_t* out_idx,
                                  float* out_dist, int strategy)
{
    return guarded([&] {
        if (scan == nullptr) throw std::invalid_argument("scan handle must not be null");
        checked_ws(scan, ws, count);
        if (strategy != UHANN_STRATEGY_DIRECT) throw std::invalid_argument("unknown strategy");
        if (count == 0) return;
        if (scan->dim != 0 && device_queries == nullptr) {
            throw std::invalid_argument("device query pointer must not be null");
        }
        select_device(scan->device);
        scan_launch(*scan, device_queries, scan->dim, count, ws->keys.data(), ws->out_idx.data(),
                    ws->out_dist.data(), ws->stream, nullptr);
        copy_results(*ws, count, out_idx, out_dist);
    });
}

int uhann_scan_query_device_async(const uhann_scan* scan, uhann_workspace* ws,
                                  const float* device_queries, uint64_t ld_queries,
                                  uint64_t count, uint64_t* device_out_idx,
                                  float* device_out_dist, int strategy, void* stream)
{
    return guarded([&] {
        if (scan == nullptr) throw std::invalid_argument("scan handle must not be null");
        checked_ws(scan, ws, count);
        if (strategy != UHANN_STRATEGY_DIRECT) throw std::invalid_argument("unknown strategy");
        if (count == 0) return;
        if (scan->dim != 0 && device_queries == nullptr) {
            throw std::invalid_argument("device query pointer must not be null");
        }
        if (ld_queries < scan->dim) throw std::invalid_argument("query row stride is below dim");
        select_device(scan->device);
        scan_launch(*scan, device_queries, ld_queries, count, ws->keys.data(), device_out_idx,
                    device_out_dist, static_cast<hipStream_t>(stream), nullptr);
    });
}

int uhann_scan_query_device_tile(const uhann_scan* scan, uhann_workspace* ws,
                                 const float* device_queries, uint64_t ld_queries, uint64_t count,
                                 uint64_t* device_out_idx, float* device_out_dist, int rpt, int tq,
                                 int tr, int group_q, void* stream)
{
    return guarded([&] {
        if (scan == nullptr) throw std::invalid_argument("scan handle must not be null");
        checked_ws(scan, ws, count);
        const kern::ScanTile tile{rpt, tq, tr, group_q};
        if (!kern::scan_tile_supported(tile)) throw std::invalid_argument("unsupported scan tile");
        if (count == 0) return;
        if (scan->dim != 0 && device_queries == nullptr) {
            throw std::invalid_argument("device query pointer must not be null");
        }
        if (ld_queries < scan->dim) throw std::invalid_argument("query row stride is below dim");
        select_device(scan->device);
        scan_launch(*scan, device_queries, ld_queries, count, ws->keys.data(), device_out_idx,
                    device_out_dist, static_cast<hipStream_t>(stream), &tile);
    });
}

}  // extern "C"
