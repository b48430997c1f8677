// Stand-alone device components of the hierarchical indexes, as C ABI
// handles: the binary32 query projection (the reference's
// detail::CudaFloatProjection, src/core/cuda_float_projection.hpp; kernel
// K7, cuda_float_projection.cu:234-263) and the median-L1 scan (its
// detail::CudaMedianL1ScanIndex, src/hnsvw25/l1/hierarchical/
// cuda_median_l1_scan_index.hpp; K9/K10, cuda_median_l1_scan_index.cu:
// 202-301), plus device metadata (cuda_runtime_info.hpp). The hierarchical
// index handles run the same kernels on gathered queries; these handles let
// the reference's own facades compose them as they compose its CUDA classes.

#include <hip/hip_runtime_api.h>

#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "common.hpp"
#include "engine.hpp"
#include "hier.hpp"
#include "kernels.hpp"

using namespace uhann;

struct uhann_projection {
    int device = 0;
    uint64_t p = 0;  // projection dimension (rows of the matrix)
    uint64_t k = 0;  // sampled dimension (columns)
    DeviceBuffer<float> proj;
};

struct uhann_median {
    int device = 0;
    uint64_t rows = 0;
    uint64_t p = 0;
    DeviceBuffer<float> refs_t;  // p x rows, transposed for coalesced loads
};

namespace {

template <class H>
uhann_workspace* owned_ws(const H* h, uhann_workspace* ws, uint64_t count, const char* what)
{
    if (ws == nullptr || ws->owner != h) throw std::invalid_argument(std::string(what) + " workspace belongs to a different index");
    if (count > ws->max_batch) throw std::invalid_argument(std::string("query count exceeds ") + what + " workspace capacity");
    return ws;
}

int read_flag(Workspace& ws, int which)
{
    int host[4] = {0, 0, 0, 0};
    check_hip(hipMemcpy(host, ws.invalid.data(), sizeof(host), hipMemcpyDeviceToHost),
              "cannot read validation flags");
    return host[which];
}

}  // namespace

extern "C" {

int uhann_projection_create(const float* host_matrix, uint64_t p, uint64_t k, int device, uhann_projection** out)
{
    return guarded([&] {
        if (out == nullptr) throw std::invalid_argument("output handle must not be null");
        *out = nullptr;
        if (p == 0) throw std::invalid_argument("CUDA projection dimension must be positive");
        if (p > 0x7fffffffULL || k > 0x7fffffffULL) throw std::length_error("projection matrix dimensions exceed the 32-bit range");
        const std::size_t n = checked_mul(p, k, "projection matrix");
        if (n != 0 && host_matrix == nullptr) throw std::invalid_argument("projection matrix must not be null");
        for (std::size_t i = 0; i < n; ++i) {
            if (!std::isfinite(host_matrix[i])) throw std::overflow_error("projection matrix is not finite");
        }
        select_device(device);
        auto h = std::make_unique<uhann_projection>();
        h->device = device;
        h->p = p;
        h->k = k;
        h->proj = DeviceBuffer<float>(device, n);
        if (n != 0) {
            check_hip(hipMemcpy(h->proj.data(), host_matrix, n * sizeof(float), hipMemcpyHostToDevice),
                      "cannot copy the projection matrix to the HIP device");
        }
        *out = h.release();
    });
}

void uhann_projection_destroy(uhann_projection* h) { delete h; }

int uhann_projection_workspace_create(const uhann_projection* h, uint64_t max_batch, uhann_workspace** out)
{
    return guarded([&] {
        if (h == nullptr || out == nullptr) throw std::invalid_argument("null handle");
        *out = nullptr;
        // staged sampled queries (max_batch x k), projected queries (max_batch x p)
        *out = new uhann_workspace(h, h->device, max_batch, h->k, h->p);
    });
}

int uhann_projection_project(const uhann_projection* h, uhann_workspace* ws, const float* host_queries,
                             uint64_t count, const float** device_projected)
{
    return guarded([&] {
        if (h == nullptr) throw std::invalid_argument("projection handle must not be null");
        if (device_projected == nullptr) throw std::invalid_argument("output pointer must not be null");
        owned_ws(h, ws, count, "CUDA projection");
        const std::size_t n = checked_mul(count, h->k, "sampled-query projection batch");
        if (n != 0 && host_queries == nullptr) throw std::invalid_argument("sampled queries must not be null");
        validate_finite_host(host_queries, n, "sampled queries");
        *device_projected = ws->extra.data();
        if (count == 0) return;
        select_device(h->device);
        kern::zero_u32(reinterpret_cast<unsigned int*>(ws->invalid.data()), 4, ws->stream);
        if (h->k == 0) {
            // no sampled coordinates: every projection is +0 (a fill kernel)
            const uint64_t words = checked_mul(count, h->p, "projected-query batch");
            if (words > 0xffffffffULL) throw std::length_error("projected-query batch exceeds the 32-bit range");
            kern::zero_u32(reinterpret_cast<unsigned int*>(ws->extra.data()), static_cast<uint32_t>(words), ws->stream);
        } else {
            check_hip(hipMemcpyAsync(ws->queries.data(), host_queries, n * sizeof(float), hipMemcpyHostToDevice,
                                     ws->stream),
                      "cannot copy sampled queries to the HIP device");
            kern::project_fp32(ws->queries.data(), h->k, static_cast<uint32_t>(count), static_cast<uint32_t>(h->k),
                               h->proj.data(), h->k, static_cast<uint32_t>(h->p), ws->extra.data(), h->p,
                               ws->invalid.data(), ws->stream);
        }
        check_hip(hipStreamSynchronize(ws->stream), "HIP projection failed");
        if (read_flag(*ws, 0) != 0) throw std::overflow_error("projected query is not finite");
    });
}

int uhann_projection_info(const uhann_projection* h, uint64_t* projection_dim, uint64_t* sampled_dim,
                          uint64_t* payload_bytes, int* device)
{
    return guarded([&] {
        if (h == nullptr) throw std::invalid_argument("projection handle must not be null");
        if (projection_dim) *projection_dim = h->p;
        if (sampled_dim) *sampled_dim = h->k;
        if (payload_bytes) *payload_bytes = h->proj.bytes();
        if (device) *device = h->device;
    });
}

int uhann_median_create(const float* host_projected_refs, uint64_t rows, uint64_t p, int device, uhann_median** out)
{
    return guarded([&] {
        if (out == nullptr) throw std::invalid_argument("output handle must not be null");
        *out = nullptr;
        if (rows == 0) throw std::invalid_argument("CUDA median L1 scan expects at least one representative");
        if (p == 0) throw std::invalid_argument("CUDA median L1 projection dimension must be positive");
        if (rows > 0x7fffffffULL) throw std::length_error("median L1 scan supports fewer than 2^31 rows");
        select_device(device);
        if (!kern::median_supported(p)) {
            throw std::invalid_argument(
                "CUDA median L1 projection dimension exceeds the device's per-block shared-memory capacity");
        }
        const std::size_t n = checked_mul(rows, p, "projected representatives");
        if (host_projected_refs == nullptr) throw std::invalid_argument("projected representatives must not be null");
        for (std::size_t i = 0; i < n; ++i) {
            if (!std::isfinite(host_projected_refs[i])) throw std::overflow_error("projected representatives is not finite");
        }
        select_device(device);
        auto h = std::make_unique<uhann_median>();
        h->device = device;
        h->rows = rows;
        h->p = p;
        std::vector<float> t(n);
        for (uint64_t r = 0; r < rows; ++r) {
            for (uint64_t c = 0; c < p; ++c) t[c * rows + r] = host_projected_refs[r * p + c];
        }
        h->refs_t = DeviceBuffer<float>(device, n);
        check_hip(hipMemcpy(h->refs_t.data(), t.data(), n * sizeof(float), hipMemcpyHostToDevice),
                  "cannot copy projected representatives to the HIP device");
        *out = h.release();
    });
}

void uhann_median_destroy(uhann_median* h) { delete h; }

int uhann_median_workspace_create(const uhann_median* h, uint64_t max_batch, uhann_workspace** out)
{
    return guarded([&] {
        if (h == nullptr || out == nullptr) throw std::invalid_argument("null handle");
        *out = nullptr;
        *out = new uhann_workspace(h, h->device, max_batch, 0, 0);
    });
}

int uhann_median_query_device_batch(const uhann_median* h, uhann_workspace* ws, const float* device_projected,
                                    uint64_t count, uint64_t* out_idx)
{
    return guarded([&] {
        if (h == nullptr) throw std::invalid_argument("median handle must not be null");
        owned_ws(h, ws, count, "CUDA median L1");
        if (count == 0) return;
        if (device_projected == nullptr) throw std::invalid_argument("CUDA median L1 projected-query pointer must not be null");
        if (out_idx == nullptr) throw std::invalid_argument("output indices must not be null");
        select_device(h->device);
        const uint32_t n = static_cast<uint32_t>(count);
        kern::zero_u32(reinterpret_cast<unsigned int*>(ws->invalid.data()), 4, ws->stream);
        kern::fill_keys(ws->keys.data(), n, ws->stream);
        kern::median_l1(h->refs_t.data(), static_cast<uint32_t>(h->rows), device_projected, h->p, n,
                        static_cast<uint32_t>(h->p), ws->keys.data(), ws->invalid.data(), ws->stream);
        kern::finalize_keys(ws->keys.data(), n, ws->out_idx.data(), nullptr, ws->stream);
        check_hip(hipMemcpyAsync(out_idx, ws->out_idx.data(), count * sizeof(uint64_t), hipMemcpyDeviceToHost,
                                 ws->stream),
                  "cannot copy median L1 results from the HIP device");
        check_hip(hipStreamSynchronize(ws->stream), "HIP median L1 scan failed");
        if (read_flag(*ws, 0) != 0) throw std::overflow_error("CUDA median L1 distance is not finite");
    });
}

int uhann_median_info(const uhann_median* h, uint64_t* rows, uint64_t* projection_dim, uint64_t* payload_bytes,
                      int* device)
{
    return guarded([&] {
        if (h == nullptr) throw std::invalid_argument("median handle must not be null");
        if (rows) *rows = h->rows;
        if (projection_dim) *projection_dim = h->p;
        if (payload_bytes) *payload_bytes = h->refs_t.bytes();
        if (device) *device = h->device;
    });
}

int uhann_device_properties(int device, char* name, size_t name_cap, int* arch_major, int* arch_minor,
                            uint64_t* total_memory_bytes, int* compiled_runtime_version, int* runtime_version,
                            int* driver_version)
{
    return guarded([&] {
        select_device(device);
        hipDeviceProp_t prop{};
        check_hip(hipGetDeviceProperties(&prop, device), "cannot read HIP device properties");
        if (name != nullptr && name_cap > 0) {
            std::strncpy(name, device_label(prop).c_str(), name_cap - 1);
            name[name_cap - 1] = '\0';
        }
        if (arch_major) *arch_major = prop.major;
        if (arch_minor) *arch_minor = prop.minor;
        if (total_memory_bytes) *total_memory_bytes = prop.totalGlobalMem;
        if (compiled_runtime_version) *compiled_runtime_version = HIP_VERSION;
        int rt = 0, drv = 0;
        check_hip(hipRuntimeGetVersion(&rt), "cannot read the HIP runtime version");
        check_hip(hipDriverGetVersion(&drv), "cannot read the HIP driver version");
        if (runtime_version) *runtime_version = rt;
        if (driver_version) *driver_version = drv;
    });
}

}  // extern "C"
