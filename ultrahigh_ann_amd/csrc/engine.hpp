// Engine objects behind the C ABI handles (uhann_scan, uhann_workspace and
// the index handles). Internal to the library.
#pragma once

#include <hip/hip_runtime_api.h>

#include <cstdint>
#include <string>

#include "common.hpp"
#include "kernels.hpp"

namespace uhann {

const char* last_error() noexcept;
uint64_t padded_ld(uint64_t dim);
void validate_finite_host(const float* values, std::size_t n, const char* what);

// Dense FP32 references resident on one device, row-major with a leading
// dimension padded to a multiple of 4 floats.
struct Scan {
    Scan(int metric, const float* host_refs, uint64_t rows, uint64_t dim, const float* host_weights,
         int device);
    Scan(const Scan&) = delete;
    Scan& operator=(const Scan&) = delete;

    int metric = 0;
    int device = 0;
    uint64_t rows = 0;
    uint64_t dim = 0;
    uint64_t ld = 0;
    bool weighted = false;
    DeviceBuffer<float> refs;
    DeviceBuffer<float> weights;
    uint64_t payload_bytes = 0;
    std::string device_name;
};

// Per-caller device buffers and a private stream. `owner` identifies the
// handle that created it (workspaces are not interchangeable, as in the
// reference).
struct Workspace {
    Workspace(const void* owner, int device, uint64_t max_batch, uint64_t query_ld,
              uint64_t extra_floats_per_query);
    Workspace(const Workspace&) = delete;
    Workspace& operator=(const Workspace&) = delete;
    ~Workspace();

    const void* owner = nullptr;
    int device = 0;
    uint64_t max_batch = 0;
    DeviceBuffer<float> queries;   // staged full-width queries (host uploads)
    DeviceBuffer<float> extra;     // gathered / projected queries
    DeviceBuffer<unsigned long long> keys;
    DeviceBuffer<uint64_t> out_idx;
    DeviceBuffer<float> out_dist;
    DeviceBuffer<int> invalid;
    hipStream_t stream = nullptr;
    uint64_t payload_bytes = 0;
};

kern::ScanTile choose_tile(uint64_t nq, bool aligned4);

// fill keys -> direct scan -> finalize, all on `stream`. forced_tile (or
// null) overrides the tile choice.
void scan_launch(const Scan& scan, const float* queries, uint64_t ld_q, uint64_t count,
                 unsigned long long* keys, uint64_t* out_idx, float* out_dist, hipStream_t stream,
                 const kern::ScanTile* forced_tile);

}  // namespace uhann
