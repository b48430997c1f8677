// Index handles of the C ABI: exact, flat (importance-sampled coordinates)
// and hierarchical (sampled coordinates projected to p dimensions) indexes,
// L2 or L1, built from host references and an explicit coordinate sample /
// projection matrix, queried from host or device memory.
//
// Reference behaviour mirrored (paths relative to the reference checkout):
//   exact        cuda_exact_l{1,2}_index.cpp -> dense scans
//   flat L2      cuda_sampled_coordinate_l2_ann_index.cpp:55-227 with the
//                embedding of l2_coordinate_embedding.cpp:29-127
//   flat L1      cuda_sampled_coordinate_l1_ann_index.cpp:55-196 with
//                l1_coordinate_embedding.cpp:29-111 (float weights m/p)
//   hier. L2     cuda_hierarchical_l2_ann_index.cpp:128-377
//   hier. L1     cuda_hierarchical_l1_ann_index.cpp:292-340 + median scan
// The reference packs sampled query coordinates on the host; here the gather
// runs on the device from full-width device-resident queries. A non-finite
// sampled coordinate raises invalid_argument and a non-finite projection or
// median raises overflow_error, as in the reference.

#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "common.hpp"
#include "engine.hpp"
#include "hier.hpp"
#include "index.hpp"
#include "kernels.hpp"
#include "ultrahigh_ann_amd/build.hpp"

namespace uab = ultrahigh_ann_amd;
using namespace uhann;


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

// binary32 -> bf16 bits, round to nearest even (finite inputs).
uint16_t bf16_bits(float f)
{
    uint32_t b;
    std::memcpy(&b, &f, sizeof b);
    b += 0x7fffU + ((b >> 16) & 1U);
    return static_cast<uint16_t>(b >> 16);
}

void check_metric(int metric)
{
    if (metric != UHANN_METRIC_L2 && metric != UHANN_METRIC_L1) throw std::invalid_argument("unknown metric");
}

void upload_cols(uhann_index& ix, const uab::CoordinateSample& s)
{
    if (s.columns.empty()) return;
    std::vector<uint32_t> c(s.columns.size());
    for (std::size_t i = 0; i < c.size(); ++i) c[i] = static_cast<uint32_t>(s.columns[i].source_column);
    ix.cols = DeviceBuffer<uint32_t>(ix.device, c.size());
    check_hip(hipMemcpy(ix.cols.data(), c.data(), ix.cols.bytes(), hipMemcpyHostToDevice),
              "cannot copy sampled columns to the HIP device");
}

std::vector<float> sampled_refs(const float* refs, uint64_t rows, uint64_t dim,
                                const uab::CoordinateSample& s, const std::vector<float>* scales)
{
    const std::size_t k = s.columns.size();
    std::vector<float> out(checked_mul(rows, k, "sampled representatives"));
    for (uint64_t r = 0; r < rows; ++r) {
        for (std::size_t j = 0; j < k; ++j) {
            const float v = refs[r * dim + s.columns[j].source_column];
            out[r * k + j] = scales != nullptr ? v * (*scales)[j] : v;
        }
    }
    return out;
}

void common_checks(const float* refs, uint64_t rows, uint64_t dim, int device)
{
    select_device(device);
    if (rows == 0) throw std::invalid_argument("the index expects at least one representative");
    if (rows * dim != 0 && refs == nullptr) throw std::invalid_argument("representatives must not be null");
    if (dim > 0xffffffffULL) throw std::length_error("dimension exceeds the 32-bit column range");
}

uhann_workspace* checked_ws(const uhann_index* ix, uhann_workspace* ws, uint64_t count)
{
    if (ws == nullptr || ws->owner != ix) throw std::invalid_argument("workspace belongs to a different index");
    if (count > ws->max_batch) throw std::invalid_argument("query count exceeds workspace capacity");
    return ws;
}

// One batch of `count` device queries (row stride ld_q) -> device outputs.
// Data errors are left in ws.invalid[0..2]: sampled value, projection, median.
void index_launch(const uhann_index& ix, Workspace& ws, const float* queries, uint64_t ld_q,
                  bool padded, uint64_t count, uint64_t* out_idx, float* out_dist, int strategy,
                  hipStream_t stream)
{
    const uint32_t n = static_cast<uint32_t>(count);
    if (ix.kind == UHANN_KIND_EXACT) {
        run_strategy(*ix.scan, ws, queries, ld_q, padded, count, out_idx, out_dist, stream, strategy);
        return;
    }
    int* flags = ws.invalid.data();
    float* gathered = ws.extra.data();
    if (ix.kind == UHANN_KIND_FLAT) {
        kern::gather_columns(queries, ld_q, n, ix.cols.data(), ix.scales.data(), static_cast<uint32_t>(ix.k),
                             gathered, ix.k_ld, flags, stream);
        run_strategy(*ix.scan, ws, gathered, ix.k_ld, true, count, out_idx, out_dist, stream, strategy);
        return;
    }
    // hierarchical
    const bool bf16 = strategy == UHANN_STRATEGY_PROJECTION_BF16;
    if (strategy != UHANN_STRATEGY_DIRECT && !(bf16 && ix.metric == UHANN_METRIC_L2)) {
        throw std::invalid_argument(ix.metric == UHANN_METRIC_L2
                                        ? "hierarchical L2 indexes support the direct and bf16 projection strategies"
                                        : "hierarchical L1 indexes support the direct strategy only");
    }
    if (ix.rows == 1) {
        kern::zero_results(out_idx, out_dist, n, stream);
        return;
    }
    float* projected = gathered + count * ix.k_ld;
    kern::gather_columns(queries, ld_q, n, ix.cols.data(), nullptr, static_cast<uint32_t>(ix.k), gathered,
                         ix.k_ld, flags, stream);
    if (bf16) {
        kern::project_bf16(gathered, ix.k_ld, n, static_cast<uint32_t>(ix.k), ix.proj16.data(), ix.proj16_ld,
                           static_cast<uint32_t>(ix.p), projected, ix.p_ld, flags + 1, stream);
    } else {
        kern::project_fp32(gathered, ix.k_ld, n, static_cast<uint32_t>(ix.k), ix.proj.data(), ix.k_ld,
                           static_cast<uint32_t>(ix.p), projected, ix.p_ld, flags + 1, stream);
    }
    if (ix.metric == UHANN_METRIC_L2) {
        scan_launch(*ix.scan, projected, ix.p_ld, count, ws.keys.data(), out_idx, out_dist, stream, true,
                    nullptr);
    } else {
        kern::fill_keys(ws.keys.data(), n, stream);
        kern::median_l1(ix.refs_t.data(), static_cast<uint32_t>(ix.rows), projected, ix.p_ld, n,
                        static_cast<uint32_t>(ix.p), ws.keys.data(), flags + 2, stream);
        kern::finalize_keys(ws.keys.data(), n, out_idx, out_dist, stream);
    }
}

void clear_flags(Workspace& ws, hipStream_t stream)
{
    kern::zero_u32(reinterpret_cast<unsigned int*>(ws.invalid.data()), 4, stream);
}

void raise_flags(Workspace& ws)
{
    int host[4] = {0, 0, 0, 0};
    check_hip(hipMemcpy(host, ws.invalid.data(), sizeof(host), hipMemcpyDeviceToHost),
              "cannot read query validation flags");
    if (host[0] != 0) throw std::invalid_argument("sampled query coordinate must be finite");
    if (host[1] != 0) throw std::overflow_error("projected query is not finite");
    if (host[2] != 0) throw std::overflow_error("median L1 distance is not finite");
}

void sync_query(const uhann_index* ix, uhann_workspace* ws, const float* queries, uint64_t ld_q,
                bool padded, uint64_t count, uint64_t* out_idx, float* out_dist, int strategy)
{
    if (out_idx == nullptr) throw std::invalid_argument("output indices must not be null");
    clear_flags(*ws, ws->stream);
    index_launch(*ix, *ws, queries, ld_q, padded, count, ws->out_idx.data(), ws->out_dist.data(), strategy,
                 ws->stream);
    check_hip(hipMemcpyAsync(out_idx, ws->out_idx.data(), count * sizeof(uint64_t), hipMemcpyDeviceToHost,
                             ws->stream),
              "cannot copy query results from the HIP device");
    if (out_dist != nullptr) {
        check_hip(hipMemcpyAsync(out_dist, ws->out_dist.data(), count * sizeof(float), hipMemcpyDeviceToHost,
                                 ws->stream),
                  "cannot copy query distances from the HIP device");
    }
    check_hip(hipStreamSynchronize(ws->stream), "HIP query failed");
    raise_flags(*ws);
}

}  // namespace

extern "C" {

int uhann_exact_create(int metric, const float* host_refs, uint64_t rows, uint64_t dim, int device,
                       uhann_index** out)
{
    return guarded([&] {
        if (out == nullptr) throw std::invalid_argument("output handle must not be null");
        *out = nullptr;
        check_metric(metric);
        auto ix = std::make_unique<uhann_index>();
        ix->kind = UHANN_KIND_EXACT;
        ix->metric = metric;
        ix->device = device;
        ix->rows = rows;
        ix->dim = dim;
        ix->k = dim;
        ix->multiplicity = dim;
        ix->scan = std::make_unique<Scan>(metric, host_refs, rows, dim, nullptr, device);
        ix->payload_bytes = ix->scan->payload_bytes;
        *out = ix.release();
    });
}

int uhann_flat_create(int metric, const float* host_refs, uint64_t rows, uint64_t dim,
                      const uhann_sampled_column* sample, uint64_t sample_count, int device,
                      uhann_index** out)
{
    return guarded([&] {
        if (out == nullptr) throw std::invalid_argument("output handle must not be null");
        *out = nullptr;
        check_metric(metric);
        common_checks(host_refs, rows, dim, device);
        const auto s = to_sample(sample, sample_count);
        auto ix = std::make_unique<uhann_index>();
        ix->kind = UHANN_KIND_FLAT;
        ix->metric = metric;
        ix->device = device;
        ix->rows = rows;
        ix->dim = dim;
        ix->multiplicity = uab::validate_sample(s, dim);
        ix->k = s.columns.size();
        ix->k_ld = padded_ld(ix->k);
        const bool l1 = metric == UHANN_METRIC_L1;
        const std::vector<float> factors = uab::flat_factors(s, l1);
        // L2: references pre-scaled by sqrt(m/p) (binary32 multiply); L1: raw
        // sampled references with per-coordinate weights m/p in the scan.
        const std::vector<float> reduced = sampled_refs(host_refs, rows, dim, s, l1 ? nullptr : &factors);
        ix->scan = std::make_unique<Scan>(metric, reduced.data(), rows, ix->k,
                                          (l1 && ix->k != 0) ? factors.data() : nullptr, device);
        upload_cols(*ix, s);
        if (!l1 && ix->k != 0) {
            ix->scales = DeviceBuffer<float>(device, ix->k);
            check_hip(hipMemcpy(ix->scales.data(), factors.data(), ix->scales.bytes(), hipMemcpyHostToDevice),
                      "cannot copy coordinate scales to the HIP device");
        }
        // Reference accounting: scan payload + size_t columns (+ float scales for L2).
        ix->payload_bytes = ix->scan->payload_bytes + ix->k * sizeof(uint64_t) + (l1 ? 0 : ix->k * sizeof(float));
        *out = ix.release();
    });
}

int uhann_hierarchical_create(int metric, const float* host_refs, uint64_t rows, uint64_t dim,
                              const uhann_sampled_column* sample, uint64_t sample_count,
                              const double* projection, uint64_t projection_dim, int device,
                              uhann_index** out)
{
    return guarded([&] {
        if (out == nullptr) throw std::invalid_argument("output handle must not be null");
        *out = nullptr;
        check_metric(metric);
        common_checks(host_refs, rows, dim, device);
        if (projection_dim == 0) throw std::invalid_argument("projection_dimension has to be a positive integer");
        const auto s = to_sample(sample, sample_count);
        if (s.columns.empty() && rows > 1) throw std::runtime_error("importance sampling selected no coordinates");
        auto ix = std::make_unique<uhann_index>();
        ix->kind = UHANN_KIND_HIERARCHICAL;
        ix->metric = metric;
        ix->device = device;
        ix->rows = rows;
        ix->dim = dim;
        ix->multiplicity = uab::validate_sample(s, dim);
        ix->k = s.columns.size();
        ix->k_ld = padded_ld(ix->k);
        ix->p = projection_dim;
        ix->p_ld = padded_ld(projection_dim);
        if (ix->k != 0 && projection == nullptr) throw std::invalid_argument("projection must not be null");
        if (metric == UHANN_METRIC_L1) {
            select_device(device);
            if (!kern::median_supported(projection_dim)) {
                throw std::invalid_argument(
                    "CUDA median L1 projection dimension exceeds the device's per-block shared-memory capacity");
            }
        }
        std::vector<float> proj32;
        std::vector<float> projected;
        uab::fp32_hierarchy(std::span<const float>(host_refs, rows * dim), rows, dim, s,
                            std::span<const double>(projection, projection_dim * ix->k), projection_dim, proj32,
                            projected);
        upload_cols(*ix, s);
        if (ix->k != 0) {
            std::vector<float> staged(checked_mul(projection_dim, ix->k_ld, "projection"), 0.0F);
            for (uint64_t o = 0; o < projection_dim; ++o) {
                std::memcpy(staged.data() + o * ix->k_ld, proj32.data() + o * ix->k, ix->k * sizeof(float));
            }
            ix->proj = DeviceBuffer<float>(device, staged.size());
            check_hip(hipMemcpy(ix->proj.data(), staged.data(), ix->proj.bytes(), hipMemcpyHostToDevice),
                      "cannot copy the projection to the HIP device");
        } else {
            // rows == 1 with an empty sample: nothing to project; keep a valid pointer.
            ix->proj = DeviceBuffer<float>(device, 4);
        }
        if (metric == UHANN_METRIC_L2) {
            // bf16 copy of the binary32 projection (round to nearest even),
            // zero padded to the MFMA tile.
            ix->proj16_ld = kern::project_bf16_ld(ix->k);
            const uint64_t prow = kern::project_bf16_rows(projection_dim);
            std::vector<uint16_t> p16(checked_mul(prow, ix->proj16_ld, "bf16 projection"), 0);
            for (uint64_t o = 0; o < projection_dim; ++o) {
                for (uint64_t j = 0; j < ix->k; ++j) p16[o * ix->proj16_ld + j] = bf16_bits(proj32[o * ix->k + j]);
            }
            ix->proj16 = DeviceBuffer<uint16_t>(device, p16.size());
            check_hip(hipMemcpy(ix->proj16.data(), p16.data(), ix->proj16.bytes(), hipMemcpyHostToDevice),
                      "cannot copy the bf16 projection to the HIP device");
            ix->scan = std::make_unique<Scan>(metric, projected.data(), rows, projection_dim, nullptr, device);
            // Reference accounting: scan payload + SampledColumn sample + FP32 matrix.
            ix->payload_bytes = ix->scan->payload_bytes + ix->k * sizeof(uhann_sampled_column) +
                                projection_dim * ix->k * sizeof(float);
        } else {
            std::vector<float> t(checked_mul(projection_dim, rows, "projected representatives"));
            for (uint64_t r = 0; r < rows; ++r) {
                for (uint64_t o = 0; o < projection_dim; ++o) t[o * rows + r] = projected[r * projection_dim + o];
            }
            ix->refs_t = DeviceBuffer<float>(device, t.size());
            check_hip(hipMemcpy(ix->refs_t.data(), t.data(), ix->refs_t.bytes(), hipMemcpyHostToDevice),
                      "cannot copy projected representatives to the HIP device");
            ix->payload_bytes = ix->k * sizeof(uhann_sampled_column) + projection_dim * ix->k * sizeof(float) +
                                rows * projection_dim * sizeof(float);
        }
        *out = ix.release();
    });
}

void uhann_index_destroy(uhann_index* index) { delete index; }

int uhann_index_info(const uhann_index* ix, int* kind, int* metric, uint64_t* rows, uint64_t* dim,
                     uint64_t* unique_coordinates, uint64_t* multiplicity, uint64_t* payload_bytes)
{
    return guarded([&] {
        if (ix == nullptr) throw std::invalid_argument("index handle must not be null");
        if (kind) *kind = ix->kind;
        if (metric) *metric = ix->metric;
        if (rows) *rows = ix->rows;
        if (dim) *dim = ix->dim;
        if (unique_coordinates) *unique_coordinates = ix->k;
        if (multiplicity) *multiplicity = ix->multiplicity;
        if (payload_bytes) {
            *payload_bytes = ix->payload_bytes + (ix->scan && ix->scan->gram
                                                      ? ix->scan->gram->refs16.bytes() + ix->scan->gram->stats.bytes()
                                                      : 0);
        }
    });
}

int uhann_index_enable_mixed(uhann_index* ix, int strategy)
{
    return guarded([&] {
        if (ix == nullptr) throw std::invalid_argument("index handle must not be null");
        if (strategy != UHANN_STRATEGY_GRAM_BF16 && strategy != UHANN_STRATEGY_GRAM_FP8 &&
            strategy != UHANN_STRATEGY_GRAM_INT8) {
            throw std::invalid_argument("unknown mixed strategy");
        }
        if (ix->kind == UHANN_KIND_HIERARCHICAL) {
            throw std::invalid_argument("the mixed strategy applies to exact and flat indexes");
        }
        ix->scan->enable_mixed(strategy);
    });
}

int uhann_index_workspace_create(const uhann_index* ix, uint64_t max_batch, uhann_workspace** out)
{
    return guarded([&] {
        if (ix == nullptr || out == nullptr) throw std::invalid_argument("null handle");
        *out = nullptr;
        const uint64_t query_ld = padded_ld(ix->dim);
        uint64_t extra = 0;
        if (ix->kind == UHANN_KIND_FLAT) extra = ix->k_ld;
        if (ix->kind == UHANN_KIND_HIERARCHICAL) extra = ix->k_ld + ix->p_ld;
        *out = new uhann_workspace(ix, ix->device, max_batch, query_ld, extra);
    });
}

int uhann_index_query_batch(const uhann_index* ix, uhann_workspace* ws, const float* host_queries,
                            uint64_t count, uint64_t* out_idx, float* out_dist, int strategy)
{
    return guarded([&] {
        if (ix == nullptr) throw std::invalid_argument("index handle must not be null");
        checked_ws(ix, ws, count);
        const std::size_t n = checked_mul(count, ix->dim, "query batch");
        if (n != 0 && host_queries == nullptr) throw std::invalid_argument("queries must not be null");
        // The exact path validates every query value (dense_l2_scan_index.cpp:137);
        // sampled paths validate only the sampled coordinates, on the device.
        if (ix->kind == UHANN_KIND_EXACT) validate_finite_host(host_queries, n, "queries");
        if (count == 0) return;
        select_device(ix->device);
        const uint64_t ld = padded_ld(ix->dim);
        if (n != 0) {
            check_hip(hipMemcpy2DAsync(ws->queries.data(), ld * sizeof(float), host_queries, ix->dim * sizeof(float),
                                       ix->dim * sizeof(float), count, hipMemcpyHostToDevice, ws->stream),
                      "cannot copy queries to the HIP device");
        }
        sync_query(ix, ws, ws->queries.data(), ld, true, count, out_idx, out_dist, strategy);
    });
}

int uhann_index_query_device_batch(const uhann_index* ix, uhann_workspace* ws, const float* device_queries,
                                   uint64_t count, uint64_t* out_idx, float* out_dist, int strategy)
{
    return guarded([&] {
        if (ix == nullptr) throw std::invalid_argument("index handle must not be null");
        checked_ws(ix, ws, count);
        if (count == 0) return;
        if (ix->dim != 0 && device_queries == nullptr) throw std::invalid_argument("device query pointer must not be null");
        select_device(ix->device);
        sync_query(ix, ws, device_queries, ix->dim, false, count, out_idx, out_dist, strategy);
    });
}

int uhann_index_query_device_async(const uhann_index* ix, uhann_workspace* ws, const float* device_queries,
                                   uint64_t ld_queries, uint64_t count, uint64_t* device_out_idx,
                                   float* device_out_dist, int strategy, void* stream)
{
    return guarded([&] {
        if (ix == nullptr) throw std::invalid_argument("index handle must not be null");
        checked_ws(ix, ws, count);
        if (count == 0) return;
        if (ix->dim != 0 && device_queries == nullptr) throw std::invalid_argument("device query pointer must not be null");
        if (ld_queries < ix->dim) throw std::invalid_argument("query row stride is below dim");
        select_device(ix->device);
        index_launch(*ix, *ws, device_queries, ld_queries, false, count, device_out_idx, device_out_dist, strategy,
                     static_cast<hipStream_t>(stream));
    });
}

int uhann_index_autotune(uhann_index* ix, uhann_workspace* ws, const float* device_queries, uint64_t ld_queries,
                         uint64_t max_batch, int* table, int capacity, int* count)
{
    return guarded([&] {
        if (ix == nullptr) throw std::invalid_argument("index handle must not be null");
        checked_ws(ix, ws, max_batch);
        if (max_batch == 0) throw std::invalid_argument("autotune batch must be positive");
        if (ix->dim != 0 && device_queries == nullptr) throw std::invalid_argument("device query pointer must not be null");
        if (ld_queries < ix->dim) throw std::invalid_argument("query row stride is below dim");
        if (!ix->scan) throw std::invalid_argument("this index has no direct scan to tune");
        select_device(ix->device);
        hipStream_t s = ws->stream;
        const uint32_t n = static_cast<uint32_t>(max_batch);
        const float* q = device_queries;
        uint64_t ld = ld_queries;
        bool padded = false;
        if (ix->kind == UHANN_KIND_FLAT) {
            kern::gather_columns(device_queries, ld_queries, n, ix->cols.data(), ix->scales.data(),
                                 static_cast<uint32_t>(ix->k), ws->extra.data(), ix->k_ld, ws->invalid.data(), s);
            q = ws->extra.data();
            ld = ix->k_ld;
            padded = true;
        } else if (ix->kind == UHANN_KIND_HIERARCHICAL) {
            float* projected = ws->extra.data() + max_batch * ix->k_ld;
            kern::gather_columns(device_queries, ld_queries, n, ix->cols.data(), nullptr, static_cast<uint32_t>(ix->k),
                                 ws->extra.data(), ix->k_ld, ws->invalid.data(), s);
            kern::project_fp32(ws->extra.data(), ix->k_ld, n, static_cast<uint32_t>(ix->k), ix->proj.data(), ix->k_ld,
                               static_cast<uint32_t>(ix->p), projected, ix->p_ld, ws->invalid.data() + 1, s);
            q = projected;
            ld = ix->p_ld;
            padded = true;
        }
        autotune_scan(*ix->scan, *ws, q, ld, padded, max_batch, s);
        int rows = 0;
        for (int b = 0; b < 24; ++b) {
            const Scan::Tuned& t = ix->scan->tuned[b];
            if (!t.set) continue;
            if (table != nullptr && rows < capacity) {
                int* row = table + 6 * rows;
                row[0] = b;
                row[1] = t.tile.rpt;
                row[2] = t.tile.tq;
                row[3] = t.tile.tr;
                row[4] = t.tile.group_q;
                row[5] = static_cast<int>(t.ms * 1000.0F);
            }
            ++rows;
        }
        if (count != nullptr) *count = rows;
    });
}

int uhann_workspace_status(uhann_workspace* ws)
{
    return guarded([&] {
        if (ws == nullptr) throw std::invalid_argument("workspace handle must not be null");
        select_device(ws->device);
        check_hip(hipDeviceSynchronize(), "HIP device synchronisation failed");
        try {
            raise_flags(*ws);
        } catch (...) {
            check_hip(hipMemset(ws->invalid.data(), 0, ws->invalid.bytes()), "cannot clear validation flags");
            throw;
        }
    });
}

}  // extern "C"
