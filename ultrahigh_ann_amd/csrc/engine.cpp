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
#include <cstdlib>
#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "common.hpp"
#include "engine.hpp"
#include "kernels.hpp"
#include "probabilities.hpp"

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
    device_name = device_label(prop);
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
        if (ld == dim) {
            check_hip(hipMemcpy(refs.data(), host_refs, refs.bytes(), hipMemcpyHostToDevice),
                      "cannot copy representatives to the HIP device");
        } else {
            std::vector<float> staged(static_cast<std::size_t>(rows * ld), 0.0F);
            for (uint64_t r = 0; r < rows; ++r) {
                std::memcpy(staged.data() + r * ld, host_refs + r * dim, dim * sizeof(float));
            }
            check_hip(hipMemcpy(refs.data(), staged.data(), refs.bytes(), hipMemcpyHostToDevice),
                      "cannot copy representatives to the HIP device");
        }
    }
    if (dim != 0 && dim <= kern::scan_pairs_max_dim() && rows < kern::kPairsMaxRows) {
        // Short rows also get a transposed copy (dim x rows, then
        // kPairsPadCols rows of zeros) for the pair-per-thread scan, whose
        // loads are then contiguous across a wave (that scan takes fewer
        // than 2^25 rows, so larger sets skip it).
        std::vector<float> t(static_cast<std::size_t>(rows * (dim + kern::kPairsPadCols)), 0.0F);
        for (uint64_t r = 0; r < rows; ++r) {
            for (uint64_t c = 0; c < dim; ++c) t[c * rows + r] = host_refs[r * dim + c];
        }
        refs_t = DeviceBuffer<float>(device, t.size());
        check_hip(hipMemcpy(refs_t.data(), t.data(), refs_t.bytes(), hipMemcpyHostToDevice),
                  "cannot copy transposed representatives to the HIP device");
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
    invalid = DeviceBuffer<int>(device, 4);
    check_hip(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking),
              "cannot create a HIP stream");
    payload_bytes = queries.bytes() + extra.bytes() + keys.bytes() + out_idx.bytes() +
                    out_dist.bytes() + invalid.bytes();
}

void Workspace::mark(int i, hipStream_t s)
{
    if (!timing) return;
    if (events[i] == nullptr) {
        check_hip(hipEventCreate(&events[i]), "cannot create a HIP event");
    }
    check_hip(hipEventRecord(events[i], s), "cannot record a HIP event");
}

Workspace::~Workspace()
{
    for (hipEvent_t& e : events) {
        if (e != nullptr) static_cast<void>(hipEventDestroy(e));
    }
    if (stream != nullptr) {
        static_cast<void>(hipSetDevice(device));
        static_cast<void>(hipStreamSynchronize(stream));
        static_cast<void>(hipStreamDestroy(stream));
    }
}

// Rows up to this length use the pair-per-thread scan by default
// (UHANN_PAIRS_MAX_DIM overrides it for tuning).
uint32_t pairs_max_dim()
{
    const char* env = std::getenv("UHANN_PAIRS_MAX_DIM");
    if (env != nullptr) {
        const long v = std::strtol(env, nullptr, 10);
        if (v >= 0) return static_cast<uint32_t>(std::min<long>(v, kern::scan_pairs_max_dim()));
    }
    return 4096;
}

int batch_bucket(uint64_t nq)
{
    int b = 0;
    while (b < 23 && (uint64_t{1} << b) < nq) ++b;
    return b;
}

void autotune_scan(Scan& scan, Workspace& ws, const float* queries, uint64_t ld_q, bool padded,
                   uint64_t max_batch, hipStream_t stream)
{
    hipEvent_t e0 = nullptr, e1 = nullptr;
    check_hip(hipEventCreate(&e0), "cannot create a HIP event");
    check_hip(hipEventCreate(&e1), "cannot create a HIP event");
    try {
        const bool aligned4 = (ld_q % 4U == 0) && (reinterpret_cast<uintptr_t>(queries) % 16U == 0) &&
                              (scan.dim % 4U == 0 || padded);
        std::vector<kern::ScanTile> cands;
        for (const kern::ScanTile& t : kern::scan_tiles()) {
            if (t.rpt != 4 || aligned4) cands.push_back(t);
        }
        if (scan.refs_t.bytes() != 0 && scan.dim != 0 && scan.rows < kern::kPairsMaxRows) {
            for (int tq : {1, 2, 4}) cands.push_back(kern::ScanTile{0, tq, 0, 0});
            for (int tq : {4, 8}) cands.push_back(kern::ScanTile{0, tq, 8, 0});
        }
        if (scan.dim != 0 && scan.dim <= kern::scan_short_max_dim()) cands.push_back(kern::ScanTile{0, 0, 0, 0});
        for (uint64_t nq = 1;; nq *= 2) {
            const uint64_t n = std::min(nq, max_batch);
            float best = -1.0F;
            kern::ScanTile best_tile{};
            for (const kern::ScanTile& t : cands) {
                // Tiles much taller than the batch only repeat clamped rows.
                if (t.rpt != 0 && static_cast<uint64_t>(t.tq) > 2 * n + 1) continue;
                float ms_min = -1.0F;
                for (int rep = 0; rep < 3; ++rep) {
                    check_hip(hipEventRecord(e0, stream), "cannot record a HIP event");
                    scan_launch(scan, queries, ld_q, n, ws.keys.data(), ws.out_idx.data(), ws.out_dist.data(),
                                stream, padded, &t);
                    check_hip(hipEventRecord(e1, stream), "cannot record a HIP event");
                    check_hip(hipEventSynchronize(e1), "autotune run failed");
                    float ms = 0.0F;
                    check_hip(hipEventElapsedTime(&ms, e0, e1), "cannot read a HIP event");
                    if (rep > 0 && (ms_min < 0.0F || ms < ms_min)) ms_min = ms;
                    // Clearly slower than the best so far: stop timing it.
                    if (rep == 0 && best > 0.0F && ms > 2.0F * best) break;
                }
                if (ms_min > 0.0F && (best < 0.0F || ms_min < best)) {
                    best = ms_min;
                    best_tile = t;
                }
            }
            if (best > 0.0F) {
                Scan::Tuned& slot = scan.tuned[batch_bucket(n)];
                slot.set = true;
                slot.tile = best_tile;
                slot.ms = best;
            }
            if (n >= max_batch) break;
        }
    } catch (...) {
        static_cast<void>(hipEventDestroy(e0));
        static_cast<void>(hipEventDestroy(e1));
        throw;
    }
    static_cast<void>(hipEventDestroy(e0));
    static_cast<void>(hipEventDestroy(e1));
}

namespace {

uint32_t device_cus(int device)
{
    int cus = 0;
    check_hip(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device),
              "cannot read the compute unit count");
    return cus > 0 ? static_cast<uint32_t>(cus) : 0U;
}

}  // namespace

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
        t = kern::ScanTile{1, 16, 8, 4};
    }
    return t;
}

bool stream_refs_enabled();
bool scan_buffers_enabled();

void scan_launch(const Scan& scan, const float* queries, uint64_t ld_q, uint64_t count,
                 unsigned long long* keys, uint64_t* out_idx, float* out_dist,
                 hipStream_t stream, bool queries_padded, const kern::ScanTile* forced_tile, Workspace* timed)
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
    // 16-byte vector loads need aligned rows, and a row whose length is not a
    // multiple of 4 must be padded (the last vector reads up to 3 floats past
    // dim): true for workspace-staged queries, not assumed for caller memory.
    const bool aligned4 = (ld_q % 4U == 0) && (scan.ld % 4U == 0) &&
                          (reinterpret_cast<uintptr_t>(queries) % 16U == 0) &&
                          (scan.dim % 4U == 0 || queries_padded);
    kern::ScanTile tile = forced_tile != nullptr ? *forced_tile : choose_tile(count, aligned4);
    // Short rows: the pair-per-thread kernel (tile.rpt == 0 selects it, tq
    // queries per thread) unless a tiled shape is forced or tuned.
    const bool pairs_ok = scan.refs_t.bytes() != 0 && a.dim != 0 && scan.rows < kern::kPairsMaxRows;
    // Rows of at most 128 coordinates: one thread per query (tile.rpt == 0,
    // tq == 0) when the launch has at least one 256-query x 128-reference
    // block per compute unit; fewer pairs go to the pair-per-thread form.
    const bool short_ok = a.dim != 0 && a.dim <= kern::scan_short_max_dim();
    const uint64_t short_blocks = ((count + 255) / 256) * ((scan.rows + 127) / 128);
    const Scan::Tuned& tuned = scan.tuned[batch_bucket(count)];
    if (forced_tile == nullptr && tuned.set && (tuned.tile.rpt != 4 || aligned4)) {
        tile = tuned.tile;
    } else if (forced_tile == nullptr && short_ok && short_blocks >= device_cus(scan.device)) {
        tile = kern::ScanTile{0, 0, 0, 0};
    } else if (forced_tile == nullptr && pairs_ok && a.dim <= pairs_max_dim() && count >= 4) {
        // chain classes, 4 queries per thread (at r = 7028, batch 4096 the
        // fastest form from 64 to 3527 coordinates, the tiled kernel included)
        tile = kern::ScanTile{0, 4, 8, 0};
    }
    if (tile.rpt == 0) {
        const kern::Metric metric = scan.metric == UHANN_METRIC_L2 ? kern::Metric::l2 : kern::Metric::l1;
        if (tile.tq == 0 && !short_ok) throw std::invalid_argument("short-row scan needs 1 to 128 coordinates");
        if (tile.tq != 0 && !pairs_ok) throw std::invalid_argument("pair scan needs rows of at most 4096 coordinates and fewer than 2^25 references");
        kern::fill_keys(keys, a.nq, stream);
        if (timed != nullptr) timed->mark(0, stream);
        if (tile.tq == 0) {
            kern::scan_short(metric, /*weight_fma=*/true, scan.refs.data(), scan.ld, a.rows, queries, ld_q, a.nq,
                             a.weights, a.dim, keys, stream);
        } else {
            kern::scan_pairs(metric, /*weight_fma=*/true, scan.refs_t.data(), scan.rows, a.rows, queries, ld_q, a.nq,
                             a.weights, a.dim, keys, tile.tq, tile.tr == 8, stream);
        }
        if (timed != nullptr) timed->mark(1, stream);
        kern::finalize_keys(keys, a.nq, out_idx, out_dist, stream);
        if (timed != nullptr && timed->timing) {
            check_hip(hipEventSynchronize(timed->events[1]), "HIP event wait failed");
            check_hip(hipEventElapsedTime(&timed->stage_ms[0], timed->events[0], timed->events[1]),
                      "cannot read HIP event time");
            for (int i = 1; i < Workspace::kStages; ++i) timed->stage_ms[i] = 0.0F;
        }
        return;
    }
    if (tile.rpt == 4 && !aligned4) {
        throw std::invalid_argument(
            "vector scan tile needs 16-byte aligned, padded query rows");
    }
    // One query tile reads every reference value exactly once: stream them.
    a.stream_refs = a.nq <= static_cast<uint32_t>(tile.tq) && stream_refs_enabled();
    // Buffer-resource loads when a tile's rows lie within 2^32 bytes of its
    // first row (references and queries).
    a.buffers = scan_buffers_enabled() && static_cast<uint64_t>(tile.tr) * scan.ld * 4U < (1ULL << 32) &&
                static_cast<uint64_t>(tile.tq) * ld_q * 4U < (1ULL << 32);
    if (a.dim == 0) {
        // No coordinates: every distance is zero and the reference returns row 0.
        a.queries = scan.refs.data();
        a.ld_q = 0;
    }
    kern::fill_keys(keys, a.nq, stream);
    if (timed != nullptr) timed->mark(0, stream);
    kern::scan_direct(scan.metric == UHANN_METRIC_L2 ? kern::Metric::l2 : kern::Metric::l1,
                      /*weight_fma=*/true, a, tile, stream);
    if (timed != nullptr) timed->mark(1, stream);
    kern::finalize_keys(keys, a.nq, out_idx, out_dist, stream);
    if (timed != nullptr && timed->timing) {
        // Stage 0 = the scan kernel alone (without the key fill / finalise launches).
        check_hip(hipEventSynchronize(timed->events[1]), "HIP event wait failed");
        check_hip(hipEventElapsedTime(&timed->stage_ms[0], timed->events[0], timed->events[1]),
                  "cannot read HIP event time");
        for (int i = 1; i < Workspace::kStages; ++i) timed->stage_ms[i] = 0.0F;
    }
}


// Gram block shape; UHANN_GRAM_TILE (0..7) overrides the default for tuning.
// Non-temporal reference loads for single-pass scans; UHANN_STREAM_REFS=0
// turns them off (for measurement).
bool stream_refs_enabled()
{
    const char* env = std::getenv("UHANN_STREAM_REFS");
    return env == nullptr || env[0] != '0';
}

// Buffer-resource loads in the tiled direct scan; UHANN_SCAN_BUFFERS=0 uses
// 64-bit global addresses instead (same arithmetic, for measurement).
bool scan_buffers_enabled()
{
    const char* env = std::getenv("UHANN_SCAN_BUFFERS");
    return env == nullptr || env[0] != '0';
}

int gram_tile()
{
    const char* env = std::getenv("UHANN_GRAM_TILE");
    if (env != nullptr) {
        char* end = nullptr;
        const long t = std::strtol(env, &end, 10);
        if (end != env && *end == '\0' && t >= 0 && t < kern::gram_tile_count()) return static_cast<int>(t);
    }
    return 3;
}

void Scan::enable_mixed(int strategy)
{
    if (strategy != UHANN_STRATEGY_GRAM_BF16 && strategy != UHANN_STRATEGY_GRAM_FP8 &&
        strategy != UHANN_STRATEGY_GRAM_INT8) {
        throw std::invalid_argument("unknown mixed strategy");
    }
    if (metric != UHANN_METRIC_L2) {
        throw std::invalid_argument("the mixed-precision strategy applies to the L2 metric only");
    }
    select_device(device);
    // Rows padded with zeros to whole Gram tiles: the Gram kernels read
    // complete tiles without clamping row indices.
    const uint64_t mt = kern::gram_m_tile();
    const uint64_t padded_rows = (rows + mt - 1) / mt * mt;
    if (strategy == UHANN_STRATEGY_GRAM_INT8) {
        if (i8) return;
        if (dim > kern::gram_i8_max_dim()) {
            throw std::length_error("the int8 strategy keeps its sums exact for at most 65536 coordinates");
        }
        auto f = std::make_unique<I8Refs>();
        const uint64_t step = kern::gram_i8_k_padding();
        f->ldk = std::max<uint64_t>(step, (dim + 63) / 64 * step);
        f->refs = DeviceBuffer<uint8_t>(device, checked_mul(padded_rows, f->ldk, "int8 references"));
        check_hip(hipMemset(f->refs.data(), 0, f->refs.bytes()), "cannot clear the int8 references");
        f->stats = DeviceBuffer<kern::I8Stats>(device, rows);
        kern::to_i8x2_rows(refs.data(), ld, static_cast<uint32_t>(rows), static_cast<uint32_t>(dim), f->refs.data(),
                           f->ldk, f->stats.data(), nullptr);
        check_hip(hipDeviceSynchronize(), "int8 reference conversion failed");
        payload_bytes += f->refs.bytes() + f->stats.bytes();
        i8 = std::move(f);
        return;
    }
    if (strategy == UHANN_STRATEGY_GRAM_FP8) {
        if (fp8) return;
        auto f = std::make_unique<Fp8Refs>();
        const uint64_t step = kern::gram_fp8_k_padding();
        f->ldk8 = std::max<uint64_t>(step, (dim + step - 1) / step * step);
        f->refs8 = DeviceBuffer<uint8_t>(device, checked_mul(padded_rows, f->ldk8, "fp8 references"));
        check_hip(hipMemset(f->refs8.data(), 0, f->refs8.bytes()), "cannot clear the fp8 references");
        f->stats = DeviceBuffer<kern::RowStats>(device, rows);
        kern::to_fp8_rows(refs.data(), ld, static_cast<uint32_t>(rows), static_cast<uint32_t>(dim), f->refs8.data(),
                          f->ldk8, f->stats.data(), nullptr);
        check_hip(hipDeviceSynchronize(), "fp8 reference conversion failed");
        payload_bytes += f->refs8.bytes() + f->stats.bytes();
        fp8 = std::move(f);
        return;
    }
    if (gram) return;
    auto g = std::make_unique<GramRefs>();
    const uint64_t step = kern::gram_k_padding();
    g->ldk = std::max<uint64_t>(step, (dim + step - 1) / step * step);
    g->refs16 = DeviceBuffer<__bf16>(device, checked_mul(padded_rows, g->ldk, "bf16 references"));
    check_hip(hipMemset(g->refs16.data(), 0, g->refs16.bytes()), "cannot clear the bf16 references");
    g->stats = DeviceBuffer<kern::RowStats>(device, rows);
    kern::to_bf16_rows(refs.data(), ld, static_cast<uint32_t>(rows), static_cast<uint32_t>(dim),
                       g->refs16.data(), g->ldk, g->stats.data(), nullptr);
    check_hip(hipDeviceSynchronize(), "bf16 reference conversion failed");
    payload_bytes += g->refs16.bytes() + g->stats.bytes();
    gram = std::move(g);
}

void Workspace::reserve_partials(uint64_t floats)
{
    const uint64_t have = gram_partials.bytes() / sizeof(float);
    if (floats <= have) return;
    gram_partials = DeviceBuffer<float>(device, floats);
    payload_bytes += (floats - have) * sizeof(float);
}

void Workspace::reserve_mixed(const Scan& scan)
{
    if (scan.gram == nullptr || q16.bytes() != 0) return;
    const uint64_t mt = kern::gram_m_tile();
    const uint64_t mrows = (max_batch + mt - 1) / mt * mt;
    q16 = DeviceBuffer<__bf16>(device, checked_mul(mrows, scan.gram->ldk, "bf16 query workspace"));
    qstats = DeviceBuffer<kern::RowStats>(device, mrows);
    payload_bytes += q16.bytes() + qstats.bytes();
    reserve_partials(kern::gram_split_floats(static_cast<uint32_t>(max_batch), static_cast<uint32_t>(scan.rows),
                                             static_cast<uint32_t>(scan.gram->ldk / kern::gram_k_padding()),
                                             device_cus(device)));
    reserve_common(scan);
}

void Workspace::reserve_fp8(const Scan& scan)
{
    if (scan.fp8 == nullptr || q8.bytes() != 0) return;
    const uint64_t mt = kern::gram_m_tile();
    const uint64_t mrows = (max_batch + mt - 1) / mt * mt;
    q8 = DeviceBuffer<uint8_t>(device, checked_mul(mrows, scan.fp8->ldk8, "fp8 query workspace"));
    q8stats = DeviceBuffer<kern::RowStats>(device, mrows);
    payload_bytes += q8.bytes() + q8stats.bytes();
    reserve_partials(kern::gram_split_floats(static_cast<uint32_t>(max_batch), static_cast<uint32_t>(scan.rows),
                                             static_cast<uint32_t>(scan.fp8->ldk8 / kern::gram_fp8_k_padding()),
                                             device_cus(device)));
    reserve_common(scan);
}

void Workspace::reserve_i8(const Scan& scan)
{
    if (scan.i8 == nullptr || qi8.bytes() != 0) return;
    const uint64_t mt = kern::gram_m_tile();
    const uint64_t mrows = (max_batch + mt - 1) / mt * mt;
    qi8 = DeviceBuffer<uint8_t>(device, checked_mul(mrows, scan.i8->ldk, "int8 query workspace"));
    qi8stats = DeviceBuffer<kern::I8Stats>(device, mrows);
    payload_bytes += qi8.bytes() + qi8stats.bytes();
    reserve_partials(kern::gram_i8_split_floats(static_cast<uint32_t>(max_batch), static_cast<uint32_t>(scan.rows),
                                                static_cast<uint32_t>(scan.i8->ldk / kern::gram_i8_k_padding()),
                                                device_cus(device)));
    reserve_common(scan);
}

// Estimate matrix, counters and escalation buffers shared by the strategies.
void Workspace::reserve_common(const Scan& scan)
{
    if (approx.bytes() != 0) return;
    approx = DeviceBuffer<float>(device, checked_mul(max_batch, scan.rows, "Gram workspace"));
    counters = DeviceBuffer<unsigned int>(device, 8);
    escalate_list = DeviceBuffer<unsigned int>(device, max_batch);
    esc_queries = DeviceBuffer<float>(device, checked_mul(max_batch, scan.ld, "escalation workspace"));
    payload_bytes += approx.bytes() + counters.bytes() + escalate_list.bytes() + esc_queries.bytes();
}

void mixed_launch(const Scan& scan, Workspace& ws, const float* queries, uint64_t ld_q,
                  uint64_t count, uint64_t* out_idx, float* out_dist, hipStream_t stream)
{
    if (scan.gram == nullptr) {
        throw std::invalid_argument("mixed-precision strategy not enabled for this index");
    }
    if (out_idx == nullptr || out_dist == nullptr) {
        throw std::invalid_argument("the mixed-precision strategy needs index and distance outputs");
    }
    ws.reserve_mixed(scan);
    const GramRefs& g = *scan.gram;
    const uint32_t n = static_cast<uint32_t>(count);
    ws.mark(0, stream);
    kern::to_bf16_rows(queries, ld_q, n, static_cast<uint32_t>(scan.dim), ws.q16.data(), g.ldk,
                       ws.qstats.data(), stream);
    ws.mark(1, stream);
    kern::gram_bf16(ws.q16.data(), g.ldk, n, g.refs16.data(), g.ldk, static_cast<uint32_t>(scan.rows),
                    static_cast<uint32_t>(g.ldk), ws.qstats.data(), g.stats.data(), ws.approx.data(),
                    scan.rows, gram_tile(), ws.gram_partials.data(), ws.gram_partials.bytes() / sizeof(float),
                    stream);
    ws.mark(2, stream);
    kern::zero_u32(ws.counters.data(), 8, stream);
    kern::SelectArgs a;
    a.nq = n;
    a.rows = static_cast<uint32_t>(scan.rows);
    a.dim = static_cast<uint32_t>(scan.dim);
    a.approx = ws.approx.data();
    a.ld_approx = scan.rows;
    a.qstats = ws.qstats.data();
    a.vstats = g.stats.data();
    // BF16 MFMA: binary32 rounding once per 16-product step, modelled as
    // random (Hoeffding, K = 7) over ceil(K/16) steps.
    a.acc_coef = 7.0F * static_cast<float>(std::sqrt(static_cast<double>((g.ldk + 15) / 16)));
    a.queries = queries;
    a.ld_q = ld_q;
    a.refs = scan.refs.data();
    a.ld_ref = scan.ld;
    const double k = std::ceil(static_cast<double>(scan.dim) / 256.0) + 10.0;
    const double u = std::ldexp(1.0, -24);
    const double gamma = k * u / (1.0 - k * u);
    a.fp32_band = (1.0 + gamma) / (1.0 - gamma);
    a.out_idx = out_idx;
    a.out_dist = out_dist;
    a.escalate_count = ws.counters.data();
    a.escalate_list = ws.escalate_list.data();
    a.stats = reinterpret_cast<kern::MixedStats*>(ws.counters.data() + 4);
    kern::select_rerank(a, stream);
    ws.mark(3, stream);
    // Escalated queries get the exact direct scan on the device (the count
    // stays in device memory; nothing waits for the host).
    kern::ScanArgs esc;
    esc.refs = scan.refs.data();
    esc.ld_ref = scan.ld;
    esc.rows = static_cast<uint32_t>(scan.rows);
    esc.queries = queries;
    esc.ld_q = ld_q;
    esc.nq = n;
    esc.dim = static_cast<uint32_t>(scan.dim);
    esc.keys = ws.keys.data();
    kern::escalate_scan(esc, ws.counters.data(), ws.escalate_list.data(), n, ws.esc_queries.data(), scan.ld,
                        out_idx, out_dist, stream);
    ws.mark(4, stream);
    if (ws.timing) {
        check_hip(hipEventSynchronize(ws.events[4]), "HIP event wait failed");
        for (int i = 0; i < Workspace::kStages; ++i) {
            check_hip(hipEventElapsedTime(&ws.stage_ms[i], ws.events[i], ws.events[i + 1]),
                      "cannot read HIP event time");
        }
    }
}

void fp8_launch(const Scan& scan, Workspace& ws, const float* queries, uint64_t ld_q, uint64_t count,
                uint64_t* out_idx, float* out_dist, hipStream_t stream)
{
    if (scan.fp8 == nullptr) throw std::invalid_argument("fp8 strategy not enabled for this index");
    if (out_idx == nullptr || out_dist == nullptr) {
        throw std::invalid_argument("the fp8 strategy needs index and distance outputs");
    }
    ws.reserve_fp8(scan);
    const Fp8Refs& f = *scan.fp8;
    const uint32_t n = static_cast<uint32_t>(count);
    ws.mark(0, stream);
    kern::to_fp8_rows(queries, ld_q, n, static_cast<uint32_t>(scan.dim), ws.q8.data(), f.ldk8, ws.q8stats.data(),
                      stream);
    ws.mark(1, stream);
    kern::gram_fp8(ws.q8.data(), f.ldk8, n, f.refs8.data(), f.ldk8, static_cast<uint32_t>(scan.rows),
                   static_cast<uint32_t>(f.ldk8), ws.q8stats.data(), f.stats.data(), ws.approx.data(), scan.rows,
                   ws.gram_partials.data(), ws.gram_partials.bytes() / sizeof(float), stream);
    ws.mark(2, stream);
    kern::zero_u32(ws.counters.data(), 8, stream);
    kern::SelectArgs a;
    a.nq = n;
    a.rows = static_cast<uint32_t>(scan.rows);
    a.dim = static_cast<uint32_t>(scan.dim);
    a.approx = ws.approx.data();
    a.ld_approx = scan.rows;
    a.qstats = ws.q8stats.data();
    a.vstats = f.stats.data();
    // FP8 MFMA: its sum of 64 products per instruction is coarser than
    // binary32 (errors up to 2.4 u_32 per product relative to sum |q8 v8|
    // were measured), so the bound allows 8 u_32 per product.
    a.acc_coef = 8.0F * static_cast<float>(f.ldk8);
    a.queries = queries;
    a.ld_q = ld_q;
    a.refs = scan.refs.data();
    a.ld_ref = scan.ld;
    const double k = std::ceil(static_cast<double>(scan.dim) / 256.0) + 10.0;
    const double u = std::ldexp(1.0, -24);
    const double gamma = k * u / (1.0 - k * u);
    a.fp32_band = (1.0 + gamma) / (1.0 - gamma);
    a.out_idx = out_idx;
    a.out_dist = out_dist;
    a.stats = reinterpret_cast<kern::MixedStats*>(ws.counters.data() + 4);
    a.shortlist = scan.fp8_shortlist;
    kern::select_topk_rerank(a, stream);
    ws.mark(3, stream);
    ws.mark(4, stream);
    if (ws.timing) {
        check_hip(hipEventSynchronize(ws.events[4]), "HIP event wait failed");
        for (int i = 0; i < Workspace::kStages; ++i) {
            check_hip(hipEventElapsedTime(&ws.stage_ms[i], ws.events[i], ws.events[i + 1]),
                      "cannot read HIP event time");
        }
    }
}

// The binary32 scan's relative error bound gamma_k (k = ceil(dim / 256) + 10:
// the longest chain, the tree depth and the subtract and square of a term)
// as the band factor (1 + gamma) / (1 - gamma).
double fp32_band(uint64_t dim)
{
    const double k = std::ceil(static_cast<double>(dim) / 256.0) + 10.0;
    const double u = std::ldexp(1.0, -24);
    const double gamma = k * u / (1.0 - k * u);
    return (1.0 + gamma) / (1.0 - gamma);
}

void i8_launch(const Scan& scan, Workspace& ws, const float* queries, uint64_t ld_q, uint64_t count,
               uint64_t* out_idx, float* out_dist, hipStream_t stream)
{
    if (scan.i8 == nullptr) throw std::invalid_argument("int8 strategy not enabled for this index");
    if (out_idx == nullptr || out_dist == nullptr) {
        throw std::invalid_argument("the int8 strategy needs index and distance outputs");
    }
    ws.reserve_i8(scan);
    const I8Refs& f = *scan.i8;
    const uint32_t n = static_cast<uint32_t>(count);
    ws.mark(0, stream);
    kern::to_i8x2_rows(queries, ld_q, n, static_cast<uint32_t>(scan.dim), ws.qi8.data(), f.ldk, ws.qi8stats.data(),
                       stream);
    ws.mark(1, stream);
    kern::gram_i8x2(ws.qi8.data(), f.ldk, n, f.refs.data(), f.ldk, static_cast<uint32_t>(scan.rows),
                    static_cast<uint32_t>(f.ldk), ws.qi8stats.data(), f.stats.data(), ws.approx.data(), scan.rows,
                    -1, ws.gram_partials.data(), ws.gram_partials.bytes() / sizeof(float), stream);
    ws.mark(2, stream);
    kern::zero_u32(ws.counters.data(), 8, stream);
    kern::SelectArgs a;
    a.nq = n;
    a.rows = static_cast<uint32_t>(scan.rows);
    a.dim = static_cast<uint32_t>(scan.dim);
    a.approx = ws.approx.data();
    a.ld_approx = scan.rows;
    a.qstats8 = ws.qi8stats.data();
    a.vstats8 = f.stats.data();
    a.queries = queries;
    a.ld_q = ld_q;
    a.refs = scan.refs.data();
    a.ld_ref = scan.ld;
    a.fp32_band = fp32_band(scan.dim);
    a.out_idx = out_idx;
    a.out_dist = out_dist;
    a.escalate_count = ws.counters.data();
    a.escalate_list = ws.escalate_list.data();
    a.stats = reinterpret_cast<kern::MixedStats*>(ws.counters.data() + 4);
    kern::select_rerank_i8(a, stream);
    ws.mark(3, stream);
    kern::ScanArgs esc;
    esc.refs = scan.refs.data();
    esc.ld_ref = scan.ld;
    esc.rows = static_cast<uint32_t>(scan.rows);
    esc.queries = queries;
    esc.ld_q = ld_q;
    esc.nq = n;
    esc.dim = static_cast<uint32_t>(scan.dim);
    esc.keys = ws.keys.data();
    kern::escalate_scan(esc, ws.counters.data(), ws.escalate_list.data(), n, ws.esc_queries.data(), scan.ld,
                        out_idx, out_dist, stream);
    ws.mark(4, stream);
    if (ws.timing) {
        check_hip(hipEventSynchronize(ws.events[4]), "HIP event wait failed");
        for (int i = 0; i < Workspace::kStages; ++i) {
            check_hip(hipEventElapsedTime(&ws.stage_ms[i], ws.events[i], ws.events[i + 1]),
                      "cannot read HIP event time");
        }
    }
}

// Runs one batch with `strategy` on `stream`.
void run_strategy(const Scan& scan, Workspace& ws, const float* queries, uint64_t ld_q, bool padded,
                  uint64_t count, uint64_t* out_idx, float* out_dist, hipStream_t stream,
                  int strategy)
{
    if (strategy == UHANN_STRATEGY_DIRECT) {
        scan_launch(scan, queries, ld_q, count, ws.keys.data(), out_idx, out_dist, stream, padded,
                    nullptr, &ws);
    } else if (strategy == UHANN_STRATEGY_GRAM_BF16) {
        mixed_launch(scan, ws, queries, ld_q, count, out_idx, out_dist, stream);
    } else if (strategy == UHANN_STRATEGY_GRAM_FP8) {
        fp8_launch(scan, ws, queries, ld_q, count, out_idx, out_dist, stream);
    } else if (strategy == UHANN_STRATEGY_GRAM_INT8) {
        i8_launch(scan, ws, queries, ld_q, count, out_idx, out_dist, stream);
    } else {
        throw std::invalid_argument("unknown query strategy");
    }
}

}  // namespace uhann

using namespace uhann;

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
        const std::string label = device_label(prop);
        std::strncpy(buf, label.c_str(), cap - 1);
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
        run_strategy(*scan, *ws, ws->queries.data(), scan->ld, true, count, ws->out_idx.data(),
                     ws->out_dist.data(), ws->stream, strategy);
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
        if (count == 0) return;
        if (scan->dim != 0 && device_queries == nullptr) {
            throw std::invalid_argument("device query pointer must not be null");
        }
        select_device(scan->device);
        run_strategy(*scan, *ws, device_queries, scan->dim, false, count, ws->out_idx.data(),
                     ws->out_dist.data(), ws->stream, strategy);
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
        if (count == 0) return;
        if (scan->dim != 0 && device_queries == nullptr) {
            throw std::invalid_argument("device query pointer must not be null");
        }
        if (ld_queries < scan->dim) throw std::invalid_argument("query row stride is below dim");
        select_device(scan->device);
        run_strategy(*scan, *ws, device_queries, ld_queries, false, count, device_out_idx,
                     device_out_dist, static_cast<hipStream_t>(stream), strategy);
    });
}

int uhann_scan_query_device_noalloc(const uhann_scan* scan, const float* device_queries, uint64_t ld_queries,
                                    uint64_t count, uint64_t* device_out_idx, float* device_out_dist, void* stream)
{
    return guarded([&] {
        if (scan == nullptr) throw std::invalid_argument("scan handle must not be null");
        if (count > kern::noalloc_max_queries()) {
            throw std::length_error("the allocation-free query takes at most 65536 queries per call");
        }
        if (count == 0) return;
        if (scan->dim != 0 && device_queries == nullptr) {
            throw std::invalid_argument("device query pointer must not be null");
        }
        if (ld_queries < scan->dim) throw std::invalid_argument("query row stride is below dim");
        select_device(scan->device);
        unsigned long long* keys = kern::noalloc_keys(scan->device);
        scan_launch(*scan, device_queries, ld_queries, count, keys, device_out_idx, device_out_dist,
                    static_cast<hipStream_t>(stream), false, nullptr, nullptr);
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
        const bool pair_tile = rpt == 0 && ((tr == 0 && (tq == 0 || tq == 1 || tq == 2 || tq == 4)) ||
                                            (tr == 8 && (tq == 4 || tq == 8)));
        if (!pair_tile && !kern::scan_tile_supported(tile)) throw std::invalid_argument("unsupported scan tile");
        if (count == 0) return;
        if (scan->dim != 0 && device_queries == nullptr) {
            throw std::invalid_argument("device query pointer must not be null");
        }
        if (ld_queries < scan->dim) throw std::invalid_argument("query row stride is below dim");
        select_device(scan->device);
        scan_launch(*scan, device_queries, ld_queries, count, ws->keys.data(), device_out_idx,
                    device_out_dist, static_cast<hipStream_t>(stream), false, &tile);
    });
}

int uhann_scan_gram_estimates(const uhann_scan* scan, uhann_workspace* ws, const float* device_queries,
                              uint64_t ld_queries, uint64_t count, float* device_out, uint64_t ld_out, int tile,
                              void* stream)
{
    return guarded([&] {
        if (scan == nullptr) throw std::invalid_argument("scan handle must not be null");
        checked_ws(scan, ws, count);
        if (scan->gram == nullptr) throw std::invalid_argument("mixed-precision strategy not enabled for this index");
        if (tile < -1 || tile > kern::gram_tile_count() - 1) throw std::invalid_argument("unknown Gram tile");
        if (count == 0) return;
        if (device_queries == nullptr || device_out == nullptr) {
            throw std::invalid_argument("device pointers must not be null");
        }
        if (ld_queries < scan->dim) throw std::invalid_argument("query row stride is below dim");
        if (ld_out < scan->rows) throw std::invalid_argument("output row stride is below the reference count");
        select_device(scan->device);
        ws->reserve_mixed(*scan);
        const GramRefs& g = *scan->gram;
        const auto s = static_cast<hipStream_t>(stream);
        const uint32_t n = static_cast<uint32_t>(count);
        kern::to_bf16_rows(device_queries, ld_queries, n, static_cast<uint32_t>(scan->dim), ws->q16.data(), g.ldk,
                           ws->qstats.data(), s);
        kern::gram_bf16(ws->q16.data(), g.ldk, n, g.refs16.data(), g.ldk, static_cast<uint32_t>(scan->rows),
                        static_cast<uint32_t>(g.ldk), ws->qstats.data(), g.stats.data(), device_out, ld_out,
                        tile < 0 ? gram_tile() : tile, tile < 0 ? ws->gram_partials.data() : nullptr,
                        ws->gram_partials.bytes() / sizeof(float), s);
    });
}

int uhann_scan_distance_table(const uhann_scan* scan, const float* device_queries, uint64_t ld_queries,
                              uint64_t count, double* device_out, uint64_t ld_out, void* stream)
{
    return guarded([&] {
        if (scan == nullptr) throw std::invalid_argument("scan handle must not be null");
        if (scan->weighted) throw std::invalid_argument("distance tables are defined for unweighted references");
        if (count == 0 || scan->rows == 0) return;
        if (device_queries == nullptr || device_out == nullptr) {
            throw std::invalid_argument("device pointers must not be null");
        }
        if (ld_queries < scan->dim) throw std::invalid_argument("query row stride is below dim");
        if (ld_out < scan->rows) throw std::invalid_argument("output row stride is below the reference count");
        if (count > 0xffffffffULL || scan->rows > 0xffffffffULL || scan->dim > 0xffffffffULL) {
            throw std::length_error("distance table dimensions exceed 32-bit indexing");
        }
        select_device(scan->device);
        kern::distance_table(scan->metric == UHANN_METRIC_L1, device_queries, ld_queries,
                             static_cast<uint32_t>(count), scan->refs.data(), scan->ld,
                             static_cast<uint32_t>(scan->rows), static_cast<uint32_t>(scan->dim), device_out, ld_out,
                             static_cast<hipStream_t>(stream));
    });
}

int uhann_scan_fp8_estimates(const uhann_scan* scan, uhann_workspace* ws, const float* device_queries,
                             uint64_t ld_queries, uint64_t count, float* device_out, uint64_t ld_out, void* stream)
{
    return guarded([&] {
        if (scan == nullptr) throw std::invalid_argument("scan handle must not be null");
        checked_ws(scan, ws, count);
        if (scan->fp8 == nullptr) throw std::invalid_argument("fp8 strategy not enabled for this index");
        if (count == 0) return;
        if (device_queries == nullptr || device_out == nullptr) {
            throw std::invalid_argument("device pointers must not be null");
        }
        if (ld_queries < scan->dim) throw std::invalid_argument("query row stride is below dim");
        if (ld_out < scan->rows) throw std::invalid_argument("output row stride is below the reference count");
        select_device(scan->device);
        ws->reserve_fp8(*scan);
        const Fp8Refs& f = *scan->fp8;
        const auto s = static_cast<hipStream_t>(stream);
        const uint32_t n = static_cast<uint32_t>(count);
        kern::to_fp8_rows(device_queries, ld_queries, n, static_cast<uint32_t>(scan->dim), ws->q8.data(), f.ldk8,
                          ws->q8stats.data(), s);
        kern::gram_fp8(ws->q8.data(), f.ldk8, n, f.refs8.data(), f.ldk8, static_cast<uint32_t>(scan->rows),
                       static_cast<uint32_t>(f.ldk8), ws->q8stats.data(), f.stats.data(), device_out, ld_out,
                       ws->gram_partials.data(), ws->gram_partials.bytes() / sizeof(float), s);
    });
}

int uhann_scan_int8_estimates(const uhann_scan* scan, uhann_workspace* ws, const float* device_queries,
                              uint64_t ld_queries, uint64_t count, float* device_out, uint64_t ld_out, int tile,
                              void* stream)
{
    return guarded([&] {
        if (scan == nullptr) throw std::invalid_argument("scan handle must not be null");
        checked_ws(scan, ws, count);
        if (scan->i8 == nullptr) throw std::invalid_argument("int8 strategy not enabled for this index");
        if (tile < -1 || tile >= kern::gram_i8_tile_count()) throw std::invalid_argument("unknown int8 Gram tile");
        if (count == 0) return;
        if (device_queries == nullptr || device_out == nullptr) {
            throw std::invalid_argument("device pointers must not be null");
        }
        if (ld_queries < scan->dim) throw std::invalid_argument("query row stride is below dim");
        if (ld_out < scan->rows) throw std::invalid_argument("output row stride is below the reference count");
        select_device(scan->device);
        ws->reserve_i8(*scan);
        const I8Refs& f = *scan->i8;
        const auto s = static_cast<hipStream_t>(stream);
        const uint32_t n = static_cast<uint32_t>(count);
        kern::to_i8x2_rows(device_queries, ld_queries, n, static_cast<uint32_t>(scan->dim), ws->qi8.data(), f.ldk,
                           ws->qi8stats.data(), s);
        kern::gram_i8x2(ws->qi8.data(), f.ldk, n, f.refs.data(), f.ldk, static_cast<uint32_t>(scan->rows),
                        static_cast<uint32_t>(f.ldk), ws->qi8stats.data(), f.stats.data(), device_out, ld_out, tile,
                        ws->gram_partials.data(), ws->gram_partials.bytes() / sizeof(float), s);
    });
}

int uhann_int8_rows(const float* device_rows, uint64_t ld_rows, uint64_t count, uint64_t dim, uint8_t* device_bytes,
                    void* device_stats, void* stream)
{
    return guarded([&] {
        if (count == 0) return;
        if (device_rows == nullptr || device_bytes == nullptr || device_stats == nullptr) {
            throw std::invalid_argument("device pointers must not be null");
        }
        if (ld_rows < dim) throw std::invalid_argument("row stride is below dim");
        if (count > 0xffffffffULL || dim > kern::gram_i8_max_dim()) throw std::length_error("too many rows or columns");
        const uint64_t ld = std::max<uint64_t>(kern::gram_i8_k_padding(), (dim + 63) / 64 * kern::gram_i8_k_padding());
        kern::to_i8x2_rows(device_rows, ld_rows, static_cast<uint32_t>(count), static_cast<uint32_t>(dim),
                           device_bytes, ld, static_cast<kern::I8Stats*>(device_stats),
                           static_cast<hipStream_t>(stream));
    });
}

int uhann_scan_autotune(uhann_scan* scan, uhann_workspace* ws, const float* device_queries,
                        uint64_t ld_queries, uint64_t max_batch, int* table, int capacity, int* count)
{
    return guarded([&] {
        if (scan == nullptr) throw std::invalid_argument("scan handle must not be null");
        checked_ws(scan, ws, max_batch);
        if (max_batch == 0) throw std::invalid_argument("autotune batch must be positive");
        if (scan->dim != 0 && device_queries == nullptr) throw std::invalid_argument("device query pointer must not be null");
        if (ld_queries < scan->dim) throw std::invalid_argument("query row stride is below dim");
        select_device(scan->device);
        autotune_scan(*scan, *ws, device_queries, ld_queries, false, max_batch, ws->stream);
        int n = 0;
        for (int b = 0; b < 24; ++b) {
            const Scan::Tuned& t = scan->tuned[b];
            if (!t.set) continue;
            if (table != nullptr && n < capacity) {
                int* row = table + 6 * n;
                row[0] = b;
                row[1] = t.tile.rpt;
                row[2] = t.tile.tq;
                row[3] = t.tile.tr;
                row[4] = t.tile.group_q;
                row[5] = static_cast<int>(t.ms * 1000.0F);  // microseconds
            }
            ++n;
        }
        if (count != nullptr) *count = n;
    });
}

int uhann_scan_enable_mixed(uhann_scan* scan, int strategy)
{
    return guarded([&] {
        if (scan == nullptr) throw std::invalid_argument("scan handle must not be null");
        scan->enable_mixed(strategy);
    });
}

int uhann_scan_set_fp8_shortlist(uhann_scan* scan, uint32_t rows)
{
    return guarded([&] {
        if (scan == nullptr) throw std::invalid_argument("scan handle must not be null");
        if (rows < 1 || rows > static_cast<uint32_t>(kern::kShortlist)) {
            throw std::invalid_argument("the FP8 shortlist holds 1 to 16 rows");
        }
        scan->fp8_shortlist = rows;
    });
}

int uhann_workspace_mixed_stats(uhann_workspace* ws, uint64_t* candidates, uint64_t* fp64_reranked,
                                uint64_t* escalated)
{
    return guarded([&] {
        if (ws == nullptr) throw std::invalid_argument("workspace handle must not be null");
        unsigned int host[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        if (ws->counters.bytes() != 0) {
            select_device(ws->device);
            check_hip(hipStreamSynchronize(ws->stream), "workspace stream failed");
            check_hip(hipMemcpy(host, ws->counters.data(), sizeof(host), hipMemcpyDeviceToHost),
                      "cannot read mixed-precision statistics");
        }
        if (candidates) *candidates = host[4];
        if (fp64_reranked) *fp64_reranked = host[5];
        if (escalated) *escalated = host[6];
    });
}

int uhann_workspace_shortlist_stats(uhann_workspace* ws, uint64_t* certified)
{
    return guarded([&] {
        if (ws == nullptr) throw std::invalid_argument("workspace handle must not be null");
        unsigned int host[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        if (ws->counters.bytes() != 0) {
            select_device(ws->device);
            check_hip(hipStreamSynchronize(ws->stream), "workspace stream failed");
            check_hip(hipMemcpy(host, ws->counters.data(), sizeof(host), hipMemcpyDeviceToHost),
                      "cannot read shortlist statistics");
        }
        if (certified) *certified = host[7];
    });
}

int uhann_workspace_set_timing(uhann_workspace* ws, int enable)
{
    return guarded([&] {
        if (ws == nullptr) throw std::invalid_argument("workspace handle must not be null");
        ws->timing = enable != 0;
    });
}

int uhann_workspace_stage_times(const uhann_workspace* ws, float* ms, int n)
{
    return guarded([&] {
        if (ws == nullptr || ms == nullptr) throw std::invalid_argument("null argument");
        for (int i = 0; i < n && i < Workspace::kStages; ++i) ms[i] = ws->stage_ms[i];
    });
}

}  // extern "C"
