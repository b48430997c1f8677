// Engine objects behind the C ABI handles (uhann_scan, uhann_workspace and
// the index handles). Internal to the library.
#pragma once

#include <hip/hip_runtime_api.h>

#include <cstdint>
#include <string>

#include <memory>

#include "common.hpp"
#include "gram.hpp"
#include "kernels.hpp"

namespace uhann {

const char* last_error() noexcept;
uint64_t padded_ld(uint64_t dim);
void validate_finite_host(const float* values, std::size_t n, const char* what);

// BF16 copy of the references for the mixed-precision strategy: rows x ldk
// (ldk = dim rounded up to the Gram K step, zero padded) plus row statistics.
struct GramRefs {
    uint64_t ldk = 0;
    DeviceBuffer<__bf16> refs16;
    DeviceBuffer<kern::RowStats> stats;
};

// Per-row-scaled e4m3 copy of the references for GRAM_FP8: rows padded to
// whole Gram tiles, ldk8 = dim rounded up to the FP8 K step (bytes), plus
// row statistics (scale included).
struct Fp8Refs {
    uint64_t ldk8 = 0;
    DeviceBuffer<uint8_t> refs8;
    DeviceBuffer<kern::RowStats> stats;
};

// 16-bit fixed-point copy of the references for GRAM_INT8: two int8 planes
// interleaved per 64 columns (ldk8x2 = 128 ceil(dim / 64) bytes per row),
// rows padded to whole Gram tiles, plus row statistics.
struct I8Refs {
    uint64_t ldk = 0;
    DeviceBuffer<uint8_t> refs;
    DeviceBuffer<kern::I8Stats> stats;
};

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
    DeviceBuffer<float> refs_t;  // (dim + kPairsPadCols) x rows copy for short rows (pair-per-thread scan)
    DeviceBuffer<float> weights;
    uint64_t payload_bytes = 0;
    std::string device_name;
    std::unique_ptr<GramRefs> gram;  // set by enable_mixed(GRAM_BF16)
    std::unique_ptr<Fp8Refs> fp8;    // set by enable_mixed(GRAM_FP8)
    std::unique_ptr<I8Refs> i8;      // set by enable_mixed(GRAM_INT8)
    uint32_t fp8_shortlist = kern::kShortlist;  // GRAM_FP8: rows re-ranked per query

    // Autotuned direct-scan kernels per batch bucket b (batches of
    // 2^(b-1)+1 .. 2^b queries); empty entries fall back to the heuristic.
    struct Tuned {
        bool set = false;
        kern::ScanTile tile;
        float ms = 0.0F;
    };
    Tuned tuned[24];

    void enable_mixed(int strategy);
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

    // Mixed-precision buffers, allocated on the first mixed query.
    DeviceBuffer<__bf16> q16;
    DeviceBuffer<kern::RowStats> qstats;
    DeviceBuffer<float> approx;
    DeviceBuffer<unsigned int> counters;  // [0] escalations, [1..4] MixedStats
    DeviceBuffer<unsigned int> escalate_list;
    DeviceBuffer<float> esc_queries;    // escalated queries, gathered on the device
    DeviceBuffer<uint8_t> q8;           // FP8 strategy: e4m3 queries
    DeviceBuffer<kern::RowStats> q8stats;
    DeviceBuffer<uint8_t> qi8;          // INT8 strategy: fixed-point queries
    DeviceBuffer<kern::I8Stats> qi8stats;
    DeviceBuffer<float> gram_partials;  // K-split partial sums of the Gram tail tiles
    void reserve_partials(uint64_t floats);
    void reserve_mixed(const Scan& scan);
    void reserve_fp8(const Scan& scan);
    void reserve_i8(const Scan& scan);
    void reserve_common(const Scan& scan);

    // Optional stage timing (HIP events on the launch stream):
    // [0] bf16 conversion, [1] Gram GEMM, [2] select + re-rank, [3] escalation.
    static constexpr int kStages = 4;
    bool timing = false;
    hipEvent_t events[kStages + 1] = {};
    float stage_ms[kStages] = {};
    void mark(int i, hipStream_t s);
};

kern::ScanTile choose_tile(uint64_t nq, bool aligned4);
uint32_t pairs_max_dim();
int batch_bucket(uint64_t nq);

// Times every candidate direct-scan kernel for batches of 1, 2, 4, ... up to
// max_batch queries (row stride ld_q at `queries`) and records the fastest per
// bucket in scan.tuned.
void autotune_scan(Scan& scan, Workspace& ws, const float* queries, uint64_t ld_q, bool padded,
                   uint64_t max_batch, hipStream_t stream);

// fill keys -> direct scan -> finalize, all on `stream`. queries_padded
// says every query row is readable up to the next multiple of 4 floats past
// dim (workspace-staged rows are). forced_tile (or null) overrides the tile.
// With `timed` (a workspace whose timing is on) stage 0 receives the scan
// kernel's own duration.
void scan_launch(const Scan& scan, const float* queries, uint64_t ld_q, uint64_t count,
                 unsigned long long* keys, uint64_t* out_idx, float* out_dist, hipStream_t stream,
                 bool queries_padded, const kern::ScanTile* forced_tile, Workspace* timed = nullptr);

// Mixed-precision L2 query (BF16 Gram -> candidate set -> exact re-rank ->
// binary64 near ties -> direct scan for escalated queries, on the device).
void mixed_launch(const Scan& scan, Workspace& ws, const float* queries, uint64_t ld_q,
                  uint64_t count, uint64_t* out_idx, float* out_dist, hipStream_t stream);

// FP8 shortlist query (e4m3 Gram -> top-16 per query -> exact re-rank ->
// binary64 near ties); no host synchronisation.
void fp8_launch(const Scan& scan, Workspace& ws, const float* queries, uint64_t ld_q, uint64_t count,
                uint64_t* out_idx, float* out_dist, hipStream_t stream);

// Certified INT8 query (fixed-point int8 Gram with exact int32 sums ->
// deterministic candidate set -> exact re-rank -> binary64 near ties ->
// direct scan for escalated queries, on the device).
void i8_launch(const Scan& scan, Workspace& ws, const float* queries, uint64_t ld_q, uint64_t count,
               uint64_t* out_idx, float* out_dist, hipStream_t stream);

// Any strategy's batch on `stream` (DIRECT, GRAM_BF16, GRAM_FP8, GRAM_INT8).
void run_strategy(const Scan& scan, Workspace& ws, const float* queries, uint64_t ld_q, bool padded,
                  uint64_t count, uint64_t* out_idx, float* out_dist, hipStream_t stream, int strategy);

}  // namespace uhann

// C ABI handle types (include/uhann.h) backed by the engine objects.
struct uhann_scan : uhann::Scan {
    using Scan::Scan;
};

struct uhann_workspace : uhann::Workspace {
    using Workspace::Workspace;
};
