// Grouped kernels of a serving tick (serve.hip) and the descriptor layout the
// host writes for them (server.cpp).
//
// A tick is a set of requests, each one batch of queries for one resident
// index. The host lays the tick out as one descriptor blob: a ServeHeader, one
// ServeItem per request and, for every stage, a list of ServeEntry records
// (item, first unit, units per row of the item's unit grid). Every stage is one
// persistent launch of a fixed grid whose blocks walk the stage's units
// (u = blockIdx.x, blockIdx.x + gridDim.x, ...), find the owning item by
// binary search over the entries and run the same tile body as the one-index
// kernels (bodies.hpp). A request's arithmetic therefore does not depend on
// which other requests share its tick.
#pragma once

#include <hip/hip_runtime_api.h>

#include <cstdint>

namespace uhann::kern {

enum ServeStage : int {
    kServeCheck = 0,   // finiteness of full-width exact queries (unit = query row)
    kServeGather,      // sampled coordinates (unit = 1024-column chunk x 8 rows)
    kServeProject32,   // binary32 projection, K7 order (unit = 4 queries x 8 outputs)
    kServeProject16,   // bf16 MFMA projection (unit = 32 queries x 32 outputs)
    kServePairsShortL2,   // pair-per-thread scans of rows up to kServePairsShortDim
    kServePairsShortL1,   // (unit = 256 references x pairs_tq queries; less LDS per block)
    kServePairsShortL1W,
    kServePairsL2,     // pair-per-thread scans (unit = 256 references x pairs_tq queries)
    kServePairsL1,
    kServePairsL1W,
    kServeTileL2,      // tiled direct scans (unit = 8 queries x 8 references)
    kServeTileL1,
    kServeTileL1W,
    kServeMedian8,     // median-L1 scans by padded p (unit = 1 query x 256 references)
    kServeMedian16,
    kServeMedian32,
    kServeMedian64,
    kServeMedian128,
    kServeStages
};

constexpr uint32_t kServeTileQ = 8;
constexpr uint32_t kServeTileR = 8;
// Pair stages take scanned dimensions up to this (queries and weights in
// (pairs_tq + 1) * kServePairsMaxDim floats of LDS, 28 more columns at
// pairs_tq 4 (the chain-class body); at most 62 KB). Rows up to
// kServePairsShortDim go to the short pair stages, whose blocks reserve LDS
// for that many columns only, so more of them share a CU.
constexpr uint32_t kServePairsMaxDim = 3072;
constexpr uint32_t kServePairsShortDim = 1024;

struct ServeItem {
    const float* queries;    // n full-width queries, row stride ld_q
    uint64_t ld_q;
    uint32_t n;
    uint32_t qoff;           // first row of the request in the tick's keys and outputs
    uint32_t k;              // sampled coordinates (flat, hierarchical)
    uint32_t p;              // projection dimension (hierarchical)
    const uint32_t* cols;    // sampled source columns
    const float* scales;     // flat L2 importance scales, else null
    float* gathered;         // n x ld_g
    uint64_t ld_g;
    const float* proj;       // binary32 projection, p x ld_proj
    uint64_t ld_proj;
    const void* proj16;      // bf16 projection, ld_proj16 elements per row
    uint64_t ld_proj16;
    float* projected;        // n x ld_p
    uint64_t ld_p;
    const float* refs;       // scanned references, row-major (tile stages)
    uint64_t ld_ref;
    const float* refs_t;     // transposed references, row stride `rows` (pair and median stages)
    const float* weights;    // flat L1 coordinate weights
    uint32_t rows;
    uint32_t dim;            // scanned coordinates
    const float* scan_q;     // queries of the scan stage (full, gathered or projected)
    uint64_t ld_sq;
    int* flags;              // [0] non-finite query value, [1] projection, [2] median
};

struct ServeEntry {
    uint32_t item;
    uint32_t begin;  // first unit of the item in this stage
    uint32_t fast;   // units along the item's first grid axis
    uint32_t pad;
};

struct ServeHeader {
    uint32_t nq;     // rows of keys and outputs used by the tick
    uint32_t items;
    uint32_t count[kServeStages];   // entries per stage
    uint32_t units[kServeStages];   // units per stage
    uint32_t offset[kServeStages];  // byte offset of the stage's entries in the blob
    uint32_t items_offset;          // byte offset of the ServeItem array
};

// Device buffers of one tick slot. blob holds the descriptor; keys, outputs
// and flags have room for max_q rows and max_items requests.
struct ServeLaunch {
    const void* blob = nullptr;
    unsigned long long* keys = nullptr;
    uint64_t* out_idx = nullptr;
    float* out_dist = nullptr;
    int* flags = nullptr;
    uint32_t max_q = 0;
    uint32_t max_items = 0;
    int pairs_tq = 4;
    uint32_t grid = 1024;
};

// Launches, on `stream`, the key/flag reset, the stages whose bit is set in
// stage_mask (in stage order) and the key finalisation. The launch shapes
// depend only on `l` and stage_mask, so the sequence can be captured once and
// replayed for any descriptor contents.
void serve_tick(const ServeLaunch& l, uint32_t stage_mask, hipStream_t stream);
bool serve_pairs_tq_supported(int tq);

}  // namespace uhann::kern
