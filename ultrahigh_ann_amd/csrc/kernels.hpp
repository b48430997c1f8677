// Launch interfaces of the HIP kernels (implemented in *.hip). Host code in
// engine.cpp validates shapes before calling any of these; every launcher
// checks the launch status and throws std::runtime_error on failure.
#pragma once

#include <hip/hip_runtime_api.h>

#include <cstdint>
#include <vector>

namespace uhann::kern {

// Tile shapes of the kernel bodies (bodies.hpp), shared with host code that
// counts tiles.
namespace body {
constexpr int kGatherThreads = 256;
constexpr int kGatherChunk = 1024;
constexpr int kGatherRows = 8;  // query rows per gather tile
constexpr int kProjThreads = 256;
constexpr int kProjQ = 4;
constexpr int kProjP = 8;
constexpr int kProjWaves = 4;  // waves of the bf16 projection tile (32 x 32 outputs)
constexpr int kMedianThreads = 256;
}  // namespace body

enum class Metric : int { l2 = 0, l1 = 1 };

// One dense FP32 scan: `nq` query rows against `rows` reference rows, both
// row-major with leading dimensions ld_q / ld_ref (in floats), `dim` scanned
// coordinates. weights (L1 only) holds one float per coordinate or is null.
// keys[q] receives atomicMin of ((distance bits) << 32 | row); it must hold
// ~0 before the launch (fill_keys).
struct ScanArgs {
    const float* refs = nullptr;
    uint64_t ld_ref = 0;
    uint32_t rows = 0;
    const float* queries = nullptr;
    uint64_t ld_q = 0;
    uint32_t nq = 0;
    const float* weights = nullptr;
    uint32_t dim = 0;
    unsigned long long* keys = nullptr;
    // Every reference value is read once by the launch (one query tile):
    // load references with the non-temporal cache policy so the stream does
    // not evict the queries from L2.
    bool stream_refs = false;
    // Tile loads through buffer resources (one at the tile's first reference
    // row, one at its first query row, rows as scalar byte offsets): set by
    // the launcher when every such offset stays below 2^32.
    bool buffers = false;
};

// Register tile of the direct scan: a block of 256/rpt threads, each thread
// owning `rpt` consecutive residues of the 256-column stride, accumulating a
// tq x tr tile of (query, reference) pairs. group_q q-tiles share one sweep
// over r-tiles (L2 locality). weight_fma selects fmaf(w,|d|,acc) for
// weighted L1 (else rounded product, then add). rpt 2 is the shared-query
// tile: group_q reference groups of tr rows and tq queries per block of
// 256 group_q threads, one chain per thread, the queries' 256-column steps
// staged in LDS for all groups.
struct ScanTile {
    int rpt = 1;
    int tq = 8;
    int tr = 8;
    int group_q = 8;
};

bool scan_tile_supported(const ScanTile& tile);
// Every tiled shape compiled in (candidates of the autotuner).
const std::vector<ScanTile>& scan_tiles();
void fill_keys(unsigned long long* keys, uint32_t n, hipStream_t stream);
// The allocation-free direct query's key array (scan.hip `scratch`): its
// address on `device`, which must be current (resolved once per device), and
// its capacity in queries.
unsigned long long* noalloc_keys(int device);
uint32_t noalloc_max_queries();
void scan_direct(Metric metric, bool weight_fma, const ScanArgs& args,
                 const ScanTile& tile, hipStream_t stream);
void finalize_keys(const unsigned long long* keys, uint32_t n, uint64_t* out_idx,
                   float* out_dist, hipStream_t stream);
// Exact binary32 L2 scan of the queries listed in list[0 .. min(*count, cap))
// (indices into base.queries, row stride base.ld_q), with the count read on
// the device: the listed rows are copied to staged (cap x ld_staged floats,
// ld_staged >= base.dim), base.keys (cap entries) receives their keys, and
// out_idx / out_dist[list[i]] the results. No host synchronisation.
void escalate_scan(const ScanArgs& base, const unsigned int* count, const unsigned int* list, uint32_t cap,
                   float* staged, uint64_t ld_staged, uint64_t* out_idx, float* out_dist, hipStream_t stream);

// Thread-per-query direct scan for 1 <= dim <= scan_short_max_dim(): refs
// row-major (row stride ld_ref), one thread per query walking 128 references
// staged in LDS. Same keys contract as scan_direct, same bits.
uint32_t scan_short_max_dim();
void scan_short(Metric metric, bool weight_fma, const float* refs, uint64_t ld_ref, uint32_t rows,
                const float* queries, uint64_t ld_q, uint32_t nq, const float* weights, uint32_t dim,
                unsigned long long* keys, hipStream_t stream);

// Pair-per-thread direct scan for dim <= scan_pairs_max_dim(): refs_t is the
// transposed reference matrix (row stride ld_t), dim rows followed by
// kPairsPadCols rows of zeros; tq queries per block, rows < kPairsMaxRows.
// classes: the chain-class body (tq 4 or 8; it reads up to 28 of the zero
// rows, whose terms add +0) instead of whole 32-chain groups (tq 1, 2 or 4).
// Same keys contract as scan_direct, same bits.
constexpr uint32_t kPairsMaxRows = 1U << 25;
constexpr uint32_t kPairsPadCols = 32;
uint32_t scan_pairs_max_dim();
void scan_pairs(Metric metric, bool weight_fma, const float* refs_t, uint64_t ld_t, uint32_t rows,
                const float* queries, uint64_t ld_q, uint32_t nq, const float* weights, uint32_t dim,
                unsigned long long* keys, int tq, bool classes, hipStream_t stream);

// Gather of sampled coordinates: out[q][j] = src[q][cols[j]] (* scales[j]
// when scales is non-null), for q < nq, j < k. Sets *invalid to 1 when a
// gathered (or scaled) value is not finite.
void gather_columns(const float* src, uint64_t ld_src, uint32_t nq,
                    const uint32_t* cols, const float* scales, uint32_t k,
                    float* out, uint64_t ld_out, int* invalid, hipStream_t stream);

// Elementwise finiteness check over n floats; sets *invalid to 1 on a
// non-finite value.
void check_finite(const float* values, uint64_t n, int* invalid, hipStream_t stream);

}  // namespace uhann::kern
