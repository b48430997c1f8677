// Serving loop behind uhann_server (include/uhann.h): resident indexes
// admitted by an HBM budget, a request queue, and a worker thread that packs
// queued requests into ticks of grouped persistent kernels (serve.hip).
//
// Tick life cycle (two tick slots, so one tick is prepared while the other
// runs): take queued requests of resident indexes in arrival order while the
// tick's rows, requests, scratch and staging last -> write the descriptor blob
// and stage host queries in pinned memory -> copy both to the device, run the
// tick's kernels (a captured hipGraph per (slot, stage set), replayed), copy
// results and error flags back, record an event -> when the event completes,
// hand each request its rows and status. Latency is measured from submit to
// that hand-off.

#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "common.hpp"
#include "engine.hpp"
#include "hier.hpp"
#include "index.hpp"
#include "kernels.hpp"
#include "serve.hpp"
#include "uhann.h"

using namespace uhann;
using Clock = std::chrono::steady_clock;

namespace {

template <class T> class PinnedBuffer {
  public:
    PinnedBuffer() = default;
    explicit PinnedBuffer(std::size_t count) : count_(count)
    {
        if (count == 0) return;
        check_hip(hipHostMalloc(reinterpret_cast<void**>(&data_), checked_mul(count, sizeof(T), "pinned buffer"),
                                hipHostMallocDefault),
                  "cannot allocate pinned host memory");
    }
    PinnedBuffer(const PinnedBuffer&) = delete;
    PinnedBuffer& operator=(const PinnedBuffer&) = delete;
    PinnedBuffer(PinnedBuffer&& o) noexcept : count_(std::exchange(o.count_, 0)), data_(std::exchange(o.data_, nullptr)) {}
    PinnedBuffer& operator=(PinnedBuffer&& o) noexcept
    {
        if (this != &o) {
            release();
            count_ = std::exchange(o.count_, 0);
            data_ = std::exchange(o.data_, nullptr);
        }
        return *this;
    }
    ~PinnedBuffer() { release(); }
    T* data() const { return data_; }
    std::size_t size() const { return count_; }

  private:
    void release()
    {
        if (data_ != nullptr) (void)hipHostFree(data_);
        data_ = nullptr;
    }
    std::size_t count_ = 0;
    T* data_ = nullptr;
};

struct ServedIndex {
    uint64_t id = 0;
    int state = UHANN_INDEX_PENDING;
    uhann_index_spec spec{};
    std::vector<uhann_sampled_column> sample;
    std::vector<double> projection;
    uint64_t estimate = 0;      // device bytes reserved before the build
    uint64_t device_bytes = 0;  // device bytes of the built index
    uhann_index* ix = nullptr;
    uint64_t inflight = 0;      // submitted, not yet completed
    bool removing = false;
    bool building = false;      // popped from the admission queue, build running
    int fail_status = UHANN_OK;
    std::string fail_message;
};

struct Request {
    uint64_t id = 0;
    ServedIndex* index = nullptr;
    const float* queries = nullptr;
    bool on_device = false;
    uint64_t ld = 0;
    uint64_t n = 0;
    uint64_t* out_idx = nullptr;
    float* out_dist = nullptr;
    Clock::time_point t_submit;
    bool done = false;
    int status = UHANN_OK;
    std::string error;
};

struct Slot {
    DeviceBuffer<unsigned char> blob;
    PinnedBuffer<unsigned char> blob_host;
    DeviceBuffer<float> staging;
    PinnedBuffer<float> staging_host;
    DeviceBuffer<float> scratch;
    DeviceBuffer<unsigned long long> keys;
    DeviceBuffer<uint64_t> out_idx;
    DeviceBuffer<float> out_dist;
    DeviceBuffer<int> flags;
    PinnedBuffer<uint64_t> res_idx;
    PinnedBuffer<float> res_dist;
    PinnedBuffer<int> res_flags;
    hipEvent_t start = nullptr;
    hipEvent_t done = nullptr;
    std::vector<Request*> requests;
    std::vector<uint32_t> qoff;
    std::vector<char> zero;  // hierarchical index with one row: results are zeros
    uint32_t nq = 0;
    bool busy = false;
    uint64_t seq = 0;
};

constexpr uint64_t kAlign = 64;

uint64_t round_up(uint64_t v, uint64_t a) { return (v + a - 1) / a * a; }

uint64_t scan_bytes(uint64_t rows, uint64_t dim, bool weighted)
{
    uint64_t b = rows * padded_ld(dim) * 4;
    if (dim != 0 && dim <= kern::scan_pairs_max_dim()) b += dim * rows * 4;
    if (weighted) b += dim * 4;
    return b;
}

// Device bytes an index of this description allocates (index.cpp).
uint64_t estimate_bytes(const ServedIndex& e)
{
    const auto& s = e.spec;
    const uint64_t k = e.sample.size();
    if (s.kind == UHANN_KIND_EXACT) return scan_bytes(s.rows, s.dim, false);
    if (s.kind == UHANN_KIND_FLAT) {
        return scan_bytes(s.rows, k, s.metric == UHANN_METRIC_L1) + k * 4 + (s.metric == UHANN_METRIC_L2 ? k * 4 : 0);
    }
    const uint64_t p = s.projection_dim;
    const uint64_t proj = k != 0 ? p * padded_ld(k) * 4 : 16;
    if (s.metric == UHANN_METRIC_L2) {
        return scan_bytes(s.rows, p, false) + k * 4 + proj +
               kern::project_bf16_rows(p) * kern::project_bf16_ld(k) * 2;
    }
    return p * s.rows * 4 + k * 4 + proj;
}

uint64_t actual_bytes(const uhann_index& ix)
{
    uint64_t b = ix.cols.bytes() + ix.scales.bytes() + ix.proj.bytes() + ix.proj16.bytes() + ix.refs_t.bytes();
    if (ix.scan) b += ix.scan->refs.bytes() + ix.scan->refs_t.bytes() + ix.scan->weights.bytes();
    return b;
}

// Scratch floats (gathered + projected rows) and staging floats of a
// request; launch() lays them out the same way.
uint64_t request_scratch(const ServedIndex& e, uint64_t n)
{
    const uint64_t k_ld = padded_ld(e.sample.size());
    const uint64_t p_ld = padded_ld(e.spec.projection_dim);
    if (e.spec.kind == UHANN_KIND_FLAT) return round_up(n * k_ld, 4) + kAlign;
    if (e.spec.kind == UHANN_KIND_HIERARCHICAL) return round_up(n * k_ld, 4) + round_up(n * p_ld, 4) + 2 * kAlign;
    return 0;
}

uint64_t request_staging(const ServedIndex& e, uint64_t n, bool on_device)
{
    return on_device ? 0 : round_up(n * e.spec.dim, 4) + kAlign;
}

}  // namespace

struct uhann_server {
    uhann_server_config cfg{};
    std::mutex mu;
    std::condition_variable cv_work;  // worker: new requests, admissions, stop
    std::condition_variable cv_done;  // waiters: completions
    std::map<uint64_t, std::unique_ptr<ServedIndex>> indexes;
    std::deque<ServedIndex*> admission;
    uint64_t used_bytes = 0;
    uint64_t next_index = 1;
    uint64_t next_request = 1;
    std::deque<Request*> queue;
    std::unordered_map<uint64_t, std::unique_ptr<Request>> requests;
    uint64_t open_requests = 0;  // submitted, not completed
    Slot slots[2];
    uint64_t launch_seq = 0;
    hipStream_t stream = nullptr;
    std::map<uint64_t, hipGraphExec_t> graphs;
    std::thread worker;
    bool stop = false;

    // statistics since the last reset
    std::vector<double> latency_ms;
    uint64_t done_queries = 0;
    uint64_t ticks = 0;
    uint64_t tick_queries = 0;
    double gpu_busy_ms = 0.0;
    bool any_submit = false;
    Clock::time_point first_submit;
    Clock::time_point last_done;

    ~uhann_server();
    void worker_main();
    bool runnable() const;
    void take(Slot& slot);
    void launch(Slot& slot, int slot_index);
    void complete(Slot& slot);
    void fail_slot(Slot& slot, const std::string& message);
    void admit(std::unique_lock<std::mutex>& lk);
    void build(ServedIndex& e);
    kern::ServeLaunch launch_params(int slot_index);
    void run_tick(int slot_index, uint32_t mask);
};

namespace {

Slot make_slot(const uhann_server_config& c)
{
    Slot s;
    const int d = c.device;
    s.blob_host = PinnedBuffer<unsigned char>(sizeof(kern::ServeHeader) + kAlign +
                                              static_cast<std::size_t>(c.max_tick_requests) *
                                                  (sizeof(kern::ServeItem) + 4 * sizeof(kern::ServeEntry)) +
                                              kern::kServeStages * kAlign);
    s.blob = DeviceBuffer<unsigned char>(d, s.blob_host.size());
    s.staging_host = PinnedBuffer<float>(c.staging_floats);
    s.staging = DeviceBuffer<float>(d, c.staging_floats);
    s.scratch = DeviceBuffer<float>(d, c.scratch_floats);
    s.keys = DeviceBuffer<unsigned long long>(d, c.max_tick_queries);
    s.out_idx = DeviceBuffer<uint64_t>(d, c.max_tick_queries);
    s.out_dist = DeviceBuffer<float>(d, c.max_tick_queries);
    s.flags = DeviceBuffer<int>(d, 4ULL * c.max_tick_requests);
    s.res_idx = PinnedBuffer<uint64_t>(c.max_tick_queries);
    s.res_dist = PinnedBuffer<float>(c.max_tick_queries);
    s.res_flags = PinnedBuffer<int>(4ULL * c.max_tick_requests);
    check_hip(hipEventCreate(&s.start), "cannot create a HIP event");
    check_hip(hipEventCreate(&s.done), "cannot create a HIP event");
    return s;
}

}  // namespace

uhann_server::~uhann_server()
{
    {
        std::lock_guard<std::mutex> lk(mu);
        stop = true;
    }
    cv_work.notify_all();
    if (worker.joinable()) worker.join();
    (void)hipSetDevice(cfg.device);
    for (auto& [key, g] : graphs) (void)hipGraphExecDestroy(g);
    for (auto& s : slots) {
        if (s.start) (void)hipEventDestroy(s.start);
        if (s.done) (void)hipEventDestroy(s.done);
    }
    for (auto& [id, e] : indexes) uhann_index_destroy(e->ix);
    if (stream) (void)hipStreamDestroy(stream);
}

bool uhann_server::runnable() const
{
    for (const Request* r : queue) {
        if (r->index->state != UHANN_INDEX_PENDING) return true;
    }
    return false;
}

// Moves queued requests of resident (or failed) indexes into the slot, in
// arrival order, while the tick's capacities last. Requests of pending
// indexes stay queued.
void uhann_server::take(Slot& slot)
{
    slot.requests.clear();
    uint64_t rows = 0, scratch = 0, staging = 0;
    for (auto it = queue.begin(); it != queue.end();) {
        Request* r = *it;
        if (r->index->state == UHANN_INDEX_PENDING) {
            ++it;
            continue;
        }
        uint64_t need_scratch = 0, need_staging = 0;
        if (r->index->state == UHANN_INDEX_RESIDENT) {
            need_scratch = request_scratch(*r->index, r->n);
            need_staging = request_staging(*r->index, r->n, r->on_device);
        }
        if (slot.requests.size() == cfg.max_tick_requests || rows + r->n > cfg.max_tick_queries ||
            scratch + need_scratch > cfg.scratch_floats || staging + need_staging > cfg.staging_floats) {
            break;  // arrival order: later requests wait for the next tick
        }
        rows += r->n;
        scratch += need_scratch;
        staging += need_staging;
        slot.requests.push_back(r);
        it = queue.erase(it);
    }
}

kern::ServeLaunch uhann_server::launch_params(int slot_index)
{
    Slot& s = slots[slot_index];
    kern::ServeLaunch l;
    l.blob = s.blob.data();
    l.keys = s.keys.data();
    l.out_idx = s.out_idx.data();
    l.out_dist = s.out_dist.data();
    l.flags = s.flags.data();
    l.max_q = cfg.max_tick_queries;
    l.max_items = cfg.max_tick_requests;
    l.pairs_tq = cfg.pairs_tq;
    l.grid = 1024;
    return l;
}

void uhann_server::run_tick(int slot_index, uint32_t mask)
{
    const kern::ServeLaunch l = launch_params(slot_index);
    if (!cfg.use_graph) {
        kern::serve_tick(l, mask, stream);
        return;
    }
    const uint64_t key = (static_cast<uint64_t>(slot_index) << 32) | mask;
    auto found = graphs.find(key);
    if (found == graphs.end()) {
        hipGraph_t graph = nullptr;
        check_hip(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal), "cannot capture a serving tick");
        try {
            kern::serve_tick(l, mask, stream);
        } catch (...) {
            (void)hipStreamEndCapture(stream, &graph);
            if (graph) (void)hipGraphDestroy(graph);
            throw;
        }
        check_hip(hipStreamEndCapture(stream, &graph), "cannot capture a serving tick");
        hipGraphExec_t exec = nullptr;
        const hipError_t st = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        check_hip(st, "cannot instantiate a serving tick graph");
        found = graphs.emplace(key, exec).first;
    }
    check_hip(hipGraphLaunch(found->second, stream), "cannot launch a serving tick graph");
}

// Writes the tick's descriptor and host queries, then queues copies, kernels
// and result copies on the server stream. Runs without the lock: the slot's
// requests are owned by the worker until complete() and their indexes cannot
// be removed while requests are in flight.
void uhann_server::launch(Slot& slot, int slot_index)
{
    using namespace uhann::kern;
    const std::size_t n_items = slot.requests.size();
    std::vector<ServeItem> items(n_items);
    std::vector<ServeEntry> entries[kServeStages];
    uint32_t units[kServeStages] = {};
    slot.qoff.assign(n_items, 0);
    slot.zero.assign(n_items, 0);
    uint32_t rows = 0;
    uint64_t scratch = 0, staging = 0;
    auto add = [&](int stage, uint32_t item, uint32_t fast, uint64_t count) {
        if (count == 0) return;
        if (units[stage] + count > 0xffffffffULL) throw std::length_error("serving tick has too many units");
        entries[stage].push_back({item, units[stage], fast, 0});
        units[stage] += static_cast<uint32_t>(count);
    };
    auto ceil_div = [](uint64_t a, uint64_t b) { return (a + b - 1) / b; };
    for (std::size_t i = 0; i < n_items; ++i) {
        Request& r = *slot.requests[i];
        ServeItem& it = items[i];
        std::memset(&it, 0, sizeof it);
        slot.qoff[i] = rows;
        it.qoff = rows;
        it.n = static_cast<uint32_t>(r.n);
        it.flags = slot.flags.data() + 4 * i;
        rows += it.n;
        if (r.index->state != UHANN_INDEX_RESIDENT || r.n == 0) continue;  // failed index or empty request
        const uhann_index& ix = *r.index->ix;
        if (r.on_device) {
            it.queries = r.queries;
            it.ld_q = r.ld;
        } else {
            float* dst = slot.staging_host.data() + staging;
            for (uint64_t q = 0; q < r.n; ++q) std::memcpy(dst + q * ix.dim, r.queries + q * r.ld, ix.dim * sizeof(float));
            it.queries = slot.staging.data() + staging;
            it.ld_q = ix.dim;
            staging += round_up(r.n * ix.dim, 4) + kAlign;
        }
        const Scan* scan = ix.scan.get();
        const uint32_t idx = static_cast<uint32_t>(i);
        if (ix.kind == UHANN_KIND_HIERARCHICAL && ix.rows == 1) {
            slot.zero[i] = 1;  // the reference returns row 0 without any device work
            continue;
        }
        if (ix.kind == UHANN_KIND_EXACT) {
            it.scan_q = it.queries;
            it.ld_sq = it.ld_q;
            it.dim = static_cast<uint32_t>(ix.dim);
            add(kServeCheck, idx, 1, r.n);
        } else {
            it.k = static_cast<uint32_t>(ix.k);
            it.cols = ix.cols.data();
            it.scales = ix.kind == UHANN_KIND_FLAT && ix.metric == UHANN_METRIC_L2 ? ix.scales.data() : nullptr;
            it.gathered = slot.scratch.data() + scratch;
            it.ld_g = ix.k_ld;
            scratch += round_up(r.n * ix.k_ld, 4) + kAlign;
            const uint64_t chunks = ceil_div(ix.k, body::kGatherChunk);
            add(kServeGather, idx, static_cast<uint32_t>(chunks), chunks * ceil_div(r.n, body::kGatherRows));
            if (ix.kind == UHANN_KIND_FLAT) {
                it.scan_q = it.gathered;
                it.ld_sq = it.ld_g;
                it.dim = static_cast<uint32_t>(ix.k);
            } else {
                it.p = static_cast<uint32_t>(ix.p);
                it.proj = ix.proj.data();
                it.ld_proj = ix.k_ld;
                it.proj16 = ix.proj16.data();
                it.ld_proj16 = ix.proj16_ld;
                it.projected = slot.scratch.data() + scratch;
                it.ld_p = ix.p_ld;
                scratch += round_up(r.n * ix.p_ld, 4) + kAlign;
                it.scan_q = it.projected;
                it.ld_sq = it.ld_p;
                it.dim = static_cast<uint32_t>(ix.p);
                if (r.index->spec.strategy == UHANN_STRATEGY_PROJECTION_BF16) {
                    const uint64_t f = ceil_div(ix.p, 32);
                    add(kServeProject16, idx, static_cast<uint32_t>(f), f * ceil_div(r.n, 32));
                } else {
                    const uint64_t f = ceil_div(ix.p, body::kProjP);
                    add(kServeProject32, idx, static_cast<uint32_t>(f), f * ceil_div(r.n, body::kProjQ));
                }
            }
        }
        if (ix.kind == UHANN_KIND_HIERARCHICAL && ix.metric == UHANN_METRIC_L1) {
            it.refs_t = ix.refs_t.data();
            it.rows = static_cast<uint32_t>(ix.rows);
            const uint64_t f = ceil_div(ix.rows, body::kMedianThreads);
            const int stage = ix.p <= 8 ? kServeMedian8 : ix.p <= 16 ? kServeMedian16 : ix.p <= 32 ? kServeMedian32
                              : ix.p <= 64 ? kServeMedian64 : kServeMedian128;
            add(stage, idx, static_cast<uint32_t>(f), f * r.n);
            continue;
        }
        it.refs = scan->refs.data();
        it.ld_ref = scan->ld;
        it.refs_t = scan->refs_t.data();
        it.weights = scan->weights.data();
        it.rows = static_cast<uint32_t>(scan->rows);
        const bool l1 = scan->metric == UHANN_METRIC_L1;
        if (scan->refs_t.bytes() != 0 && it.dim != 0 && it.dim <= kServePairsMaxDim) {
            // Query tiles fastest: the units that run together share a
            // 256-row reference block, which then comes from L2.
            const uint64_t f = ceil_div(scan->rows, 256);
            const uint64_t nb = ceil_div(r.n, static_cast<uint64_t>(cfg.pairs_tq));
            const bool short_rows = it.dim <= kServePairsShortDim;
            const int stage = !l1 ? (short_rows ? kServePairsShortL2 : kServePairsL2)
                              : scan->weighted ? (short_rows ? kServePairsShortL1W : kServePairsL1W)
                                               : (short_rows ? kServePairsShortL1 : kServePairsL1);
            add(stage, idx, static_cast<uint32_t>(nb), f * nb);
        } else {
            const uint64_t f = ceil_div(scan->rows, kServeTileR);
            const int stage = !l1 ? kServeTileL2 : scan->weighted ? kServeTileL1W : kServeTileL1;
            add(stage, idx, static_cast<uint32_t>(f), f * ceil_div(r.n, kServeTileQ));
        }
    }
    slot.nq = rows;

    // Descriptor blob: header | items | entries of each stage.
    ServeHeader h{};
    h.nq = rows;
    h.items = static_cast<uint32_t>(n_items);
    uint64_t off = round_up(sizeof(ServeHeader), kAlign);
    h.items_offset = static_cast<uint32_t>(off);
    off += round_up(n_items * sizeof(ServeItem), kAlign);
    uint32_t mask = 0;
    for (int s = 0; s < kServeStages; ++s) {
        h.count[s] = static_cast<uint32_t>(entries[s].size());
        h.units[s] = units[s];
        h.offset[s] = static_cast<uint32_t>(off);
        off += round_up(entries[s].size() * sizeof(ServeEntry), kAlign);
        if (units[s] != 0) mask |= 1U << s;
    }
    if (off > slot.blob_host.size()) throw std::length_error("serving tick descriptor exceeds its buffer");
    unsigned char* b = slot.blob_host.data();
    std::memcpy(b, &h, sizeof h);
    if (n_items) std::memcpy(b + h.items_offset, items.data(), n_items * sizeof(ServeItem));
    for (int s = 0; s < kServeStages; ++s) {
        if (!entries[s].empty()) std::memcpy(b + h.offset[s], entries[s].data(), entries[s].size() * sizeof(ServeEntry));
    }

    check_hip(hipEventRecord(slot.start, stream), "cannot record a HIP event");
    check_hip(hipMemcpyAsync(slot.blob.data(), b, off, hipMemcpyHostToDevice, stream),
              "cannot copy a serving tick descriptor");
    if (staging != 0) {
        check_hip(hipMemcpyAsync(slot.staging.data(), slot.staging_host.data(), staging * sizeof(float),
                                 hipMemcpyHostToDevice, stream),
                  "cannot copy serving queries to the device");
    }
    run_tick(slot_index, mask);
    if (rows != 0) {
        check_hip(hipMemcpyAsync(slot.res_idx.data(), slot.out_idx.data(), rows * sizeof(uint64_t),
                                 hipMemcpyDeviceToHost, stream),
                  "cannot copy serving results");
        check_hip(hipMemcpyAsync(slot.res_dist.data(), slot.out_dist.data(), rows * sizeof(float),
                                 hipMemcpyDeviceToHost, stream),
                  "cannot copy serving results");
    }
    if (n_items != 0) {
        check_hip(hipMemcpyAsync(slot.res_flags.data(), slot.flags.data(), 4 * n_items * sizeof(int),
                                 hipMemcpyDeviceToHost, stream),
                  "cannot copy serving flags");
    }
    check_hip(hipEventRecord(slot.done, stream), "cannot record a HIP event");
}

// Hands results and statuses to the slot's requests (lock held).
void uhann_server::complete(Slot& slot)
{
    float ms = 0.0F;
    if (hipEventElapsedTime(&ms, slot.start, slot.done) == hipSuccess) gpu_busy_ms += ms;
    const auto now = Clock::now();
    ++ticks;
    tick_queries += slot.nq;
    for (std::size_t i = 0; i < slot.requests.size(); ++i) {
        Request& r = *slot.requests[i];
        ServedIndex& e = *r.index;
        const int* f = slot.res_flags.data() + 4 * i;
        if (e.state == UHANN_INDEX_FAILED) {
            r.status = e.fail_status;
            r.error = e.fail_message;
        } else if (r.n != 0 && slot.zero[i] == 0 && f[0] != 0) {
            r.status = UHANN_ERR_INVALID_ARGUMENT;
            r.error = e.spec.kind == UHANN_KIND_EXACT ? "query values must be finite"
                                                      : "sampled query coordinate must be finite";
        } else if (r.n != 0 && slot.zero[i] == 0 && f[1] != 0) {
            r.status = UHANN_ERR_OVERFLOW;
            r.error = "projected query is not finite";
        } else if (r.n != 0 && slot.zero[i] == 0 && f[2] != 0) {
            r.status = UHANN_ERR_OVERFLOW;
            r.error = "median L1 distance is not finite";
        } else {
            const uint32_t q0 = slot.qoff[i];
            for (uint64_t q = 0; q < r.n; ++q) {
                r.out_idx[q] = slot.zero[i] ? 0 : slot.res_idx.data()[q0 + q];
                if (r.out_dist) r.out_dist[q] = slot.zero[i] ? 0.0F : slot.res_dist.data()[q0 + q];
            }
        }
        r.done = true;
        latency_ms.push_back(std::chrono::duration<double, std::milli>(now - r.t_submit).count());
        done_queries += r.n;
        --e.inflight;
        --open_requests;
    }
    last_done = now;
    slot.requests.clear();
    slot.busy = false;
    cv_done.notify_all();
}

void uhann_server::fail_slot(Slot& slot, const std::string& message)
{
    for (Request* r : slot.requests) {
        r->status = UHANN_ERR_RUNTIME;
        r->error = message;
        r->done = true;
        --r->index->inflight;
        --open_requests;
    }
    slot.requests.clear();
    slot.busy = false;
    cv_done.notify_all();
}

void uhann_server::worker_main()
{
    (void)hipSetDevice(cfg.device);
    std::unique_lock<std::mutex> lk(mu);
    while (true) {
        // The oldest tick in flight completes first (one stream).
        Slot* oldest = nullptr;
        for (auto& s : slots) {
            if (s.busy && (oldest == nullptr || s.seq < oldest->seq)) oldest = &s;
        }
        if (oldest != nullptr) {
            const hipError_t q = hipEventQuery(oldest->done);
            if (q == hipSuccess) {
                complete(*oldest);
                continue;
            }
            if (q != hipErrorNotReady) {
                fail_slot(*oldest, std::string("serving tick failed: ") + hipGetErrorString(q));
                continue;
            }
        }
        Slot* free_slot = nullptr;
        int free_index = 0;
        for (int i = 0; i < 2; ++i) {
            if (!slots[i].busy) {
                free_slot = &slots[i];
                free_index = i;
                break;
            }
        }
        if (free_slot != nullptr && runnable()) {
            take(*free_slot);
            if (!free_slot->requests.empty()) {
                free_slot->busy = true;
                free_slot->seq = ++launch_seq;
                lk.unlock();
                std::string failure;
                try {
                    launch(*free_slot, free_index);
                } catch (const std::exception& ex) {
                    failure = ex.what();
                }
                lk.lock();
                if (!failure.empty()) {
                    (void)hipStreamSynchronize(stream);
                    fail_slot(*free_slot, failure);
                }
                continue;
            }
        }
        if (stop && oldest == nullptr && queue.empty()) break;
        if (stop && oldest == nullptr && !runnable()) {
            // Requests of indexes that never became resident cannot run.
            for (Request* r : queue) {
                r->status = UHANN_ERR_RUNTIME;
                r->error = "the server stopped before the index was admitted";
                r->done = true;
                --r->index->inflight;
                --open_requests;
            }
            queue.clear();
            cv_done.notify_all();
            continue;
        }
        if (oldest != nullptr) {
            cv_work.wait_for(lk, std::chrono::microseconds(20));
        } else {
            cv_work.wait(lk);
        }
    }
}

void uhann_server::build(ServedIndex& e)
{
    const auto& s = e.spec;
    uhann_index* ix = nullptr;
    int st;
    if (s.kind == UHANN_KIND_EXACT) {
        st = uhann_exact_create(s.metric, s.host_refs, s.rows, s.dim, cfg.device, &ix);
    } else if (s.kind == UHANN_KIND_FLAT) {
        st = uhann_flat_create(s.metric, s.host_refs, s.rows, s.dim, e.sample.data(), e.sample.size(), cfg.device,
                               &ix);
    } else {
        st = uhann_hierarchical_create(s.metric, s.host_refs, s.rows, s.dim, e.sample.data(), e.sample.size(),
                                       e.projection.data(), s.projection_dim, cfg.device, &ix);
    }
    if (st != UHANN_OK) {
        e.fail_status = st;
        e.fail_message = uhann_last_error();
        return;
    }
    e.ix = ix;
}

// Builds waiting indexes in arrival order while they fit the budget. The
// bytes are reserved before the build, which runs without the lock.
void uhann_server::admit(std::unique_lock<std::mutex>& lk)
{
    while (!admission.empty()) {
        ServedIndex* e = admission.front();
        if (used_bytes + e->estimate > cfg.hbm_budget_bytes) break;
        admission.pop_front();
        used_bytes += e->estimate;
        e->building = true;
        lk.unlock();
        build(*e);
        lk.lock();
        e->building = false;
        used_bytes -= e->estimate;
        if (e->ix != nullptr) {
            e->device_bytes = actual_bytes(*e->ix);
            used_bytes += e->device_bytes;
            e->state = UHANN_INDEX_RESIDENT;
        } else {
            e->state = UHANN_INDEX_FAILED;
        }
        e->spec.host_refs = nullptr;  // no longer borrowed
        cv_work.notify_all();
        cv_done.notify_all();
    }
}

namespace {

Request* find_request(uhann_server* s, uint64_t id)
{
    auto it = s->requests.find(id);
    if (it == s->requests.end()) throw std::invalid_argument("unknown request id");
    return it->second.get();
}

uint64_t submit_locked(uhann_server* s, uint64_t index_id, const float* queries, int on_device, uint64_t ld,
                       uint64_t count)
{
    auto it = s->indexes.find(index_id);
    if (it == s->indexes.end() || it->second->removing) throw std::invalid_argument("unknown index id");
    ServedIndex* e = it->second.get();
    if (count > s->cfg.max_tick_queries) throw std::invalid_argument("request exceeds the tick's query rows");
    if (count != 0 && queries == nullptr) throw std::invalid_argument("queries must not be null");
    if (ld < e->spec.dim) throw std::invalid_argument("query row stride is smaller than the index dimension");
    if (request_scratch(*e, count) > s->cfg.scratch_floats ||
        request_staging(*e, count, on_device != 0) > s->cfg.staging_floats) {
        throw std::invalid_argument("request exceeds the tick's scratch or staging memory");
    }
    auto r = std::make_unique<Request>();
    r->id = s->next_request++;
    r->index = e;
    r->queries = queries;
    r->on_device = on_device != 0;
    r->ld = ld;
    r->n = count;
    r->t_submit = Clock::now();
    if (!s->any_submit) {
        s->any_submit = true;
        s->first_submit = r->t_submit;
    }
    const uint64_t id = r->id;
    Request* raw = r.get();
    s->requests.emplace(id, std::move(r));
    ++e->inflight;
    ++s->open_requests;
    s->queue.push_back(raw);
    return id;
}

}  // namespace

extern "C" {

int uhann_server_create(const uhann_server_config* config, uhann_server** out)
{
    return guarded([&] {
        if (out == nullptr || config == nullptr) throw std::invalid_argument("arguments must not be null");
        *out = nullptr;
        const uhann_server_config& c = *config;
        select_device(c.device);
        if (c.max_tick_queries == 0 || c.max_tick_requests == 0) {
            throw std::invalid_argument("a tick needs at least one query row and one request");
        }
        if (!kern::serve_pairs_tq_supported(c.pairs_tq)) throw std::invalid_argument("pairs_tq must be 1, 2 or 4");
        auto s = std::make_unique<uhann_server>();
        s->cfg = c;
        check_hip(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking), "cannot create a HIP stream");
        for (auto& slot : s->slots) slot = make_slot(c);
        uhann_server* raw = s.get();
        s->worker = std::thread([raw] { raw->worker_main(); });
        *out = s.release();
    });
}

void uhann_server_destroy(uhann_server* server)
{
    if (server == nullptr) return;
    (void)uhann_server_drain(server);
    delete server;
}

int uhann_server_register(uhann_server* s, const uhann_index_spec* spec, uint64_t* index_id)
{
    return guarded([&] {
        if (s == nullptr || spec == nullptr || index_id == nullptr) throw std::invalid_argument("arguments must not be null");
        const uhann_index_spec& sp = *spec;
        if (sp.kind != UHANN_KIND_EXACT && sp.kind != UHANN_KIND_FLAT && sp.kind != UHANN_KIND_HIERARCHICAL) {
            throw std::invalid_argument("unknown index kind");
        }
        if (sp.metric != UHANN_METRIC_L2 && sp.metric != UHANN_METRIC_L1) throw std::invalid_argument("unknown metric");
        if (sp.rows == 0) throw std::invalid_argument("the index expects at least one representative");
        if (sp.rows * sp.dim != 0 && sp.host_refs == nullptr) throw std::invalid_argument("representatives must not be null");
        if (sp.strategy != UHANN_STRATEGY_DIRECT &&
            !(sp.strategy == UHANN_STRATEGY_PROJECTION_BF16 && sp.kind == UHANN_KIND_HIERARCHICAL &&
              sp.metric == UHANN_METRIC_L2)) {
            throw std::invalid_argument("serving supports DIRECT, and PROJECTION_BF16 for hierarchical L2 indexes");
        }
        auto e = std::make_unique<ServedIndex>();
        e->spec = sp;
        if (sp.kind != UHANN_KIND_EXACT) {
            if (sp.sample_count != 0 && sp.sample == nullptr) throw std::invalid_argument("sample must not be null");
            e->sample.assign(sp.sample, sp.sample + sp.sample_count);
        }
        if (sp.kind == UHANN_KIND_HIERARCHICAL) {
            if (sp.projection_dim == 0) throw std::invalid_argument("projection_dimension has to be a positive integer");
            const uint64_t n = checked_mul(sp.projection_dim, sp.sample_count, "projection");
            if (n != 0 && sp.projection == nullptr) throw std::invalid_argument("projection must not be null");
            e->projection.assign(sp.projection, sp.projection + n);
            if (sp.metric == UHANN_METRIC_L1 && sp.projection_dim > kern::median_register_dimension()) {
                throw std::invalid_argument("serving supports median-L1 projection dimensions up to 128");
            }
        }
        e->spec.sample = nullptr;
        e->spec.projection = nullptr;
        e->estimate = estimate_bytes(*e);
        std::unique_lock<std::mutex> lk(s->mu);
        if (e->estimate > s->cfg.hbm_budget_bytes) throw std::invalid_argument("index exceeds the server's HBM budget");
        e->id = s->next_index++;
        ServedIndex* raw = e.get();
        s->indexes.emplace(e->id, std::move(e));
        s->admission.push_back(raw);
        *index_id = raw->id;
        s->admit(lk);
    });
}

int uhann_server_index_state(uhann_server* s, uint64_t id, int* state, uint64_t* device_bytes)
{
    return guarded([&] {
        if (s == nullptr) throw std::invalid_argument("server must not be null");
        std::lock_guard<std::mutex> lk(s->mu);
        auto it = s->indexes.find(id);
        if (it == s->indexes.end()) throw std::invalid_argument("unknown index id");
        if (state) *state = it->second->state;
        if (device_bytes) *device_bytes = it->second->state == UHANN_INDEX_RESIDENT ? it->second->device_bytes
                                                                                    : it->second->estimate;
    });
}

int uhann_server_remove(uhann_server* s, uint64_t id)
{
    return guarded([&] {
        if (s == nullptr) throw std::invalid_argument("server must not be null");
        std::unique_lock<std::mutex> lk(s->mu);
        auto it = s->indexes.find(id);
        if (it == s->indexes.end() || it->second->removing) throw std::invalid_argument("unknown index id");
        ServedIndex* e = it->second.get();
        e->removing = true;
        s->cv_work.notify_all();
        s->cv_done.wait(lk, [&] { return !e->building && (e->inflight == 0 || e->state == UHANN_INDEX_PENDING); });
        if (e->state == UHANN_INDEX_PENDING) {
            s->admission.erase(std::find(s->admission.begin(), s->admission.end(), e));
            for (auto q = s->queue.begin(); q != s->queue.end();) {
                if ((*q)->index == e) {
                    (*q)->status = UHANN_ERR_INVALID_ARGUMENT;
                    (*q)->error = "the index was removed before it was admitted";
                    (*q)->done = true;
                    --s->open_requests;
                    q = s->queue.erase(q);
                } else {
                    ++q;
                }
            }
            e->inflight = 0;
            s->cv_done.notify_all();
        }
        uhann_index* ix = e->ix;
        const uint64_t bytes = e->state == UHANN_INDEX_RESIDENT ? e->device_bytes : 0;
        // Requests already completed keep a pointer to their index; the record
        // stays until they are waited for, so only the device state is freed.
        e->ix = nullptr;
        e->state = UHANN_INDEX_FAILED;
        e->fail_status = UHANN_ERR_INVALID_ARGUMENT;
        e->fail_message = "the index was removed";
        lk.unlock();
        uhann_index_destroy(ix);
        lk.lock();
        s->used_bytes -= bytes;
        s->admit(lk);
    });
}

int uhann_server_submit(uhann_server* s, uint64_t index_id, const float* queries, int on_device, uint64_t ld,
                        uint64_t count, uint64_t* out_idx, float* out_dist, uint64_t* request_id)
{
    return guarded([&] {
        if (s == nullptr || request_id == nullptr) throw std::invalid_argument("arguments must not be null");
        if (count != 0 && out_idx == nullptr) throw std::invalid_argument("output indices must not be null");
        std::lock_guard<std::mutex> lk(s->mu);
        const uint64_t id = submit_locked(s, index_id, queries, on_device, ld, count);
        Request* r = s->requests[id].get();
        r->out_idx = out_idx;
        r->out_dist = out_dist;
        *request_id = id;
        s->cv_work.notify_one();
    });
}

int uhann_server_wait(uhann_server* s, uint64_t request_id, double timeout_s, int* done, int* request_status)
{
    std::string message;
    int st = UHANN_OK;
    const int call = guarded([&] {
        if (s == nullptr || done == nullptr) throw std::invalid_argument("arguments must not be null");
        std::unique_lock<std::mutex> lk(s->mu);
        Request* r = find_request(s, request_id);
        auto ready = [&] { return r->done; };
        if (timeout_s < 0) {
            s->cv_done.wait(lk, ready);
        } else {
            s->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), ready);
        }
        *done = r->done ? 1 : 0;
        if (r->done) {
            st = r->status;
            message = r->error;
            if (request_status) *request_status = st;
            s->requests.erase(request_id);
        }
    });
    if (call == UHANN_OK && st != UHANN_OK) set_last_error(message);
    return call;
}

int uhann_server_run_trace(uhann_server* s, const uhann_trace_entry* entries, uint64_t count, uint64_t* ids)
{
    return guarded([&] {
        if (s == nullptr || (count != 0 && (entries == nullptr || ids == nullptr))) {
            throw std::invalid_argument("arguments must not be null");
        }
        for (uint64_t i = 1; i < count; ++i) {
            if (entries[i].arrival_s < entries[i - 1].arrival_s) throw std::invalid_argument("trace is not in arrival order");
        }
        const auto t0 = Clock::now();
        for (uint64_t i = 0; i < count; ++i) {
            const uhann_trace_entry& t = entries[i];
            const auto due = t0 + std::chrono::duration_cast<Clock::duration>(std::chrono::duration<double>(t.arrival_s));
            // Sleep most of the gap, then spin: arrivals keep ~10 us accuracy.
            while (true) {
                const auto now = Clock::now();
                if (now >= due) break;
                if (due - now > std::chrono::microseconds(200)) {
                    std::this_thread::sleep_for(due - now - std::chrono::microseconds(150));
                } else {
                    std::this_thread::yield();
                }
            }
            if (t.count != 0 && t.out_idx == nullptr) throw std::invalid_argument("output indices must not be null");
            std::lock_guard<std::mutex> lk(s->mu);
            ids[i] = submit_locked(s, t.index_id, t.queries, t.on_device, t.ld_queries, t.count);
            Request* r = s->requests[ids[i]].get();
            r->out_idx = t.out_idx;
            r->out_dist = t.out_dist;
            s->cv_work.notify_one();
        }
    });
}

int uhann_server_drain(uhann_server* s)
{
    return guarded([&] {
        if (s == nullptr) throw std::invalid_argument("server must not be null");
        std::unique_lock<std::mutex> lk(s->mu);
        s->cv_work.notify_all();
        s->cv_done.wait(lk, [&] {
            if (s->open_requests == 0) return true;
            // Requests of indexes still waiting for memory cannot finish.
            for (const Request* r : s->queue) {
                if (r->index->state != UHANN_INDEX_PENDING) return false;
            }
            for (const auto& slot : s->slots) {
                if (slot.busy) return false;
            }
            return true;
        });
    });
}

int uhann_server_stats(uhann_server* s, uhann_serving_stats* out)
{
    return guarded([&] {
        if (s == nullptr || out == nullptr) throw std::invalid_argument("arguments must not be null");
        std::lock_guard<std::mutex> lk(s->mu);
        uhann_serving_stats st{};
        st.requests = s->latency_ms.size();
        st.queries = s->done_queries;
        st.ticks = s->ticks;
        if (!s->latency_ms.empty()) {
            std::vector<double> v = s->latency_ms;
            std::sort(v.begin(), v.end());
            auto pct = [&](double q) {
                const std::size_t i = static_cast<std::size_t>(q * static_cast<double>(v.size() - 1) + 0.5);
                return v[std::min(i, v.size() - 1)];
            };
            st.p50_ms = pct(0.50);
            st.p99_ms = pct(0.99);
            double sum = 0.0;
            for (double x : v) sum += x;
            st.mean_ms = sum / static_cast<double>(v.size());
            st.max_ms = v.back();
            st.wall_s = std::chrono::duration<double>(s->last_done - s->first_submit).count();
            st.queries_per_s = st.wall_s > 0 ? static_cast<double>(st.queries) / st.wall_s : 0.0;
        }
        st.gpu_busy_s = s->gpu_busy_ms * 1e-3;
        st.mean_tick_queries = s->ticks ? static_cast<double>(s->tick_queries) / static_cast<double>(s->ticks) : 0.0;
        for (const auto& [id, e] : s->indexes) {
            if (e->state == UHANN_INDEX_RESIDENT) ++st.resident_indexes;
            if (e->state == UHANN_INDEX_PENDING) ++st.pending_indexes;
        }
        st.resident_bytes = s->used_bytes;
        *out = st;
    });
}

int uhann_server_reset_stats(uhann_server* s)
{
    return guarded([&] {
        if (s == nullptr) throw std::invalid_argument("server must not be null");
        std::lock_guard<std::mutex> lk(s->mu);
        s->latency_ms.clear();
        s->done_queries = 0;
        s->ticks = 0;
        s->tick_queries = 0;
        s->gpu_busy_ms = 0.0;
        s->any_submit = false;
    });
}

int uhann_server_autotune(uhann_server* s, const float* device_queries, uint64_t ld, uint32_t rows_per_index,
                          int* pairs_tq)
{
    return guarded([&] {
        if (s == nullptr || device_queries == nullptr) throw std::invalid_argument("arguments must not be null");
        std::lock_guard<std::mutex> lk(s->mu);
        if (s->open_requests != 0) throw std::logic_error("autotune needs an idle server");
        select_device(s->cfg.device);
        // A synthetic tick: rows_per_index queries for every resident index.
        std::vector<std::unique_ptr<Request>> fake;
        std::vector<uint64_t> sink_idx;
        Slot& slot = s->slots[0];
        for (auto& [id, e] : s->indexes) {
            if (e->state != UHANN_INDEX_RESIDENT) continue;
            if (ld < e->spec.dim) throw std::invalid_argument("query row stride is smaller than an index dimension");
            auto r = std::make_unique<Request>();
            r->index = e.get();
            r->queries = device_queries;
            r->on_device = true;
            r->ld = ld;
            r->n = rows_per_index;
            fake.push_back(std::move(r));
        }
        if (fake.empty()) throw std::logic_error("autotune needs a resident index");
        int best = s->cfg.pairs_tq;
        float best_ms = -1.0F;
        hipEvent_t a = nullptr, b = nullptr;
        check_hip(hipEventCreate(&a), "cannot create a HIP event");
        check_hip(hipEventCreate(&b), "cannot create a HIP event");
        const int saved_graph = s->cfg.use_graph;
        s->cfg.use_graph = 0;
        try {
            for (int tq : {1, 2, 4}) {
                s->cfg.pairs_tq = tq;
                float total = 0.0F;
                for (int rep = 0; rep < 4; ++rep) {
                    // Pack as many of the synthetic requests as one tick holds.
                    slot.requests.clear();
                    uint64_t rows = 0, scratch = 0;
                    for (auto& r : fake) {
                        const uint64_t need = request_scratch(*r->index, r->n);
                        if (slot.requests.size() == s->cfg.max_tick_requests || rows + r->n > s->cfg.max_tick_queries ||
                            scratch + need > s->cfg.scratch_floats) {
                            break;
                        }
                        rows += r->n;
                        scratch += need;
                        slot.requests.push_back(r.get());
                    }
                    check_hip(hipEventRecord(a, s->stream), "cannot record a HIP event");
                    s->launch(slot, 0);
                    check_hip(hipEventRecord(b, s->stream), "cannot record a HIP event");
                    check_hip(hipEventSynchronize(b), "serving autotune tick failed");
                    float ms = 0.0F;
                    check_hip(hipEventElapsedTime(&ms, a, b), "cannot read a HIP event");
                    if (rep > 0) total += ms;  // first run warms up
                }
                if (best_ms < 0.0F || total < best_ms) {
                    best_ms = total;
                    best = tq;
                }
            }
        } catch (...) {
            slot.requests.clear();
            s->cfg.pairs_tq = best;
            s->cfg.use_graph = saved_graph;
            (void)hipEventDestroy(a);
            (void)hipEventDestroy(b);
            throw;
        }
        slot.requests.clear();
        s->cfg.pairs_tq = best;
        s->cfg.use_graph = saved_graph;
        for (auto& [key, g] : s->graphs) (void)hipGraphExecDestroy(g);
        s->graphs.clear();
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(b);
        if (pairs_tq) *pairs_tq = best;
    });
}

}  // extern "C"
