// The index handle behind uhann_index (include/uhann.h): exact, flat or
// hierarchical, with the device state the query kernels read. Internal to the
// library; index.cpp builds and queries it, server.cpp schedules it.
#pragma once

#include <cstdint>
#include <memory>

#include "common.hpp"
#include "engine.hpp"

struct uhann_index {
    int kind = UHANN_KIND_EXACT;
    int metric = UHANN_METRIC_L2;
    int device = 0;
    uint64_t rows = 0;
    uint64_t dim = 0;
    uint64_t k = 0;             // sampled coordinates |I|
    uint64_t k_ld = 0;          // padded row length of gathered queries
    uint64_t multiplicity = 0;  // sum of multiplicities
    uint64_t p = 0;             // projection dimension
    uint64_t p_ld = 0;
    std::unique_ptr<uhann::Scan> scan;  // over dim (exact), k (flat), p (hierarchical L2)
    uhann::DeviceBuffer<uint32_t> cols;
    uhann::DeviceBuffer<float> scales;  // flat L2 importance scales
    uhann::DeviceBuffer<float> proj;    // p x k_ld binary32 projection
    uhann::DeviceBuffer<uint16_t> proj16;  // hierarchical L2: bf16 projection for PROJECTION_BF16
    uint64_t proj16_ld = 0;
    uhann::DeviceBuffer<float> refs_t;  // hierarchical L1: p x rows projected references
    uint64_t payload_bytes = 0;
};
