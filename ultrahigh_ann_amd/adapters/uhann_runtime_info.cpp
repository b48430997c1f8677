// cuda_runtime_info (include/ultrahigh_ann/core/cuda_runtime_info.hpp) on
// this engine: replaces src/core/cuda_runtime_info.cu. The "compute
// capability" fields carry the GPU architecture's major and minor numbers
// (gfx950: 9 and 5) and the version fields the HIP versions.

#include "ultrahigh_ann/core/cuda_runtime_info.hpp"

#include <uhann.h>

namespace ultrahigh_ann {

std::optional<CudaRuntimeInfo> cuda_runtime_info(int device) noexcept
{
    try {
        if (uhann_device_available() == 0) return std::nullopt;
        char name[256] = {};
        CudaRuntimeInfo info{};
        uint64_t total = 0;
        if (uhann_device_properties(device, name, sizeof(name), &info.compute_capability_major,
                                    &info.compute_capability_minor, &total, &info.compiled_runtime_version,
                                    &info.runtime_version, &info.driver_version) != UHANN_OK) {
            return std::nullopt;
        }
        info.device = device;
        info.device_name = name;
        info.total_global_memory_bytes = static_cast<std::size_t>(total);
        return info;
    } catch (...) {
        return std::nullopt;
    }
}

}  // namespace ultrahigh_ann
