// Host-side helpers shared by the engine's translation units: HIP error
// checks, overflow-checked sizes, RAII device buffers and the mapping from
// C++ exception types to the C ABI status codes of include/uhann.h.
#pragma once

#include <hip/hip_runtime_api.h>

#include <cstddef>
#include <cstdint>
#include <limits>
#include <stdexcept>
#include <string>
#include <string_view>
#include <utility>

#include "uhann.h"

namespace uhann {

inline void check_hip(hipError_t status, std::string_view what)
{
    if (status != hipSuccess) {
        throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(status));
    }
}

inline std::size_t checked_mul(std::size_t a, std::size_t b, std::string_view what)
{
    if (b != 0 && a > std::numeric_limits<std::size_t>::max() / b) {
        throw std::length_error(std::string(what) + " size overflows");
    }
    return a * b;
}

inline std::size_t checked_add(std::size_t a, std::size_t b, std::string_view what)
{
    if (a > std::numeric_limits<std::size_t>::max() - b) {
        throw std::length_error(std::string(what) + " size overflows");
    }
    return a + b;
}

// Selects a device for the lifetime of a call and validates its index the
// way the reference's SelectedDevice does (cuda_dense_l2_scan_index.cu:94-123).
inline void select_device(int device)
{
    if (device < 0) {
        throw std::invalid_argument("HIP device index must be nonnegative");
    }
    int count = 0;
    check_hip(hipGetDeviceCount(&count), "cannot enumerate HIP devices");
    if (device >= count) {
        throw std::invalid_argument("HIP device index is out of range");
    }
    check_hip(hipSetDevice(device), "cannot select HIP device");
}

// The device's marketing name, or "AMD GPU <architecture>" when the
// runtime has no name for it (a missing amdgpu.ids table leaves it empty).
inline std::string device_label(const hipDeviceProp_t& prop)
{
    if (prop.name[0] != '\0') return std::string(prop.name);
    std::string arch(prop.gcnArchName);
    const std::size_t colon = arch.find(':');
    if (colon != std::string::npos) arch.resize(colon);
    return "AMD GPU " + (arch.empty() ? std::string("(unnamed)") : arch);
}

template <class T> class DeviceBuffer {
  public:
    DeviceBuffer() = default;
    DeviceBuffer(int device, std::size_t count)
        : device_(device), bytes_(checked_mul(count, sizeof(T), "HIP device buffer"))
    {
        if (bytes_ == 0) return;
        check_hip(hipSetDevice(device_), "cannot select HIP device");
        check_hip(hipMalloc(reinterpret_cast<void**>(&data_), bytes_),
                  "cannot allocate HIP device memory");
    }
    DeviceBuffer(const DeviceBuffer&) = delete;
    DeviceBuffer& operator=(const DeviceBuffer&) = delete;
    DeviceBuffer(DeviceBuffer&& o) noexcept
        : device_(o.device_), bytes_(std::exchange(o.bytes_, 0)),
          data_(std::exchange(o.data_, nullptr))
    {
    }
    DeviceBuffer& operator=(DeviceBuffer&& o) noexcept
    {
        if (this != &o) {
            release();
            device_ = o.device_;
            bytes_ = std::exchange(o.bytes_, 0);
            data_ = std::exchange(o.data_, nullptr);
        }
        return *this;
    }
    ~DeviceBuffer() { release(); }

    T* data() noexcept { return data_; }
    const T* data() const noexcept { return data_; }
    std::size_t bytes() const noexcept { return bytes_; }

  private:
    void release() noexcept
    {
        if (data_ != nullptr) {
            static_cast<void>(hipSetDevice(device_));
            static_cast<void>(hipFree(data_));
            data_ = nullptr;
        }
    }
    int device_ = 0;
    std::size_t bytes_ = 0;
    T* data_ = nullptr;
};

// Thread-local last error message for uhann_last_error().
void set_last_error(std::string message) noexcept;

// Runs body(), translating exceptions into status codes. The mapping keeps
// the reference's exception taxonomy so the C++ facade can rethrow the same
// types (invalid_argument, length_error, runtime_error, overflow_error,
// domain_error, out_of_range).
template <class F> int guarded(F&& body) noexcept
{
    try {
        body();
        set_last_error({});
        return UHANN_OK;
    } catch (const std::invalid_argument& e) {
        set_last_error(e.what());
        return UHANN_ERR_INVALID_ARGUMENT;
    } catch (const std::length_error& e) {
        set_last_error(e.what());
        return UHANN_ERR_LENGTH;
    } catch (const std::domain_error& e) {
        set_last_error(e.what());
        return UHANN_ERR_DOMAIN;
    } catch (const std::out_of_range& e) {
        set_last_error(e.what());
        return UHANN_ERR_OUT_OF_RANGE;
    } catch (const std::overflow_error& e) {
        set_last_error(e.what());
        return UHANN_ERR_OVERFLOW;
    } catch (const std::logic_error& e) {
        set_last_error(e.what());
        return UHANN_ERR_LOGIC;
    } catch (const std::bad_alloc&) {
        set_last_error("host allocation failed");
        return UHANN_ERR_RUNTIME;
    } catch (const std::exception& e) {
        set_last_error(e.what());
        return UHANN_ERR_RUNTIME;
    } catch (...) {
        set_last_error("unknown error");
        return UHANN_ERR_RUNTIME;
    }
}

}  // namespace uhann
