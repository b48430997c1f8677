// Helpers shared by the C++ drop-in adapters: the translation units in this
// directory implement the reference's device classes (declared in the
// reference's own headers) on top of the C ABI in include/uhann.h, so the
// reference's facades, benchmark and tests run on this engine unchanged.
// Built inside the reference's tree in place of its CUDA translation units
// (the CMake lines are in INTEGRATION.md); nothing here includes HIP.
#pragma once

#include <uhann.h>

#include <cstddef>
#include <cstdint>
#include <new>
#include <stdexcept>
#include <string>
#include <type_traits>

namespace ultrahigh_ann::uhann_adapter {

static_assert(sizeof(std::size_t) == sizeof(uint64_t) && alignof(std::size_t) == alignof(uint64_t),
              "the adapters pass std::size_t spans as uint64_t arrays (LP64)");

// Rethrows a failed call with the exception type the reference throws for
// that condition (the status codes map one to one, see uhann.h).
inline void check(int status)
{
    if (status == UHANN_OK) return;
    const char* m = uhann_last_error();
    const std::string msg = m != nullptr && m[0] != '\0' ? m : "libuhann call failed";
    switch (status) {
    case UHANN_ERR_INVALID_ARGUMENT: throw std::invalid_argument(msg);
    case UHANN_ERR_LENGTH: throw std::length_error(msg);
    case UHANN_ERR_OVERFLOW: throw std::overflow_error(msg);
    case UHANN_ERR_DOMAIN: throw std::domain_error(msg);
    case UHANN_ERR_OUT_OF_RANGE: throw std::out_of_range(msg);
    case UHANN_ERR_LOGIC: throw std::logic_error(msg);
    default: throw std::runtime_error(msg);
    }
}

inline uint64_t* as_u64(std::size_t* p) { return reinterpret_cast<uint64_t*>(p); }

inline std::string device_name(int device)
{
    char buf[256] = {};
    check(uhann_device_name(device, buf, sizeof(buf)));
    return std::string(buf);
}

// a * b for payload sizes, with the reference's overflow behaviour.
inline std::size_t product(std::size_t a, std::size_t b, const char* what)
{
    if (a != 0 && b > static_cast<std::size_t>(-1) / a) throw std::length_error(std::string(what) + " size overflows");
    return a * b;
}

}  // namespace ultrahigh_ann::uhann_adapter
