// Result codes, taints and checking macros of the Blade C++ API
// (reference include/blade/macros.hh:11-154; same names and values).
#ifndef BLADE_MACROS_HH
#define BLADE_MACROS_HH

#include <stdint.h>

#include "blade/logger.hh"

namespace Blade {

enum class Taint : uint8_t {
    NONE     = 0,
    CONSUMER = 1 << 0,   // consumes externally allocated input data
    PRODUCER = 1 << 1,   // produces internally allocated output data
    MODIFIER = 1 << 2,   // modifies externally allocated input data
    CHRONOUS = 1 << 3,   // has a compute ratio other than one
};

inline constexpr Taint operator|(Taint lhs, Taint rhs) {
    return static_cast<Taint>(static_cast<uint8_t>(lhs) | static_cast<uint8_t>(rhs));
}

inline constexpr Taint operator&(Taint lhs, Taint rhs) {
    return static_cast<Taint>(static_cast<uint8_t>(lhs) & static_cast<uint8_t>(rhs));
}

// Hard errors are below 16; soft results have bit 4 set.
enum class Result : uint8_t {
    SUCCESS = 0,
    ERROR = 1,
    CUDA_ERROR,          // a HIP runtime/launch error (the reference's name)
    ASSERTION_ERROR,
    RUNNER_QUEUE_FULL           = 0 | (1 << 4),
    RUNNER_QUEUE_EMPTY          = 1 | (1 << 4),
    RUNNER_QUEUE_NONE_AVAILABLE = 2 | (1 << 4),
    PIPELINE_EXHAUSTED          = 3 | (1 << 4),
};

inline constexpr bool IsSoft(Result r) { return (static_cast<uint8_t>(r) & (1 << 4)) != 0; }

}  // namespace Blade

// Return the result of `x` if it is a hard error (soft results pass).
#define BL_CHECK(x)                                                        \
    do {                                                                   \
        const ::Blade::Result _r = (x);                                    \
        if (_r != ::Blade::Result::SUCCESS && !::Blade::IsSoft(_r)) {      \
            return _r;                                                     \
        }                                                                  \
    } while (0)

// Throw the result of `x` if it is not SUCCESS (constructor validation).
#define BL_CHECK_THROW(x)                                                  \
    do {                                                                   \
        const ::Blade::Result _r = (x);                                    \
        if (_r != ::Blade::Result::SUCCESS) {                              \
            throw _r;                                                      \
        }                                                                  \
    } while (0)

// Map a C-ABI return code (include/blade_amd.h) to a Result, logging the
// library's message on failure.
#define BL_ABI_CHECK(call)                                                 \
    do {                                                                   \
        const int _rc = (call);                                            \
        if (_rc != 0) {                                                    \
            BL_FATAL("{}", blade_amd_last_error());                        \
            return static_cast<::Blade::Result>(_rc);                      \
        }                                                                  \
    } while (0)

#define BL_HIP_CHECK(call, ...)                                            \
    do {                                                                   \
        const hipError_t _e = (call);                                      \
        if (_e != hipSuccess) {                                            \
            BL_FATAL("HIP error: {}", hipGetErrorString(_e));              \
            return ::Blade::Result::CUDA_ERROR;                            \
        }                                                                  \
    } while (0)

#define BL_HIP_CHECK_THROW(call)                                           \
    do {                                                                   \
        const hipError_t _e = (call);                                      \
        if (_e != hipSuccess) {                                            \
            BL_FATAL("HIP error: {}", hipGetErrorString(_e));              \
            throw ::Blade::Result::CUDA_ERROR;                             \
        }                                                                  \
    } while (0)

#endif
