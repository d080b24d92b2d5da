// Scalar types, devices and streams of the Blade C++ API
// (reference include/blade/memory/types.hh; same names).
#ifndef BLADE_MEMORY_TYPES_HH
#define BLADE_MEMORY_TYPES_HH

#include <hip/hip_runtime_api.h>

#include <cmath>
#include <complex>
#include <cstddef>
#include <cstdint>

#include "blade_amd.h"

#ifndef BL_PHYSICAL_CONSTANT_C
#define BL_PHYSICAL_CONSTANT_C (double)299792458.0   // speed of light (m/s)
#endif

#ifndef BL_PHYSICAL_CONSTANT_PI
#define BL_PHYSICAL_CONSTANT_PI M_PI
#endif

#define BLADE_API __attribute__((visibility("default")))
#define BLADE_HIDDEN __attribute__((visibility("hidden")))

namespace Blade {

// A HIP stream; the default-constructed Stream is the null stream.
struct Stream {
    hipStream_t handle = nullptr;

    operator hipStream_t() const { return handle; }
    operator hipStream_t*() { return &handle; }
    blade_stream_t abi() const { return reinterpret_cast<blade_stream_t>(handle); }
};

typedef _Float16 F16;
typedef float    F32;
typedef double   F64;
typedef int8_t   I8;
typedef int16_t  I16;
typedef int32_t  I32;
typedef int64_t  I64;
typedef uint8_t  U8;
typedef uint16_t U16;
typedef uint32_t U32;
typedef uint64_t U64;
typedef bool     BOOL;

typedef std::complex<F16> CF16;
typedef std::complex<F32> CF32;
typedef std::complex<F64> CF64;
typedef std::complex<I8>  CI8;
typedef std::complex<I32> CI32;

// Memory spaces.  CUDA is the reference's name for accelerator memory and is
// kept as the enumerator name of HIP device memory so that user code
// spelling Device::CUDA keeps compiling.
enum class Device : uint8_t {
    NONE = 0,
    CPU  = 1 << 0,
    CUDA = 1 << 1,
};

template <Device D = Device::NONE>
struct DeviceInfo;
template <>
struct DeviceInfo<Device::CUDA> { inline static const char* name = "HIP"; };
template <>
struct DeviceInfo<Device::CPU> { inline static const char* name = "CPU"; };

template <typename T = void>
struct TypeInfo;

#define BLADE_TYPEINFO(T, SUB, CPLX, NAME, CODE)                      \
    template <>                                                        \
    struct TypeInfo<T> {                                               \
        using type = T;                                                \
        using subtype = SUB;                                           \
        inline static const bool is_complex = CPLX;                    \
        inline static const char* name = NAME;                         \
        inline static const int abi = CODE;                            \
        inline static const U64 components = CPLX ? 2 : 1;             \
    };

BLADE_TYPEINFO(U8, U8, false, "U8", BLADE_I8)
BLADE_TYPEINFO(I8, I8, false, "I8", BLADE_I8)
BLADE_TYPEINFO(F16, F16, false, "F16", BLADE_F16)
BLADE_TYPEINFO(F32, F32, false, "F32", BLADE_F32)
BLADE_TYPEINFO(F64, F64, false, "F64", BLADE_F64)
BLADE_TYPEINFO(I32, I32, false, "I32", BLADE_I32)
BLADE_TYPEINFO(CI8, I8, true, "CI8", BLADE_CI8)
BLADE_TYPEINFO(CF16, F16, true, "CF16", BLADE_CF16)
BLADE_TYPEINFO(CF32, F32, true, "CF32", BLADE_CF32)
BLADE_TYPEINFO(CF64, F64, true, "CF64", BLADE_CF64)

#undef BLADE_TYPEINFO

}  // namespace Blade

#endif
