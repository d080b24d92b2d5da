// Shared device/host helpers for the gfx950 kernels of the ATA beamforming path.
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>
#include <stdint.h>
#include <string>

#include "blade_amd.h"

namespace bl {

// Vector types used by the MFMA builtins.
typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef float floatx16 __attribute__((ext_vector_type(16)));
typedef _Float16 halfx8 __attribute__((ext_vector_type(8)));
typedef _Float16 halfx2 __attribute__((ext_vector_type(2)));

constexpr int kWave = 64;

// Error plumbing: the C-ABI returns a Result code and keeps a thread-local
// message for blade_amd_last_error().
void set_error(const std::string& msg);

inline int fail(const std::string& msg) {
    set_error(msg);
    return BLADE_ERROR;
}

inline int check_launch(const char* what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        set_error(std::string(what) + ": " + hipGetErrorString(e));
        return BLADE_HIP_ERROR;
    }
    return BLADE_SUCCESS;
}

inline uint64_t ceil_div(uint64_t a, uint64_t b) { return (a + b - 1) / b; }

// Compute units of the current device (cached per device).
inline int device_cus() {
    static int cached[64] = {0};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
    if (!cached[dev]) {
        int n = 0;
        if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n <= 0) n = 256;
        cached[dev] = n;
    }
    return cached[dev];
}

// Library-owned scratch for launches that need it, one buffer per (device,
// stream, consumer) so that calls on different streams never share one;
// grown on demand (the old buffer is released after the stream drains).
enum WorkspaceUser { WS_BEAMFORMER_SINK = 0, WS_ADJOINT = 1, WS_CHANNELIZER_BIG = 2 };
int library_workspace(int user, size_t bytes, hipStream_t s, unsigned char** out);

// Number of scalar components and bytes per element of a blade_type.
inline int type_components(int t) {
    switch (t) {
        case BLADE_CI8: case BLADE_CF16: case BLADE_CF32: case BLADE_CF64: return 2;
        default: return 1;
    }
}

inline int type_scalar_bytes(int t) {
    switch (t) {
        case BLADE_I8: case BLADE_CI8: return 1;
        case BLADE_F16: case BLADE_CF16: return 2;
        case BLADE_F32: case BLADE_CF32: case BLADE_I32: return 4;
        case BLADE_F64: case BLADE_CF64: return 8;
        default: return 0;
    }
}

inline int type_bytes(int t) { return type_components(t) * type_scalar_bytes(t); }

__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

__device__ __forceinline__ float2 cadd(float2 a, float2 b) {
    return make_float2(a.x + b.x, a.y + b.y);
}

__device__ __forceinline__ float2 csub(float2 a, float2 b) {
    return make_float2(a.x - b.x, a.y - b.y);
}

}  // namespace bl
