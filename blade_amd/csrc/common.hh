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

__device__ __forceinline__ floatx4 mfma_f16(halfx8 a, halfx8 b, floatx4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
}

// Exact CI8 -> FP16: int8 b becomes the biased byte u = b + 128, placed in the
// mantissa of 1024.0 (0x6400) to give 1024 + u exactly; subtracting 1152 gives b.
// Both pols of a dword at once: v_perm_b32 places each biased byte below a
// 0x64 byte (0x64uu = 1024 + u in FP16), one packed subtract removes the bias.
// {src0, src1} is read as bytes 7..4 = src0, 3..0 = src1.
__device__ __forceinline__ void ci8_to_half2x2(uint32_t d, halfx2& p0, halfx2& p1) {
    const uint32_t u = d ^ 0x80808080u;
    const uint32_t b0 = __builtin_amdgcn_perm(u, 0x64646464u, 0x00050004u);   // (u0, 64, u1, 64)
    const uint32_t b1 = __builtin_amdgcn_perm(u, 0x64646464u, 0x00070006u);   // (u2, 64, u3, 64)
    const halfx2 bias = {(_Float16)1152.0f, (_Float16)1152.0f};
    p0 = __builtin_bit_cast(halfx2, b0) - bias;
    p1 = __builtin_bit_cast(halfx2, b1) - bias;
}

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
