// Detector: integrate |X|^2+|Y|^2 (Stokes-I) or [XX*, YY*, Re XY*, Im XY*]
// over `integration` consecutive samples.  No atomics: every output sample is
// owned by one lane (short integrations) or one lane group that reduces with
// cross-lane shuffles (long integrations), so the result is deterministic and
// there is no zero-then-add race (the reference zeroes with a block-local
// barrier before cross-block atomicAdds, detector.cu:11-14).
#include "common.hh"

namespace bl {

struct Pow4 {
    float xx, yy, re, im;
};

__device__ __forceinline__ void accum(Pow4& a, float4 v) {
    // v = (Xr, Xi, Yr, Yi)
    a.xx += v.x * v.x + v.y * v.y;
    a.yy += v.z * v.z + v.w * v.w;
    // X * conj(Y) = (Xr Yr + Xi Yi) + j (Xi Yr - Xr Yi)
    a.re += v.x * v.z + v.y * v.w;
    a.im += v.y * v.z - v.x * v.w;
}

template <int NPOL>
__device__ __forceinline__ void store_out(float* out, uint64_t o, const Pow4& a, bool acc) {
    if constexpr (NPOL == 1) {
        float v = a.xx + a.yy;
        if (acc) v += out[o];
        out[o] = v;
    } else {
        float4* o4 = reinterpret_cast<float4*>(out) + o;
        float4 v = make_float4(a.xx, a.yy, a.re, a.im);
        if (acc) {
            const float4 p = *o4;
            v.x += p.x; v.y += p.y; v.z += p.z; v.w += p.w;
        }
        *o4 = v;
    }
}

// One lane per output sample (integrations of 1 and of odd or other short
// lengths).
template <int NPOL>
__global__ __launch_bounds__(256) void detector_lane_kernel(const float4* __restrict__ in,
                                                            float* __restrict__ out,
                                                            uint64_t n_out, uint64_t intg,
                                                            int accumulate) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t o = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; o < n_out; o += stride) {
        Pow4 a = {0.f, 0.f, 0.f, 0.f};
        const float4* p = in + o * intg;
        for (uint64_t i = 0; i < intg; i++) accum(a, p[i]);
        store_out<NPOL>(out, o, a, accumulate != 0);
    }
}

// G lanes per output sample; lanes read consecutive samples, then a
// butterfly reduction over the G-lane group.
template <int NPOL, int G>
__global__ __launch_bounds__(256) void detector_group_kernel(const float4* __restrict__ in,
                                                             float* __restrict__ out,
                                                             uint64_t n_out, uint64_t intg,
                                                             int accumulate) {
    const int lane_g = threadIdx.x % G;
    const uint64_t groups_per_grid = (uint64_t)gridDim.x * (blockDim.x / G);
    for (uint64_t o = (uint64_t)blockIdx.x * (blockDim.x / G) + threadIdx.x / G; o < n_out;
         o += groups_per_grid) {
        Pow4 a = {0.f, 0.f, 0.f, 0.f};
        const float4* p = in + o * intg;
        for (uint64_t i = lane_g; i < intg; i += G) accum(a, p[i]);
#pragma unroll
        for (int m = G / 2; m >= 1; m /= 2) {
            a.xx += __shfl_xor(a.xx, m, G);
            a.yy += __shfl_xor(a.yy, m, G);
            if constexpr (NPOL == 4) {
                a.re += __shfl_xor(a.re, m, G);
                a.im += __shfl_xor(a.im, m, G);
            }
        }
        if (lane_g == 0) store_out<NPOL>(out, o, a, accumulate != 0);
    }
}

static unsigned grid_cap(uint64_t blocks) {
    const uint64_t cap = 256ull * 32;
    if (blocks > cap) blocks = cap;
    return blocks ? (unsigned)blocks : 1u;
}

template <int NPOL>
static void launch(const float4* in, float* out, uint64_t n_out, uint64_t intg, int acc,
                   hipStream_t s) {
    // 64 lanes per output only where outputs are too few to fill the chip
    // with 16: a 64-lane reduction costs four times the shuffles per byte
    // (at INTG 64 with 4 pols, 3.7 against 5.7 TB/s).
    if (intg >= 64 && n_out < 65536) {
        hipLaunchKernelGGL((detector_group_kernel<NPOL, 64>), dim3(grid_cap(ceil_div(n_out, 4))),
                           dim3(256), 0, s, in, out, n_out, intg, acc);
    } else if (intg >= 16) {
        hipLaunchKernelGGL((detector_group_kernel<NPOL, 16>), dim3(grid_cap(ceil_div(n_out, 16))),
                           dim3(256), 0, s, in, out, n_out, intg, acc);
    } else if (intg == 8 || intg == 4 || intg == 2) {
        // One lane per input sample: every load instruction reads 64
        // consecutive samples; the group of `intg` lanes reduces by shuffles.
        const unsigned g = grid_cap(ceil_div(n_out * intg, 256));
        if (intg == 8)
            hipLaunchKernelGGL((detector_group_kernel<NPOL, 8>), dim3(g), dim3(256), 0, s, in, out, n_out, intg, acc);
        else if (intg == 4)
            hipLaunchKernelGGL((detector_group_kernel<NPOL, 4>), dim3(g), dim3(256), 0, s, in, out, n_out, intg, acc);
        else
            hipLaunchKernelGGL((detector_group_kernel<NPOL, 2>), dim3(g), dim3(256), 0, s, in, out, n_out, intg, acc);
    } else {
        hipLaunchKernelGGL((detector_lane_kernel<NPOL>), dim3(grid_cap(ceil_div(n_out, 256))),
                           dim3(256), 0, s, in, out, n_out, intg, acc);
    }
}

}  // namespace bl

using namespace bl;

extern "C" int blade_detector(const void* input, void* output, uint64_t X, uint64_t F,
                              uint64_t T, uint64_t integration, uint64_t out_pols,
                              int accumulate, blade_stream_t stream) {
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    // Validation mirrors src/modules/detector/base.cc:17-51.
    if (integration == 0) return fail("detector: integration size should be more than zero");
    if (T % integration) return fail("detector: number of time samples should be divisible by the integration size");
    if (X == 0) return fail("detector: number of aspects should be more than zero");
    if (out_pols != 1 && out_pols != 4) return fail("detector: number of output polarizations not supported");
    const uint64_t n_out = X * F * (T / integration);
    if (n_out == 0) return BLADE_SUCCESS;
    if (!input || !output) return fail("detector: null pointer");
    if ((uintptr_t)input % 16 || (uintptr_t)output % (out_pols == 4 ? 16 : 4))
        return fail("detector: misaligned buffer");
    const float4* in = static_cast<const float4*>(input);
    float* out = static_cast<float*>(output);
    if (out_pols == 1) launch<1>(in, out, n_out, integration, accumulate, s);
    else launch<4>(in, out, n_out, integration, accumulate, s);
    return check_launch("detector");
}
