// ATA phasors from delays, on the GPU, with an FP64 phase argument.
//
// w[b,a,f,p] = exp(-j 2pi tau[b,a] ((f + fStart) dF + f_obs - BW/2)) * cal[a,f,0,p]
// (reference host loop src/modules/phasor/ata.cc:175-208).
//
// GHz frequencies times ns delays give phases of hundreds of turns, which
// FP32 cannot carry to 1e-7.  The phase is therefore formed in FP64 as a
// number of turns, reduced exactly to [-1/2, 1/2] by subtracting the nearest
// integer (exact in binary floating point), and only then passed to
// sincospi (ROCm device math), so no multiplication by a rounded pi
// precedes the reduction.
#include "common.hh"

namespace bl {

template <typename OT>
__global__ __launch_bounds__(256) void phasor_kernel(const double* __restrict__ delays,
                                                     const double2* __restrict__ cal,
                                                     OT* __restrict__ out, uint64_t B,
                                                     uint64_t A, uint64_t F, uint64_t P,
                                                     double dF, double fringe, double fStart) {
    const uint64_t n = B * A * F;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const uint64_t f = i % F;
        const uint64_t ba = i / F;            // b * A + a
        const uint64_t a = ba % A;
        const double tau = delays[ba];
        const double freq = ((double)f + fStart) * dF + fringe;
        const double turns = -tau * freq;
        const double r = turns - rint(turns);   // exact
        double sn, cs;
        sincospi(2.0 * r, &sn, &cs);            // 2*r is exact
        for (uint64_t p = 0; p < P; p++) {
            const double2 c = cal[(a * F + f) * P + p];
            const double re = cs * c.x - sn * c.y;
            const double im = cs * c.y + sn * c.x;
            if constexpr (sizeof(OT) == 8) {
                out[i * P + p] = make_float2((float)re, (float)im);
            } else {
                out[i * P + p] = make_double2(re, im);
            }
        }
    }
}

}  // namespace bl

using namespace bl;

extern "C" int blade_phasor_ata(const void* delays, const void* calibration, void* phasors,
                                int out_type, uint64_t B, uint64_t A, uint64_t F, uint64_t P,
                                double observation_frequency_hz, double channel_bandwidth_hz,
                                double total_bandwidth_hz, uint64_t frequency_start_index,
                                blade_stream_t stream) {
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    const uint64_t n = B * A * F;
    if (n == 0 || P == 0) return BLADE_SUCCESS;
    if (!delays || !calibration || !phasors) return fail("phasor: null pointer");
    const double fringe = observation_frequency_hz - total_bandwidth_hz / 2.0;
    uint64_t blocks = ceil_div(n, 256);
    if (blocks > 8192) blocks = 8192;
    if (out_type == BLADE_CF32) {
        hipLaunchKernelGGL((phasor_kernel<float2>), dim3((unsigned)blocks), dim3(256), 0, s,
                           static_cast<const double*>(delays), static_cast<const double2*>(calibration),
                           static_cast<float2*>(phasors), B, A, F, P, channel_bandwidth_hz, fringe,
                           (double)frequency_start_index);
    } else if (out_type == BLADE_CF64) {
        hipLaunchKernelGGL((phasor_kernel<double2>), dim3((unsigned)blocks), dim3(256), 0, s,
                           static_cast<const double*>(delays), static_cast<const double2*>(calibration),
                           static_cast<double2*>(phasors), B, A, F, P, channel_bandwidth_hz, fringe,
                           (double)frequency_start_index);
    } else {
        return fail("phasor: output type must be CF32 or CF64");
    }
    return check_launch("phasor");
}
