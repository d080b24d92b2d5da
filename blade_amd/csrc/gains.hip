// Per-antenna gain steering and its gradient, for gain calibration on top of
// the fused Stokes-I loss + phasor gradient (adjoint.hip).  No reference
// counterpart: the reference's Phasor module applies fixed calibration
// coefficients (src/modules/phasor/ata.cc:188-204, the `cal` factor); these
// kernels let that factor be learned as unit-modulus gains u = g / |g|.
//
//   steer:    w_eff[b,a,f,0,p] = w[b,a,f,0,p] * u[a,f,p] * r[a,f], with
//             r = exp(-j 2 pi freq[f] tau[a]) for per-antenna delays tau (FP64
//             phase, reduced to turns before sincospi as the phasor kernel
//             does); either factor may be absent.
//   gradient: given gw = dL/dconj(w_eff) for a real loss L,
//             G = dL/dconj(u) = sum_b conj(w) gw, and through u = g / |g|
//             (dconj(u)/dconj(g) = 1 / (2|g|), du/dconj(g) = -u^2 / (2|g|))
//             dL/dconj(g) = (G - u^2 conj(G)) / (2 |g|),
//             tangent to the unit circle.  |g| is floored at 1e-30 as in
//             g / |g|.clamp_min(1e-30); below the floor u = g * 1e30 and
//             dL/dconj(g) = G * 1e30.
//   Adam: the gradient above, times 2 (PyTorch's convention for complex
//             parameters) and a loss scale, then one Adam step on the real
//             and imaginary parts of g, as torch.optim.Adam (no weight
//             decay, no amsgrad) computes it:
//               m = m + (1 - b1)(grad - m),  v = b2 v + (1 - b2) grad^2,
//               g = g - lr / (1 - b1^t) * m / (sqrt(v) / sqrt(1 - b2^t) + eps),
//             with the step count t a device scalar (incremented first, so a
//             captured graph replays successive steps).
//   delay Adam: with S[a,f,p] = sum_b conj(w_eff) gw, and w_eff depending
//             on tau through r, dL/dtau_a = 2 Re sum gw conj(dw_eff/dtau)
//             = -4 pi sum_{f,p} freq[f] Im S[a,f,p]; times a loss scale, then
//             one Adam step on tau in FP64 (torch.optim.Adam on an FP64
//             parameter), one workgroup per antenna (FP64 sums in a fixed
//             order).
// All are HBM streams over B*A*F*P complex values (one thread per output).
#include "common.hh"

namespace bl {

constexpr float kGainFloor = 1e-30f;

__device__ __forceinline__ float2 unit_gain(float2 g, float& n) {
    n = sqrtf(g.x * g.x + g.y * g.y);
    const float d = 1.f / fmaxf(n, kGainFloor);
    return make_float2(g.x * d, g.y * d);
}

// exp(-j 2 pi freq tau), the phase in FP64 turns reduced to [-1/2, 1/2].
__device__ __forceinline__ float2 delay_rot(double freq, double tau) {
    const double turns = -freq * tau;
    const double r = turns - rint(turns);
    double sn, cs;
    sincospi(2.0 * r, &sn, &cs);
    return make_float2((float)cs, (float)sn);
}

__global__ __launch_bounds__(256) void steer_gains_kernel(const float2* __restrict__ w, const float2* __restrict__ g,
                                                         const double* __restrict__ tau,
                                                         const double* __restrict__ freq, float2* __restrict__ out,
                                                         uint64_t n, uint64_t afp, int F, int P) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint64_t e = i % afp;                       // (a, f, p)
    float2 v = w[i];
    if (g) {
        float norm;
        v = cmul(v, unit_gain(g[e], norm));
    }
    if (tau) {
        const uint64_t af = e / P;
        v = cmul(v, delay_rot(freq[af % F], tau[af / F]));
    }
    out[i] = v;
}

// dL/dconj(g[i]) from the phasors, their gradient and the gains.
__device__ __forceinline__ float2 gain_grad_at(const float2* __restrict__ w, const float2* __restrict__ gw, float2 gi,
                                               uint64_t i, uint64_t afp, int B) {
    // G = sum_b conj(w) gw, beams in order.
    float2 G = make_float2(0.f, 0.f);
    for (int b = 0; b < B; b++) {
        const float2 wv = w[b * afp + i], gv = gw[b * afp + i];
        G.x += wv.x * gv.x + wv.y * gv.y;
        G.y += wv.x * gv.y - wv.y * gv.x;
    }
    float norm;
    const float2 u = unit_gain(gi, norm);
    if (norm < kGainFloor) return make_float2(G.x / kGainFloor, G.y / kGainFloor);
    // (G - u^2 conj(G)) / (2 |g|)
    const float2 u2 = cmul(u, u);
    const float2 t = cmul(u2, make_float2(G.x, -G.y));
    const float s = 0.5f / norm;
    return make_float2((G.x - t.x) * s, (G.y - t.y) * s);
}

__global__ __launch_bounds__(256) void gain_gradient_kernel(const float2* __restrict__ w, const float2* __restrict__ gw,
                                                           const float2* __restrict__ g, float2* __restrict__ gg,
                                                           uint64_t afp, int B) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= afp) return;
    gg[i] = gain_grad_at(w, gw, g[i], i, afp, B);
}

__global__ void adam_tick_kernel(float* step) {
    if (threadIdx.x == 0) *step += 1.f;
}

struct AdamArgs {
    float lr, beta1, beta2, eps, grad_scale;
};

__device__ __forceinline__ void adam_update(float& p, float& m, float& v, float grad, const AdamArgs& h, float bc1,
                                            float bc2_sqrt) {
    m = m + (1.f - h.beta1) * (grad - m);
    v = h.beta2 * v + (1.f - h.beta2) * grad * grad;
    p = p - (h.lr / bc1) * m / (sqrtf(v) / bc2_sqrt + h.eps);
}

__global__ __launch_bounds__(256) void gain_adam_kernel(const float2* __restrict__ w, const float2* __restrict__ gw,
                                                       float2* __restrict__ g, float2* __restrict__ m,
                                                       float2* __restrict__ v, const float* __restrict__ step,
                                                       AdamArgs h, uint64_t afp, int B) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= afp) return;
    float2 gi = g[i];
    const float2 d = gain_grad_at(w, gw, gi, i, afp, B);
    const float t = *step;
    const float bc1 = 1.f - powf(h.beta1, t);
    const float bc2_sqrt = sqrtf(1.f - powf(h.beta2, t));
    float2 mi = m[i], vi = v[i];
    adam_update(gi.x, mi.x, vi.x, 2.f * h.grad_scale * d.x, h, bc1, bc2_sqrt);
    adam_update(gi.y, mi.y, vi.y, 2.f * h.grad_scale * d.y, h, bc1, bc2_sqrt);
    g[i] = gi;
    m[i] = mi;
    v[i] = vi;
}

__global__ __launch_bounds__(256) void delay_adam_kernel(const float2* __restrict__ weff,
                                                        const float2* __restrict__ gw,
                                                        const double* __restrict__ freq, double* __restrict__ tau,
                                                        double* __restrict__ m, double* __restrict__ v,
                                                        const float* __restrict__ step, double lr, double beta1,
                                                        double beta2, double eps, double grad_scale, int B, int F,
                                                        int P, uint64_t afp) {
    __shared__ double s_sum[256];
    const int a = blockIdx.x;
    double acc = 0.0;
    // -4 pi sum_{f,p} freq Im(sum_b conj(w_eff) gw): fp over the threads, beams in order.
    for (int fp = threadIdx.x; fp < F * P; fp += blockDim.x) {
        const uint64_t e = (uint64_t)a * F * P + fp;
        double im = 0.0;
        for (int b = 0; b < B; b++) {
            const float2 wv = weff[b * afp + e], gv = gw[b * afp + e];
            im += (double)wv.x * gv.y - (double)wv.y * gv.x;
        }
        acc += freq[fp / P] * im;
    }
    s_sum[threadIdx.x] = acc;
    __syncthreads();
    for (int off = blockDim.x / 2; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) s_sum[threadIdx.x] += s_sum[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const double grad = grad_scale * (-4.0 * 3.14159265358979323846 * s_sum[0]);
        const double t = (double)*step;
        const double bc1 = 1.0 - pow(beta1, t), bc2_sqrt = sqrt(1.0 - pow(beta2, t));
        double mi = m[a], vi = v[a];
        mi = mi + (1.0 - beta1) * (grad - mi);
        vi = beta2 * vi + (1.0 - beta2) * grad * grad;
        m[a] = mi;
        v[a] = vi;
        tau[a] = tau[a] - (lr / bc1) * mi / (sqrt(vi) / bc2_sqrt + eps);
    }
}

// Shape checks first: empty inputs (a zero axis, whose tensors may have null
// data pointers) succeed without a launch, as every entry point does.
static bool gain_shape_empty(uint64_t B, uint64_t A, uint64_t F) { return B == 0 || A == 0 || F == 0; }

static int check_gain_args(const char* who, const void* a, const void* b, const void* c, uint64_t B, uint64_t A,
                           uint64_t F, uint64_t P) {
    if (!a || !b || !c) return fail(std::string(who) + ": null pointer");
    if ((uintptr_t)a % 8 || (uintptr_t)b % 8 || (uintptr_t)c % 8) return fail(std::string(who) + ": misaligned buffer");
    if (P != 2) return fail(std::string(who) + ": number of polarizations must be 2");
    if (B > 4096 || A > 4096 || F > (1u << 24)) return fail(std::string(who) + ": shape too large");
    return BLADE_SUCCESS;
}

}  // namespace bl

using namespace bl;

extern "C" int blade_steer_gains(const void* phasors, const void* gains, const double* delays,
                                 const double* freqs_hz, void* output, uint64_t B, uint64_t A, uint64_t F,
                                 uint64_t P, blade_stream_t stream) {
    if (P != 2) return fail("gain steering: number of polarizations must be 2");
    if (gain_shape_empty(B, A, F)) return BLADE_SUCCESS;
    int rc = check_gain_args("gain steering", phasors, output, output, B, A, F, P);
    if (rc) return rc;
    if (gains && (uintptr_t)gains % 8) return fail("gain steering: misaligned gains");
    if (delays && (!freqs_hz || (uintptr_t)delays % 8 || (uintptr_t)freqs_hz % 8))
        return fail("gain steering: delays need aligned FP64 delays and channel frequencies");
    const uint64_t afp = A * F * P, n = B * afp;
    if (n == 0) return BLADE_SUCCESS;
    hipLaunchKernelGGL(steer_gains_kernel, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0,
                       reinterpret_cast<hipStream_t>(stream), static_cast<const float2*>(phasors),
                       static_cast<const float2*>(gains), delays, delays ? freqs_hz : nullptr,
                       static_cast<float2*>(output), n, afp, (int)F, (int)P);
    return check_launch("gain steering");
}

extern "C" int blade_gain_gradient(const void* phasors, const void* grad_phasors, const void* gains,
                                   void* grad_gains, uint64_t B, uint64_t A, uint64_t F, uint64_t P,
                                   blade_stream_t stream) {
    if (P != 2) return fail("gain gradient: number of polarizations must be 2");
    if (gain_shape_empty(B, A, F)) return BLADE_SUCCESS;
    int rc = check_gain_args("gain gradient", phasors, grad_phasors, gains, B, A, F, P);
    if (!rc) rc = check_gain_args("gain gradient", grad_gains, gains, gains, B, A, F, P);
    if (rc) return rc;
    const uint64_t afp = A * F * P;
    if (afp == 0) return BLADE_SUCCESS;
    hipLaunchKernelGGL(gain_gradient_kernel, dim3((unsigned)ceil_div(afp, 256)), dim3(256), 0,
                       reinterpret_cast<hipStream_t>(stream), static_cast<const float2*>(phasors),
                       static_cast<const float2*>(grad_phasors), static_cast<const float2*>(gains),
                       static_cast<float2*>(grad_gains), afp, (int)B);
    return check_launch("gain gradient");
}

extern "C" int blade_gain_adam(const void* phasors, const void* grad_phasors, void* gains, void* exp_avg,
                               void* exp_avg_sq, float* step, float lr, float beta1, float beta2, float eps,
                               float grad_scale, uint64_t B, uint64_t A, uint64_t F, uint64_t P,
                               blade_stream_t stream) {
    if (P != 2) return fail("gain Adam: number of polarizations must be 2");
    if (gain_shape_empty(B, A, F)) return BLADE_SUCCESS;
    int rc = check_gain_args("gain Adam", phasors, grad_phasors, gains, B, A, F, P);
    if (!rc) rc = check_gain_args("gain Adam", exp_avg, exp_avg_sq, gains, B, A, F, P);
    if (rc) return rc;
    if (!step || (uintptr_t)step % 4) return fail("gain Adam: null or misaligned step counter");
    if (!(lr > 0.f) || !(beta1 >= 0.f && beta1 < 1.f) || !(beta2 >= 0.f && beta2 < 1.f) || !(eps >= 0.f))
        return fail("gain Adam: need lr > 0, 0 <= beta1, beta2 < 1 and eps >= 0");
    const uint64_t afp = A * F * P;
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    hipLaunchKernelGGL(adam_tick_kernel, dim3(1), dim3(64), 0, s, step);
    if (afp > 0) {
        const AdamArgs h{lr, beta1, beta2, eps, grad_scale};
        hipLaunchKernelGGL(gain_adam_kernel, dim3((unsigned)ceil_div(afp, 256)), dim3(256), 0, s,
                           static_cast<const float2*>(phasors), static_cast<const float2*>(grad_phasors),
                           static_cast<float2*>(gains), static_cast<float2*>(exp_avg),
                           static_cast<float2*>(exp_avg_sq), static_cast<const float*>(step), h, afp, (int)B);
    }
    return check_launch("gain Adam");
}

extern "C" int blade_delay_adam(const void* steered, const void* grad_phasors, const double* freqs_hz,
                                double* delays, double* exp_avg, double* exp_avg_sq, float* step, double lr,
                                double beta1, double beta2, double eps, double grad_scale, uint64_t B,
                                uint64_t A, uint64_t F, uint64_t P, blade_stream_t stream) {
    if (P != 2) return fail("delay Adam: number of polarizations must be 2");
    if (gain_shape_empty(B, A, F)) return BLADE_SUCCESS;
    int rc = check_gain_args("delay Adam", steered, grad_phasors, freqs_hz, B, A, F, P);
    if (!rc) rc = check_gain_args("delay Adam", delays, exp_avg, exp_avg_sq, B, A, F, P);
    if (rc) return rc;
    if (!step || (uintptr_t)step % 4) return fail("delay Adam: null or misaligned step counter");
    if (!(lr > 0.0) || !(beta1 >= 0.0 && beta1 < 1.0) || !(beta2 >= 0.0 && beta2 < 1.0) || !(eps >= 0.0))
        return fail("delay Adam: need lr > 0, 0 <= beta1, beta2 < 1 and eps >= 0");
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    hipLaunchKernelGGL(adam_tick_kernel, dim3(1), dim3(64), 0, s, step);
    if (A > 0)
        hipLaunchKernelGGL(delay_adam_kernel, dim3((unsigned)A), dim3(256), 0, s,
                           static_cast<const float2*>(steered), static_cast<const float2*>(grad_phasors), freqs_hz,
                           delays, exp_avg, exp_avg_sq, static_cast<const float*>(step), lr, beta1, beta2, eps,
                           grad_scale, (int)B, (int)F, (int)P, A * F * P);
    return check_launch("delay Adam");
}
