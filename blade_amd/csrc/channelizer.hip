// FFT channelizer:  out[a, f*C + k, s, p] = fftshift(DFT_C(in[a, f, s*C:(s+1)*C, p]))[k]
// (oracle tests/modules/channelizer/advanced.py:49-65; forward, unnormalised,
// e^{-j 2pi k n / C}).  Written from the Stockham autosort formulation of the
// mixed-radix FFT (each pass reads stride C/R, twiddles by W_{Ns R}^{(j mod Ns) r},
// butterflies, writes to (j/Ns)*Ns*R + j mod Ns + r*Ns), so the output lands
// in natural order without a bit-reversal pass.  fftshift is an index remap on
// the store.  Twiddles come from a per-C table built in FP64 and rounded
// once to FP32.
#include "common.hh"

#include <algorithm>
#include <cmath>
#include <map>
#include <mutex>
#include <vector>

namespace bl {

// ---------------------------------------------------------------------------
// Twiddle tables  W_C^m = exp(-j 2 pi m / C),  m in [0, C), one per (device, C).
// ---------------------------------------------------------------------------

__global__ void twiddle_kernel(float2* tw, int C) {
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m < C) {
        double s, c;
        sincospi(-2.0 * (double)m / (double)C, &s, &c);
        tw[m] = make_float2((float)c, (float)s);
    }
}

// Each table is followed by a small sink that idle lanes of the radix-16
// kernel store into, so that every lane issues the same stores.
constexpr size_t kTwiddleSinkBytes = 512 * sizeof(float4);

static std::mutex g_tw_mutex;
static std::map<std::pair<int, int>, float2*> g_tw;

static int get_twiddles(int C, hipStream_t s, const float2** out) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return fail("channelizer: hipGetDevice failed");
    std::lock_guard<std::mutex> lock(g_tw_mutex);
    auto it = g_tw.find({dev, C});
    if (it != g_tw.end()) {
        *out = it->second;
        return BLADE_SUCCESS;
    }
    float2* p = nullptr;
    hipError_t e = hipMalloc(&p, sizeof(float2) * C + kTwiddleSinkBytes);
    if (e != hipSuccess) {
        set_error(std::string("channelizer: twiddle allocation: ") + hipGetErrorString(e));
        return BLADE_HIP_ERROR;
    }
    hipLaunchKernelGGL(twiddle_kernel, dim3((C + 255) / 256), dim3(256), 0, s, p, C);
    int rc = check_launch("channelizer twiddles");
    if (rc) return rc;
    g_tw[{dev, C}] = p;
    *out = p;
    return BLADE_SUCCESS;
}

// ---------------------------------------------------------------------------
// Small DFT in registers.
// ---------------------------------------------------------------------------

__device__ __forceinline__ float2 mul_neg_j(float2 a) { return make_float2(a.y, -a.x); }

template <int R>
__device__ __forceinline__ void dft_reg(float2 (&v)[R], const float2* __restrict__ tw, int twstep) {
    if constexpr (R == 1) {
        return;
    } else if constexpr (R == 2) {
        const float2 a = v[0], b = v[1];
        v[0] = cadd(a, b);
        v[1] = csub(a, b);
    } else if constexpr (R == 4) {
        const float2 s02 = cadd(v[0], v[2]), d02 = csub(v[0], v[2]);
        const float2 s13 = cadd(v[1], v[3]), d13 = mul_neg_j(csub(v[1], v[3]));
        v[0] = cadd(s02, s13);
        v[2] = csub(s02, s13);
        v[1] = cadd(d02, d13);
        v[3] = csub(d02, d13);
    } else if constexpr (R == 8) {
        float2 e[4] = {v[0], v[2], v[4], v[6]};
        float2 o[4] = {v[1], v[3], v[5], v[7]};
        dft_reg<4>(e, tw, 0);
        dft_reg<4>(o, tw, 0);
        const float c = 0.70710678118654752440f;
        const float2 o1 = make_float2(c * (o[1].x + o[1].y), c * (o[1].y - o[1].x));   // * e^{-j pi/4}
        const float2 o2 = mul_neg_j(o[2]);                                            // * e^{-j pi/2}
        const float2 o3 = make_float2(c * (o[3].y - o[3].x), -c * (o[3].x + o[3].y)); // * e^{-j 3pi/4}
        v[0] = cadd(e[0], o[0]); v[4] = csub(e[0], o[0]);
        v[1] = cadd(e[1], o1);   v[5] = csub(e[1], o1);
        v[2] = cadd(e[2], o2);   v[6] = csub(e[2], o2);
        v[3] = cadd(e[3], o3);   v[7] = csub(e[3], o3);
    } else {
        // Direct DFT of size R with W_R^q = tw[q * twstep].
        float2 r[R];
#pragma unroll
        for (int k = 0; k < R; k++) {
            float2 acc = v[0];
#pragma unroll
            for (int n = 1; n < R; n++) {
                const float2 w = tw[((n * k) % R) * twstep];
                acc.x += v[n].x * w.x - v[n].y * w.y;
                acc.y += v[n].x * w.y + v[n].y * w.x;
            }
            r[k] = acc;
        }
#pragma unroll
        for (int k = 0; k < R; k++) v[k] = r[k];
    }
}

// Constant-twiddle DFT for the register kernel (W computed at compile time
// from the table argument in a small struct).
template <int C>
struct SmallTw {
    float2 w[C];
};

template <int C>
__device__ __forceinline__ void dft_small(float2 (&v)[C], const SmallTw<C>& tw) {
    if constexpr (C == 2 || C == 4 || C == 8) {
        dft_reg<C>(v, nullptr, 0);
    } else {
        float2 r[C];
#pragma unroll
        for (int k = 0; k < C; k++) {
            float2 acc = v[0];
#pragma unroll
            for (int n = 1; n < C; n++) {
                const float2 w = tw.w[(n * k) % C];
                acc.x += v[n].x * w.x - v[n].y * w.y;
                acc.y += v[n].x * w.y + v[n].y * w.x;
            }
            r[k] = acc;
        }
#pragma unroll
        for (int k = 0; k < C; k++) v[k] = r[k];
    }
}

// One lane per C-sample spectrum (all P pols), C <= 16.  Reads C*P contiguous
// complex samples, writes one sample per output row (coalesced across lanes).
template <int C, int P>
__global__ __launch_bounds__(256) void chan_small_kernel(const float2* __restrict__ in,
                                                         float2* __restrict__ out,
                                                         uint64_t n_spec, uint64_t S,
                                                         SmallTw<C> tw) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t sp = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; sp < n_spec; sp += stride) {
        const uint64_t af = sp / S, s = sp - af * S;
        float2 v[P][C];
        const float2* src = in + sp * C * P;
#pragma unroll
        for (int n = 0; n < C; n++) {
            if constexpr (P == 2) {
                const float4 q = reinterpret_cast<const float4*>(src)[n];
                v[0][n] = make_float2(q.x, q.y);
                v[1][n] = make_float2(q.z, q.w);
            } else {
                v[0][n] = src[n];
            }
        }
#pragma unroll
        for (int p = 0; p < P; p++) dft_small<C>(v[p], tw);
#pragma unroll
        for (int k = 0; k < C; k++) {
            const int ko = (k + C / 2) % C;
            const uint64_t o = ((af * C + ko) * S + s) * P;
            if constexpr (P == 2) {
                reinterpret_cast<float4*>(out + o)[0] =
                    make_float4(v[0][k].x, v[0][k].y, v[1][k].x, v[1][k].y);
            } else {
                out[o] = v[0][k];
            }
        }
    }
}

// ---------------------------------------------------------------------------
// LDS Stockham kernel: one workgroup per spectrum (all P pols), C up to 8192
// (P = 2) or 16384 (P = 1).
// ---------------------------------------------------------------------------

struct Radices {
    int r[16];
    int n;
};

constexpr int kMaxLdsComplex = 16384;   // 128 KiB of LDS

template <int R>
constexpr int groups_max() {
    return (kMaxLdsComplex / R + 1023) / 1024;
}

// One radix-R pass over PC / C sequences of C points in LDS, sequence p at
// data + p * pstride (pstride = C by default).
template <int R>
__device__ __forceinline__ void stockham_pass(float2* __restrict__ data, int C, int PC, int Ns,
                                              const float2* __restrict__ tw, int pstride = 0) {
    constexpr int G = groups_max<R>();
    const int nthr = blockDim.x;
    const int per_pol = C / R;
    const int total = PC / R;
    const int twstride = C / (Ns * R);
    if (pstride == 0) pstride = C;
    float2 v[G][R];
#pragma unroll
    for (int g = 0; g < G; g++) {
        const int gi = threadIdx.x + g * nthr;
        if (gi < total) {
            const int p = gi / per_pol, j = gi - p * per_pol;
            const float2* base = data + p * pstride;
            const int jm = j % Ns;
#pragma unroll
            for (int r = 0; r < R; r++) {
                float2 x = base[j + r * per_pol];
                if (r > 0 && jm > 0) {
                    const float2 w = tw[jm * r * twstride];
                    x = cmul(x, w);
                }
                v[g][r] = x;
            }
            dft_reg<R>(v[g], tw, C / R);
        }
    }
    __syncthreads();
#pragma unroll
    for (int g = 0; g < G; g++) {
        const int gi = threadIdx.x + g * nthr;
        if (gi < total) {
            const int p = gi / per_pol, j = gi - p * per_pol;
            float2* base = data + p * pstride;
            const int jm = j % Ns;
            const int d = (j - jm) * R + jm;
#pragma unroll
            for (int r = 0; r < R; r++) base[d + r * Ns] = v[g][r];
        }
    }
    __syncthreads();
}

template <int P>
__global__ __launch_bounds__(1024) void chan_lds_kernel(const float2* __restrict__ in,
                                                        float2* __restrict__ out, int C,
                                                        uint64_t S, uint64_t n_spec,
                                                        const float2* __restrict__ tw,
                                                        Radices rad) {
    extern __shared__ float2 data[];
    const int PC = P * C;
    for (uint64_t sp = blockIdx.x; sp < n_spec; sp += gridDim.x) {
        const uint64_t af = sp / S, s = sp - af * S;
        const float2* src = in + sp * (uint64_t)C * P;
        for (int n = threadIdx.x; n < C; n += blockDim.x) {
            if constexpr (P == 2) {
                const float4 q = reinterpret_cast<const float4*>(src)[n];
                data[n] = make_float2(q.x, q.y);
                data[C + n] = make_float2(q.z, q.w);
            } else {
                data[n] = src[n];
            }
        }
        __syncthreads();
        int Ns = 1;
        for (int i = 0; i < rad.n; i++) {
            const int R = rad.r[i];
            switch (R) {
                case 2: stockham_pass<2>(data, C, PC, Ns, tw); break;
                case 3: stockham_pass<3>(data, C, PC, Ns, tw); break;
                case 4: stockham_pass<4>(data, C, PC, Ns, tw); break;
                case 5: stockham_pass<5>(data, C, PC, Ns, tw); break;
                case 7: stockham_pass<7>(data, C, PC, Ns, tw); break;
                default: stockham_pass<8>(data, C, PC, Ns, tw); break;
            }
            Ns *= R;
        }
        const int half = C / 2;
        for (int k = threadIdx.x; k < C; k += blockDim.x) {
            int ko = k + half;
            if (ko >= C) ko -= C;
            const uint64_t o = ((af * C + ko) * S + s) * P;
            if constexpr (P == 2) {
                reinterpret_cast<float4*>(out + o)[0] =
                    make_float4(data[k].x, data[k].y, data[C + k].x, data[C + k].y);
            } else {
                out[o] = data[k];
            }
        }
        __syncthreads();
    }
}

// All passes of `rad` over PC / N sequences of N points at stride pstride.
__device__ __forceinline__ void stockham_all(float2* data, int N, int PC, const float2* tw, const Radices& rad,
                                             int pstride) {
    int Ns = 1;
    for (int i = 0; i < rad.n; i++) {
        switch (rad.r[i]) {
            case 2: stockham_pass<2>(data, N, PC, Ns, tw, pstride); break;
            case 3: stockham_pass<3>(data, N, PC, Ns, tw, pstride); break;
            case 4: stockham_pass<4>(data, N, PC, Ns, tw, pstride); break;
            case 5: stockham_pass<5>(data, N, PC, Ns, tw, pstride); break;
            case 7: stockham_pass<7>(data, N, PC, Ns, tw, pstride); break;
            default: stockham_pass<8>(data, N, PC, Ns, tw, pstride); break;
        }
        Ns *= rad.r[i];
    }
}

// ---------------------------------------------------------------------------
// Rates past one workgroup's LDS (P * C > 16384): two passes through a
// scratch buffer (the four-step FFT).  With C = N1 * N2, n = n1 + N1 n2 and
// k = k2 + N2 k1:
//   pass 1, each n1:  Y[n1, k2] = W_C^(n1 k2) sum_n2 x[n1 + N1 n2] W_N2^(n2 k2)
//   pass 2, each k2:  X[k2 + N2 k1] = sum_n1 Y[n1, k2] W_N1^(n1 k1)
// and the fftshift of X moves k1 by N1 / 2 (N1 even).  Scratch layout
// [spectrum][pol][n1][k2].  A workgroup (1024 threads) transforms a tile of
// `cols` consecutive n1 (pass 1) or k2 (pass 2) of one spectrum, both pols,
// so every global access moves `cols` consecutive samples; the LDS sequences
// are padded by one sample against bank conflicts.
// ---------------------------------------------------------------------------

template <int P>
__global__ __launch_bounds__(1024) void chan_big_pass1(const float2* __restrict__ in, float2* __restrict__ y,
                                                       int C, int N1, int N2, int cols, uint64_t n_spec,
                                                       const float2* __restrict__ twN2,
                                                       const float2* __restrict__ twC, Radices rad) {
    extern __shared__ float2 data[];
    const int stride = N2 + 1;
    const int L = P * cols;
    const uint64_t tiles = (uint64_t)(N1 / cols);
    for (uint64_t it = blockIdx.x; it < n_spec * tiles; it += gridDim.x) {
        const uint64_t sp = it / tiles;
        const int n1_0 = (int)(it - sp * tiles) * cols;
        const float2* src = in + sp * (uint64_t)C * P;
        for (int e = threadIdx.x; e < cols * N2; e += blockDim.x) {
            const int n2 = e / cols, c = e - n2 * cols;
            const uint64_t n = (uint64_t)(n1_0 + c) + (uint64_t)N1 * n2;
            if constexpr (P == 2) {
                const float4 q = reinterpret_cast<const float4*>(src)[n];
                data[c * stride + n2] = make_float2(q.x, q.y);
                data[(cols + c) * stride + n2] = make_float2(q.z, q.w);
            } else {
                data[c * stride + n2] = src[n];
            }
        }
        __syncthreads();
        stockham_all(data, N2, L * N2, twN2, rad, stride);
        for (int e = threadIdx.x; e < L * N2; e += blockDim.x) {
            const int q = e / N2, k2 = e - q * N2;
            const int pp = q / cols, c = q - pp * cols;
            // n1 k2 < N1 N2 = C: the exponent needs no reduction
            const int n1 = n1_0 + c;
            const float2 w = twC[n1 * k2];
            y[((sp * P + pp) * N1 + n1) * (uint64_t)N2 + k2] = cmul(data[q * stride + k2], w);
        }
        __syncthreads();
    }
}

template <int P>
__global__ __launch_bounds__(1024) void chan_big_pass2(const float2* __restrict__ y, float2* __restrict__ out,
                                                       int C, int N1, int N2, int cols, uint64_t n_spec,
                                                       uint64_t S, const float2* __restrict__ twN1, Radices rad) {
    extern __shared__ float2 data[];
    const int stride = N1 + 1;
    const int L = P * cols;
    const uint64_t tiles = (uint64_t)(N2 / cols);
    const int half1 = N1 / 2;
    for (uint64_t it = blockIdx.x; it < n_spec * tiles; it += gridDim.x) {
        const uint64_t sp = it / tiles;
        const int k2_0 = (int)(it - sp * tiles) * cols;
        for (int e = threadIdx.x; e < L * N1; e += blockDim.x) {
            const int c = e % cols, r = e / cols;
            const int n1 = r % N1, pp = r / N1;
            data[(pp * cols + c) * stride + n1] = y[((sp * P + pp) * N1 + n1) * (uint64_t)N2 + k2_0 + c];
        }
        __syncthreads();
        stockham_all(data, N1, L * N1, twN1, rad, stride);
        const uint64_t af = sp / S, s = sp - af * S;
        for (int e = threadIdx.x; e < cols * N1; e += blockDim.x) {
            const int k1 = e / cols, c = e - k1 * cols;
            int k1o = k1 + half1;
            if (k1o >= N1) k1o -= N1;
            const uint64_t ko = (uint64_t)(k2_0 + c) + (uint64_t)N2 * k1o;
            const uint64_t o = ((af * C + ko) * S + s) * P;
            if constexpr (P == 2) {
                const float2 X = data[c * stride + k1], Y = data[(cols + c) * stride + k1];
                reinterpret_cast<float4*>(out + o)[0] = make_float4(X.x, X.y, Y.x, Y.y);
            } else {
                out[o] = data[c * stride + k1];
            }
        }
        __syncthreads();
    }
}

// Channelizer with the detector fused on the output: one workgroup per
// (a, f) walks its S = T/C spectra, transforms both pols in LDS and
// integrates |X|^2+|Y|^2 (or XX*, YY*, Re XY*, Im XY*) of every fine channel
// over N consecutive spectra in registers, writing
//   out[a, f*C + k', s/N, q]   (N <= S, N | S)
//   out[a, f*C + k', 0, q]    (+=) when N spans several blocks (accumulate).
// The CF32 spectra never reach HBM.
// Detector sums of (X, Y) turned into those of (L, R) = (X + jY, X - jY),
// the polarizer (polarizer.cu:5-25) applied before the detector: sums are
// linear in [XX*, YY*, Re XY*, Im XY*], so the map applies after integration.
//   Stokes I doubles; |L|^2 = XX + YY + 2 Im, |R|^2 = XX + YY - 2 Im,
//   Re LR* = XX - YY, Im LR* = 2 Re.
template <int NPOL>
__device__ inline void circular_sums(float (&a)[NPOL]) {
    if constexpr (NPOL == 1) {
        a[0] *= 2.f;
    } else {
        const float xx = a[0], yy = a[1], re = a[2], im = a[3];
        a[0] = (xx + yy) + 2.f * im;
        a[1] = (xx + yy) - 2.f * im;
        a[2] = xx - yy;
        a[3] = 2.f * re;
    }
}

template <int NPOL, int KPT>
__global__ __launch_bounds__(1024) void chan_detect_kernel(const float2* __restrict__ in,
                                                           float* __restrict__ out, int C,
                                                           int S, int N, uint64_t n_af,
                                                           const float2* __restrict__ tw,
                                                           Radices rad, int flags) {
    extern __shared__ float2 data[];
    const int accumulate = flags & 1;
    const bool circular = flags & 2;
    const int PC = 2 * C;
    const int groups = (N <= S) ? S / N : 1;
    const int half = C / 2;
    for (uint64_t af = blockIdx.x; af < n_af; af += gridDim.x) {
        float acc[KPT][NPOL];
#pragma unroll
        for (int j = 0; j < KPT; j++)
#pragma unroll
            for (int q = 0; q < NPOL; q++) acc[j][q] = 0.f;
        for (int s = 0; s < S; s++) {
            const float4* src = reinterpret_cast<const float4*>(in + ((af * S + s) * (uint64_t)C) * 2);
            for (int n = threadIdx.x; n < C; n += blockDim.x) {
                const float4 q = src[n];
                data[n] = make_float2(q.x, q.y);
                data[C + n] = make_float2(q.z, q.w);
            }
            __syncthreads();
            int Ns = 1;
            for (int i = 0; i < rad.n; i++) {
                const int R = rad.r[i];
                switch (R) {
                    case 2: stockham_pass<2>(data, C, PC, Ns, tw); break;
                    case 3: stockham_pass<3>(data, C, PC, Ns, tw); break;
                    case 4: stockham_pass<4>(data, C, PC, Ns, tw); break;
                    case 5: stockham_pass<5>(data, C, PC, Ns, tw); break;
                    case 7: stockham_pass<7>(data, C, PC, Ns, tw); break;
                    default: stockham_pass<8>(data, C, PC, Ns, tw); break;
                }
                Ns *= R;
            }
#pragma unroll
            for (int j = 0; j < KPT; j++) {
                const int k = threadIdx.x + j * blockDim.x;
                if (k < C) {
                    const float2 X = data[k], Y = data[C + k];
                    acc[j][0] += X.x * X.x + X.y * X.y + (NPOL == 1 ? Y.x * Y.x + Y.y * Y.y : 0.f);
                    if constexpr (NPOL == 4) {
                        acc[j][1] += Y.x * Y.x + Y.y * Y.y;
                        acc[j][2] += X.x * Y.x + X.y * Y.y;
                        acc[j][3] += X.y * Y.x - X.x * Y.y;
                    }
                }
            }
            __syncthreads();
            const bool flush = (N <= S) ? ((s + 1) % N == 0) : (s == S - 1);
            if (flush) {
                const int grp = (N <= S) ? s / N : 0;
                if (circular) {
#pragma unroll
                    for (int j = 0; j < KPT; j++) circular_sums<NPOL>(acc[j]);
                }
#pragma unroll
                for (int j = 0; j < KPT; j++) {
                    const int k = threadIdx.x + j * blockDim.x;
                    if (k < C) {
                        int ko = k + half;
                        if (ko >= C) ko -= C;
                        float* o = out + ((af * C + ko) * groups + grp) * NPOL;
#pragma unroll
                        for (int q = 0; q < NPOL; q++) {
                            o[q] = accumulate ? o[q] + acc[j][q] : acc[j][q];
                            acc[j][q] = 0.f;
                        }
                    }
                }
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Power-of-two rates 16 < C < 256 with both polarizations (fft16_kernel
// below takes 256 <= C <= 8192).
//
// A spectrum is handled by C/8 threads; a workgroup holds SPW = max(1,
// 2048/C) spectra.  Pass 1 (radix 8) runs on registers loaded straight from
// HBM (eight 16-byte loads per thread, both pols, all in flight at once);
// the radix-8 middle passes run in LDS.  The last pass (radix RL = 2, 4 or 8)
// reads LDS and writes the fftshifted spectrum (MODE 0) or integrates the
// detector (MODE 1: Stokes I, MODE 4: XX*, YY*, Re XY*, Im XY*) in registers,
// so the spectra of the fused channelizer + detector never reach HBM.
// ---------------------------------------------------------------------------
template <int MODE, int RL>
__global__ __launch_bounds__(1024) void fft2_kernel(const float4* __restrict__ in, void* __restrict__ outv,
                                                    int C, int nmid, uint64_t S, uint64_t n_items,
                                                    int N, int flags,
                                                    const float2* __restrict__ tw) {
    extern __shared__ float2 data[];
    const int accumulate = flags & 1;
    const bool circular = flags & 2;
    constexpr int G = 8 / RL;                   // last-pass butterflies per thread and pol
    const int tps = C >> 3;                     // threads per spectrum
    const int spw = blockDim.x / tps;
    const int slot = threadIdx.x / tps;
    const int j = threadIdx.x - slot * tps;
    const int PC = spw * 2 * C;
    float2* const blk = data + (size_t)slot * 2 * C;    // pol 0; pol 1 at + C
    const uint64_t per_round = (uint64_t)gridDim.x * spw;
    const uint64_t rounds = (n_items + per_round - 1) / per_round;
    const int nspec = MODE == 0 ? 1 : (int)S;
    const int groups = MODE == 0 ? 1 : ((N <= (int)S) ? (int)S / N : 1);
    const int lastR = C / RL;                   // stride of the last pass

    auto item_of = [&](uint64_t round) { return (round * gridDim.x + blockIdx.x) * spw + slot; };
    auto spec_of = [&](uint64_t round, int s) {
        const uint64_t it = item_of(round);
        const uint64_t sp = MODE == 0 ? it : it * S + s;
        return it < n_items ? sp : 0;           // idle slots re-read spectrum 0
    };
    float4 q[8];
    auto load = [&](uint64_t sp) {
        const float4* src = in + sp * (uint64_t)C;
#pragma unroll
        for (int r = 0; r < 8; r++) q[r] = src[j + r * tps];
    };

    constexpr int NPOL = MODE == 0 ? 1 : MODE;
    float acc[G][RL][NPOL];
#pragma unroll
    for (int g = 0; g < G; g++)
#pragma unroll
        for (int r = 0; r < RL; r++)
#pragma unroll
            for (int c = 0; c < NPOL; c++) acc[g][r][c] = 0.f;

    for (uint64_t round = 0; round < rounds; round++) {
        const uint64_t item = item_of(round);
        const bool valid = item < n_items;
        for (int s = 0; s < nspec; s++) {
            load(spec_of(round, s));
            // ---- pass 1: radix 8 on registers -> LDS (Stockham, Ns = 1)
#pragma unroll
            for (int p = 0; p < 2; p++) {
                float2 v[8];
#pragma unroll
                for (int r = 0; r < 8; r++)
                    v[r] = p == 0 ? make_float2(q[r].x, q[r].y) : make_float2(q[r].z, q[r].w);
                dft_reg<8>(v, tw, 0);
#pragma unroll
                for (int r = 0; r < 8; r++) blk[p * C + j * 8 + r] = v[r];
            }
            __syncthreads();
            int Ns = 8;
            for (int i = 0; i < nmid; i++) {
                stockham_pass<8>(data, C, PC, Ns, tw);
                Ns *= 8;
            }
            // ---- last pass: radix RL, LDS -> output (Ns = C / RL)
            const uint64_t sp = MODE == 0 ? item : item * S + s;
#pragma unroll
            for (int g = 0; g < G; g++) {
                const int jb = j + g * tps;
                float2 x[2][RL];
#pragma unroll
                for (int p = 0; p < 2; p++) {
#pragma unroll
                    for (int r = 0; r < RL; r++) {
                        float2 v = blk[p * C + jb + r * lastR];
                        if (r > 0 && jb > 0) v = cmul(v, tw[jb * r]);
                        x[p][r] = v;
                    }
                    dft_reg<RL>(x[p], tw, C / RL);
                }
                if constexpr (MODE == 0) {
                    if (valid) {
                        const uint64_t af = sp / S, si = sp - af * S;
                        float4* out = static_cast<float4*>(outv);
#pragma unroll
                        for (int r = 0; r < RL; r++) {
                            const int k = jb + r * lastR;
                            const int ko = (k + (C >> 1)) & (C - 1);
                            out[(af * C + ko) * S + si] = make_float4(x[0][r].x, x[0][r].y, x[1][r].x, x[1][r].y);
                        }
                    }
                } else {
#pragma unroll
                    for (int r = 0; r < RL; r++) {
                        const float2 X = x[0][r], Y = x[1][r];
                        if constexpr (MODE == 1) {
                            acc[g][r][0] += X.x * X.x + X.y * X.y + Y.x * Y.x + Y.y * Y.y;
                        } else {
                            acc[g][r][0] += X.x * X.x + X.y * X.y;
                            acc[g][r][NPOL > 1 ? 1 : 0] += Y.x * Y.x + Y.y * Y.y;
                            acc[g][r][NPOL > 2 ? 2 : 0] += X.x * Y.x + X.y * Y.y;
                            acc[g][r][NPOL > 3 ? 3 : 0] += X.y * Y.x - X.x * Y.y;
                        }
                    }
                }
            }
            if constexpr (MODE != 0) {
                const bool flush = (N <= (int)S) ? ((s + 1) % N == 0) : (s == (int)S - 1);
                if (flush) {
                    const int grp = (N <= (int)S) ? s / N : 0;
                    float* out = static_cast<float*>(outv);
                    if (circular) {
#pragma unroll
                        for (int g = 0; g < G; g++)
#pragma unroll
                            for (int r = 0; r < RL; r++) circular_sums<NPOL>(acc[g][r]);
                    }
#pragma unroll
                    for (int g = 0; g < G; g++)
#pragma unroll
                        for (int r = 0; r < RL; r++) {
                            if (valid) {
                                const int k = j + g * tps + r * lastR;
                                const int ko = (k + (C >> 1)) & (C - 1);
                                float* o = out + ((item * C + ko) * groups + grp) * NPOL;
#pragma unroll
                                for (int c = 0; c < NPOL; c++) o[c] = accumulate ? o[c] + acc[g][r][c] : acc[g][r][c];
                            }
#pragma unroll
                            for (int c = 0; c < NPOL; c++) acc[g][r][c] = 0.f;
                        }
                }
            }
            __syncthreads();   // LDS is rewritten by the next spectrum's pass 1
        }
    }
}

// Passes of fft2_kernel for C = 2^k: radix 8 first, nmid radix-8 middle
// passes, last radix RL with 3 + 3 nmid + log2(RL) = k.
struct Fft2Plan {
    int rl, nmid, threads, spw;
};

static bool fft2_plan(int C, Fft2Plan* pl) {
    if (C < 16 || C > 8192 || (C & (C - 1))) return false;
    int k = 0;
    while ((1 << k) < C) k++;
    const int rem = (k - 3) % 3;
    pl->rl = rem == 0 ? 8 : (rem == 1 ? 2 : 4);
    const int lr = rem == 0 ? 3 : rem;
    pl->nmid = (k - 3 - lr) / 3;
    const int tps = C / 8;
    pl->threads = tps >= 256 ? tps : 256;
    pl->spw = pl->threads / tps;
    return true;
}

template <int MODE>
static int launch_fft2(const Fft2Plan& pl, const float2* in, void* out, int C, uint64_t S, uint64_t n_items,
                       int N, int accumulate, const float2* tw, hipStream_t s) {
    const size_t lds = sizeof(float2) * 2 * (size_t)C * pl.spw;
    // One workgroup per SPW spectra: the hardware refills each CU as
    // workgroups finish, which beats persistent rounds here by 2-6%.
    const uint64_t blocks = std::min<uint64_t>(ceil_div(n_items, pl.spw), 0x7fffffffull);
    const void* fns[3] = {reinterpret_cast<const void*>(&fft2_kernel<MODE, 2>),
                          reinterpret_cast<const void*>(&fft2_kernel<MODE, 4>),
                          reinterpret_cast<const void*>(&fft2_kernel<MODE, 8>)};
    const void* fn = fns[pl.rl == 2 ? 0 : (pl.rl == 4 ? 1 : 2)];
    if (lds > 65536) {
        hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) {
            set_error(std::string("channelizer: LDS attribute: ") + hipGetErrorString(e));
            return BLADE_HIP_ERROR;
        }
    }
    const float4* in4 = reinterpret_cast<const float4*>(in);
    const dim3 grid((unsigned)blocks), block((unsigned)pl.threads);
    if (pl.rl == 2)
        hipLaunchKernelGGL((fft2_kernel<MODE, 2>), grid, block, lds, s, in4, out, C, pl.nmid, S, n_items, N, accumulate, tw);
    else if (pl.rl == 4)
        hipLaunchKernelGGL((fft2_kernel<MODE, 4>), grid, block, lds, s, in4, out, C, pl.nmid, S, n_items, N, accumulate, tw);
    else
        hipLaunchKernelGGL((fft2_kernel<MODE, 8>), grid, block, lds, s, in4, out, C, pl.nmid, S, n_items, N, accumulate, tw);
    return check_launch("channelizer");
}

// ---------------------------------------------------------------------------
// Power-of-two rates 256 <= C <= 8192 with both polarizations, radix 16.
//
// Every thread holds 16 points of each polarization, so a spectrum takes
// C/16 threads and a workgroup of NT threads holds NT*16/C spectra.  The
// passes are radix 16 (floor(log2 C / 4) of them) plus one trailing radix 2,
// 4 or 8 pass when log2 C is not a multiple of 4:
//   first   radix 16 on registers loaded straight from HBM: sixteen 16-byte
//           loads per thread (both pols), issued one spectrum ahead so that
//           they are in flight during the LDS passes of the current one;
//   middle  radix 16, LDS -> LDS;
//   last    LDS -> the fftshifted spectrum in HBM (MODE 0), or detector sums
//           in registers (MODE 1: Stokes I; MODE 4: XX*, YY*, Re XY*, Im XY*).
// LDS holds one plane per polarization with one float2 of padding after every
// 16 points: the 16 consecutive points a lane writes in the first pass, and
// the groups of 16 written in the Ns = 16 pass, then fall on distinct banks.
// Twiddles depend only on the thread's position in its spectrum: W, W^2, W^4
// and W^8 of each pass are read from the table once per launch and the other
// powers are products of these, so the loop's only global loads are the next
// spectrum's samples.
// ---------------------------------------------------------------------------

__device__ __forceinline__ int lpad(int pos) { return pos + (pos >> 4); }

// a * W_16^E with W_16 = exp(-j 2 pi / 16); cos(2 pi e / 16) by table,
// -sin(2 pi e / 16) = cos(2 pi (e + 4) / 16).
template <int E>
__device__ __forceinline__ float2 mul_w16(float2 a) {
    constexpr int e = E & 15;
    constexpr float c1 = 0.92387953251128675613f;   // cos(pi/8)
    constexpr float s1 = 0.38268343236508977173f;   // sin(pi/8)
    constexpr float h = 0.70710678118654752440f;    // cos(pi/4)
    constexpr float cs[16] = {1.f, c1, h, s1, 0.f, -s1, -h, -c1, -1.f, -c1, -h, -s1, 0.f, s1, h, c1};
    if constexpr (e == 0) return a;
    else if constexpr (e == 4) return make_float2(a.y, -a.x);
    else if constexpr (e == 8) return make_float2(-a.x, -a.y);
    else if constexpr (e == 12) return make_float2(-a.y, a.x);
    else return cmul(a, make_float2(cs[e], cs[(e + 4) & 15]));
}

// 16-point DFT in natural order: n = 4 n1 + n2, k = k1 + 4 k2,
// X[k1 + 4 k2] = sum_n2 W4^(n2 k2) W16^(n2 k1) sum_n1 W4^(n1 k1) x[4 n1 + n2].
__device__ __forceinline__ void dft16(float2 (&v)[16]) {
    float2 y[4][4];
#pragma unroll
    for (int n2 = 0; n2 < 4; n2++) {
        float2 t[4] = {v[n2], v[n2 + 4], v[n2 + 8], v[n2 + 12]};
        dft_reg<4>(t, nullptr, 0);
#pragma unroll
        for (int k1 = 0; k1 < 4; k1++) y[n2][k1] = t[k1];
    }
    y[1][1] = mul_w16<1>(y[1][1]); y[1][2] = mul_w16<2>(y[1][2]); y[1][3] = mul_w16<3>(y[1][3]);
    y[2][1] = mul_w16<2>(y[2][1]); y[2][2] = mul_w16<4>(y[2][2]); y[2][3] = mul_w16<6>(y[2][3]);
    y[3][1] = mul_w16<3>(y[3][1]); y[3][2] = mul_w16<6>(y[3][2]); y[3][3] = mul_w16<9>(y[3][3]);
#pragma unroll
    for (int k1 = 0; k1 < 4; k1++) {
        float2 t[4] = {y[0][k1], y[1][k1], y[2][k1], y[3][k1]};
        dft_reg<4>(t, nullptr, 0);
#pragma unroll
        for (int k2 = 0; k2 < 4; k2++) v[k1 + 4 * k2] = t[k2];
    }
}

template <int R>
__device__ __forceinline__ void dft_pow2(float2 (&v)[R]) {
    if constexpr (R == 16) dft16(v);
    else dft_reg<R>(v, nullptr, 0);
}

// w[r] = W^r for r < R from b = {W, W^2, W^4, W^8}.
template <int R>
__device__ __forceinline__ void twiddle_powers(const float2 (&b)[4], float2 (&w)[R]) {
    w[0] = make_float2(1.f, 0.f);
    if constexpr (R > 1) w[1] = b[0];
    if constexpr (R > 2) { w[2] = b[1]; w[3] = cmul(b[0], b[1]); }
    if constexpr (R > 4) {
        w[4] = b[2];
#pragma unroll
        for (int r = 5; r < 8; r++) w[r] = cmul(w[r - 4], b[2]);
    }
    if constexpr (R > 8) {
        w[8] = b[3];
#pragma unroll
        for (int r = 9; r < 16; r++) w[r] = cmul(w[r - 8], b[3]);
    }
}

// One radix-16 Stockham pass in LDS, Ns = NS, butterfly j = tl.
template <int C, int NS>
__device__ __forceinline__ void fft16_mid(float2* __restrict__ p0, float2* __restrict__ p1, int tl,
                                          const float2* __restrict__ twl) {
    constexpr int STRIDE = C / 16;
    float2 tb[4];
#pragma unroll
    for (int e = 0; e < 4; e++) tb[e] = twl[e * NS + (tl & (NS - 1))];
    float2 a[16], b[16];
    const int rd = lpad(tl);
#pragma unroll
    for (int r = 0; r < 16; r++) {
        a[r] = p0[rd + r * (STRIDE + STRIDE / 16)];
        b[r] = p1[rd + r * (STRIDE + STRIDE / 16)];
    }
    float2 w[16];
    twiddle_powers<16>(tb, w);
#pragma unroll
    for (int r = 1; r < 16; r++) {
        a[r] = cmul(a[r], w[r]);
        b[r] = cmul(b[r], w[r]);
    }
    dft16(a);
    dft16(b);
    __syncthreads();
    const int jm = tl & (NS - 1);
    const int wr = lpad((tl - jm) * 16 + jm);
#pragma unroll
    for (int r = 0; r < 16; r++) {
        p0[wr + r * (NS + NS / 16)] = a[r];
        p1[wr + r * (NS + NS / 16)] = b[r];
    }
    __syncthreads();
}

// LDS twiddle table of the middle passes: pass i (Ns = 16^(i+1)) holds
// W_{16 Ns}^(e' jm) for e' = 1, 2, 4, 8 and jm < Ns, laid out [e][jm] so that
// lanes with consecutive jm read consecutive entries.
__host__ __device__ constexpr int fft16_twl_offset(int i) { return i == 0 ? 0 : 4 * 16 + (i - 1) * 4 * 256; }

template <int C, int NMID, int I>
__device__ __forceinline__ void fft16_mids(float2* p0, float2* p1, int tl, const float2* twl) {
    if constexpr (I < NMID) {
        constexpr int NS = 1 << (4 * (I + 1));
        fft16_mid<C, NS>(p0, p1, tl, twl + fft16_twl_offset(I));
        fft16_mids<C, NMID, I + 1>(p0, p1, tl, twl);
    }
}

// The last radix-16 pass for C = 16^k * 2 kept in registers: butterfly j
// (Ns = NS) reads LDS and its 16 results stay in a[] and b[], at positions
// (j / NS) * 16 NS + (j mod NS) + r NS.  With lane l of wave w taking
// j = 32 w + (l mod 32) + (l >= 32 ? C/32 : 0), the final radix-2 pass's
// partners (positions p and p + C/2) sit in register r of lanes l and l ^ 32.
template <int C, int NS>
__device__ __forceinline__ void fft16_mid_regs(const float2* __restrict__ p0, const float2* __restrict__ p1,
                                               int j, const float2* __restrict__ twl, float2 (&a)[16],
                                               float2 (&b)[16]) {
    constexpr int STRIDE = C / 16;
    float2 tb[4];
#pragma unroll
    for (int e = 0; e < 4; e++) tb[e] = twl[e * NS + (j & (NS - 1))];
    const int rd = lpad(j);
#pragma unroll
    for (int r = 0; r < 16; r++) {
        a[r] = p0[rd + r * (STRIDE + STRIDE / 16)];
        b[r] = p1[rd + r * (STRIDE + STRIDE / 16)];
    }
    float2 w[16];
    twiddle_powers<16>(tb, w);
#pragma unroll
    for (int r = 1; r < 16; r++) {
        a[r] = cmul(a[r], w[r]);
        b[r] = cmul(b[r], w[r]);
    }
    dft16(a);
    dft16(b);
}

// One polarization's plane of fft16_mid_regs.
template <int C, int NS>
__device__ __forceinline__ void fft16_mid_regs1(const float2* __restrict__ pl, int j, const float2* __restrict__ twl,
                                                float2 (&a)[16]) {
    constexpr int STRIDE = C / 16;
    float2 tb[4];
#pragma unroll
    for (int e = 0; e < 4; e++) tb[e] = twl[e * NS + (j & (NS - 1))];
    const int rd = lpad(j);
#pragma unroll
    for (int r = 0; r < 16; r++) a[r] = pl[rd + r * (STRIDE + STRIDE / 16)];
    float2 w[16];
    twiddle_powers<16>(tb, w);
#pragma unroll
    for (int r = 1; r < 16; r++) a[r] = cmul(a[r], w[r]);
    dft16(a);
}

// W_32^r = exp(-j 2 pi r / 32), r < 16, from cos(k pi / 16).
__device__ __forceinline__ float2 w32(int r) {
    constexpr float c1 = 0.98078528040323044913f, c2 = 0.92387953251128675613f;
    constexpr float c3 = 0.83146961230254523708f, c4 = 0.70710678118654752440f;
    constexpr float c5 = 0.55557023301960222474f, c6 = 0.38268343236508977173f;
    constexpr float c7 = 0.19509032201612826785f;
    constexpr float re[16] = {1.f, c1, c2, c3, c4, c5, c6, c7, 0.f, -c7, -c6, -c5, -c4, -c3, -c2, -c1};
    constexpr float im[16] = {0.f, -c7, -c6, -c5, -c4, -c3, -c2, -c1, -1.f, -c1, -c2, -c3, -c4, -c5, -c6, -c7};
    return make_float2(re[r], im[r]);
}

// x at lanes l < 32 holds u, at lanes l >= 32 holds v (same register):
// returns u and v in every lane.
__device__ __forceinline__ void halves(float x, float& u, float& v) {
    const auto p = __builtin_amdgcn_permlane32_swap(__float_as_uint(x), __float_as_uint(x), false, false);
    u = __uint_as_float(p[0]);
    v = __uint_as_float(p[1]);
}

// Workgroup order with consecutive values on one XCD: the dispatcher deals
// workgroups to the 8 XCDs round-robin (b mod 8), so b -> (b mod 8) G/8 + b/8
// hands workgroups that take consecutive spectra to one XCD, whose L2 then
// merges their interleaved pieces of each output line.  G is a multiple of 8.
__device__ __forceinline__ unsigned xcd_order(unsigned b, unsigned G) { return (b & 7u) * (G >> 3) + (b >> 3); }

// SM (MODE 0 stores): 0 strided, 1 nontemporal with one spectrum per
// (antenna, channel), 2 transposed through LDS so that the workgroup's SPW
// consecutive spectra of one (antenna, channel) land as whole lines.
template <int LC, int NT, int MODE, int SM>
__global__ __launch_bounds__(NT, 2) void fft16_kernel(const float4* __restrict__ in, void* __restrict__ outv,
                                                      uint64_t S, uint64_t n_items, int N, int flags,
                                                      const float2* __restrict__ tw, float4* __restrict__ sink) {
    const int accumulate = flags & 1;
    const bool circular = flags & 2;
    constexpr int C = 1 << LC;
    constexpr int TPS = C / 16;                        // threads per spectrum
    static_assert(NT % TPS == 0 && LC >= 8, "workgroup must hold whole spectra");
    constexpr int SPW = NT / TPS;                      // spectra per workgroup
    constexpr int RL = 1 << (LC % 4);                  // trailing radix, 1 = none
    constexpr int NMID = LC / 4 - (RL > 1 ? 1 : 2);    // radix-16 passes in LDS
    constexpr int R = RL > 1 ? RL : 16;                // radix of the last pass
    constexpr int G = 16 / R;                          // last-pass butterflies per thread
    constexpr int NSL = C / R;                         // Ns (and stride) of the last pass
    constexpr int PLANE = C + C / 16;
    constexpr int NPOL = MODE == 0 ? 1 : MODE;
    // Final radix-2 pass across lane halves (whole waves per spectrum); the
    // full-Stokes sums need both polarizations at once and keep the LDS pass.
    constexpr bool SW = RL == 2 && TPS % 64 == 0 && NMID >= 1 && MODE != 4;
    extern __shared__ float2 lds[];
    const int slot = threadIdx.x / TPS;
    const int tl = threadIdx.x % TPS;
    float2* const p0 = lds + slot * 2 * PLANE;
    float2* const p1 = p0 + PLANE;

    // Twiddle bases.  Middle pass i has Ns = 16^(i+1) and uses
    // W_{16 Ns}^(tl mod Ns) = tw[(tl mod Ns) * C / (16 Ns)] and its powers 2,
    // 4, 8, from an LDS table behind the data planes (filled here, first read
    // after the first pass's barrier); the last pass uses W_C^(tl r) W_16^(g r)
    // for butterfly j = tl + g * C/16, from per-thread registers.
    float2* const twl = lds + SPW * 2 * PLANE;
#pragma unroll
    for (int i = 0; i < NMID; i++) {
        const int ns = 1 << (4 * (i + 1));
        for (int idx = threadIdx.x; idx < 4 * ns; idx += NT) {
            const int e = idx / ns, jm = idx - e * ns;
            twl[fft16_twl_offset(i) + idx] = tw[((jm * (C / (16 * ns))) << e) & (C - 1)];
        }
    }
    float2 tb[4];
#pragma unroll
    for (int e = 0; e < 4; e++) tb[e] = tw[(tl << e) & (C - 1)];
    // SW: lane l of wave w handles jlow = 32 w + (l mod 32) in the upper
    // (l >= 32) or lower half; its final-pass twiddle base is W_C^jlow.
    const bool hi = (tl & 63) >= 32;
    const int jlow = (tl >> 6) * 32 + (tl & 31);
    const float2 tsw = SW ? tw[jlow] : make_float2(1.f, 0.f);
    // Output channel (fftshifted) of result i: lane base + per-result offset.
    const int ch0 = SW ? jlow + (hi ? 0 : C / 2) : tl;
    auto kofs = [](int i) -> int {
        if constexpr (SW) return i * (C / 32);
        else return ((i / R) * TPS + (i % R) * NSL + (C >> 1)) & (C - 1);
    };

    // Strided stores: workgroups of consecutive spectra share an XCD.
    const unsigned wgi = (MODE == 0 && SM == 0) ? xcd_order(blockIdx.x, gridDim.x) : blockIdx.x;
    const uint64_t per_round = (uint64_t)gridDim.x * SPW;
    const uint64_t rounds = (n_items + per_round - 1) / per_round;
    const int nspec = MODE == 0 ? 1 : (int)S;
    const int groups = MODE == 0 ? 1 : ((N <= (int)S) ? (int)S / N : 1);

    // The spectra of a workgroup in one round are consecutive (items first ..
    // first + SPW - 1): addresses are a workgroup-uniform base plus a 32-bit
    // lane offset.  Idle slots re-read their workgroup's first spectrum (or
    // spectrum 0) and store nothing.
    const unsigned SC = (unsigned)(MODE == 0 ? 1 : S) * C;      // float4s per item
    float4 q[16];
    auto load = [&](uint64_t round, int s) {
        const uint64_t first = (round * gridDim.x + wgi) * SPW;
        const bool any = first < n_items;
        const uint64_t b0 = any ? (MODE == 0 ? first * C : (first * S + s) * C) : 0;
        const unsigned lo = ((any && first + slot < n_items) ? slot * SC : 0u) + tl;
#pragma unroll
        for (int r = 0; r < 16; r++) q[r] = (in + (b0 + r * TPS))[lo];
    };

    float acc[16][NPOL];
#pragma unroll
    for (int i = 0; i < 16; i++)
#pragma unroll
        for (int c = 0; c < NPOL; c++) acc[i][c] = 0.f;

    // The full-Stokes detector keeps 64 sums per thread, which leaves no
    // room for the next spectrum's registers: it loads at the top instead.
    constexpr bool PF = MODE != 4;
    if constexpr (PF) {
        load(0, 0);
        // Nothing pending at loop entry, so the first pass of every spectrum
        // waits only for its own loads, not for the previous spectrum's stores.
        __builtin_amdgcn_s_waitcnt(0);
    }
    for (uint64_t round = 0; round < rounds; round++) {
        const uint64_t first = (round * gridDim.x + wgi) * SPW;
        const uint64_t item = first + slot;
        const bool valid = item < n_items;
        for (int s = 0; s < nspec; s++) {
            if constexpr (!PF) load(round, s);
            // ---- first pass: radix 16 on registers -> LDS (Ns = 1)
            {
                float2 a[16], b[16];
#pragma unroll
                for (int r = 0; r < 16; r++) {
                    a[r] = make_float2(q[r].x, q[r].y);
                    b[r] = make_float2(q[r].z, q[r].w);
                }
                dft16(a);
                dft16(b);
                const int wr = tl * 17;
#pragma unroll
                for (int r = 0; r < 16; r++) {
                    p0[wr + r] = a[r];
                    p1[wr + r] = b[r];
                }
                // Next spectrum into the registers just consumed (a fresh set
                // would be copied back at the loop latch behind a full wait).
                // After the last spectrum this re-reads spectrum 0 (an L2 hit)
                // rather than branching, which keeps the waits exact.
                if constexpr (PF) {
                    uint64_t nr = round;
                    int ns = s + 1;
                    if (ns == nspec) { ns = 0; nr++; }
                    load(nr, ns);
                }
            }
            __syncthreads();
            float2 a[16], b[16];
            if constexpr (SW) {
                // ---- last radix-16 pass in registers, then radix 2 across
                // lane halves: X(p) = u + W v in the lower half, X(p + C/2) =
                // u - W v in the upper, W = W_C^(jlow + r C/32).
                fft16_mids<C, NMID - 1, 0>(p0, p1, tl, twl);
                constexpr int NSR = 1 << (4 * NMID);
                const int jr = jlow + (hi ? C / 32 : 0);
                const float2* const twr = twl + fft16_twl_offset(NMID - 1);
                const float sg = hi ? -1.f : 1.f;
                auto radix2 = [&](float2& x, const float2 wr) {
                    float2 u, v;
                    halves(x.x, u.x, v.x);
                    halves(x.y, u.y, v.y);
                    const float2 t = cmul(wr, v);
                    x = make_float2(u.x + sg * t.x, u.y + sg * t.y);
                };
                // W_32^(r + 8) = -j W_32^r: one twiddle product serves
                // results r and r + 8.
                auto negj = [](float2 w) { return make_float2(w.y, -w.x); };
                if constexpr (MODE == 1) {
                    // Stokes I: one polarization at a time (half the registers).
#pragma unroll
                    for (int q = 0; q < 2; q++) {
                        fft16_mid_regs1<C, NSR>(q == 0 ? p0 : p1, jr, twr, a);
#pragma unroll
                        for (int r = 0; r < 8; r++) {
                            const float2 wr = cmul(tsw, w32(r));
                            radix2(a[r], wr);
                            radix2(a[r + 8], negj(wr));
                            acc[r][0] += a[r].x * a[r].x + a[r].y * a[r].y;
                            acc[r + 8][0] += a[r + 8].x * a[r + 8].x + a[r + 8].y * a[r + 8].y;
                        }
                    }
                } else {
                    fft16_mid_regs<C, NSR>(p0, p1, jr, twr, a, b);
#pragma unroll
                    for (int r = 0; r < 8; r++) {
                        const float2 wr = cmul(tsw, w32(r)), wr8 = negj(wr);
                        radix2(a[r], wr);
                        radix2(b[r], wr);
                        radix2(a[r + 8], wr8);
                        radix2(b[r + 8], wr8);
                    }
                }
            } else {
            fft16_mids<C, NMID, 0>(p0, p1, tl, twl);
            // ---- last pass: radix R, LDS -> output (Ns = C / R)
            const int rd = lpad(tl);
#pragma unroll
            for (int g = 0; g < G; g++)
#pragma unroll
                for (int r = 0; r < R; r++) {
                    const int off = (g * TPS + r * NSL) + (g * TPS + r * NSL) / 16;
                    a[g * R + r] = p0[rd + off];
                    b[g * R + r] = p1[rd + off];
                }
            {
                float2 w[R];
                twiddle_powers<R>(tb, w);
#pragma unroll
                for (int g = 0; g < G; g++) {
#pragma unroll
                    for (int r = 1; r < R; r++) {
                        float2 wr = w[r];
                        switch ((g * r) & 15) {     // W_16^(g r), resolved at compile time
                            case 0: break;
                            case 1: wr = mul_w16<1>(wr); break;
                            case 2: wr = mul_w16<2>(wr); break;
                            case 3: wr = mul_w16<3>(wr); break;
                            case 4: wr = mul_w16<4>(wr); break;
                            case 5: wr = mul_w16<5>(wr); break;
                            case 6: wr = mul_w16<6>(wr); break;
                            case 7: wr = mul_w16<7>(wr); break;
                            case 8: wr = mul_w16<8>(wr); break;
                            case 9: wr = mul_w16<9>(wr); break;
                            case 10: wr = mul_w16<10>(wr); break;
                            case 11: wr = mul_w16<11>(wr); break;
                            case 12: wr = mul_w16<12>(wr); break;
                            case 13: wr = mul_w16<13>(wr); break;
                            case 14: wr = mul_w16<14>(wr); break;
                            default: wr = mul_w16<15>(wr); break;
                        }
                        a[g * R + r] = cmul(a[g * R + r], wr);
                        b[g * R + r] = cmul(b[g * R + r], wr);
                    }
                    float2 xa[R], xb[R];
#pragma unroll
                    for (int r = 0; r < R; r++) { xa[r] = a[g * R + r]; xb[r] = b[g * R + r]; }
                    dft_pow2<R>(xa);
                    dft_pow2<R>(xb);
#pragma unroll
                    for (int r = 0; r < R; r++) { a[g * R + r] = xa[r]; b[g * R + r] = xb[r]; }
                }
            }
            }
            // Channel of result i (fftshifted): ch0 + kofs(i).  Without SW,
            // k = tl + K with K = g C/16 + r C/R and ko = tl + ((K + C/2) mod
            // C), the bracket a multiple of C/16.
            // Idle lanes store into the sink (same store count on every lane
            // and path, so waits on the next spectrum's loads stay exact).
            if constexpr (MODE == 0 && SM == 2) {
                // S is a multiple of SPW: the workgroup's spectra are
                // si0 .. si0 + SPW - 1 of one (antenna, channel), whose outputs
                // interleave at (ko S + si).  Stage them [ko][spectrum] in LDS
                // (one float4 of padding per 16) and store runs of SPW.
                __syncthreads();                    // the last pass's LDS reads are done
                float4* const L = reinterpret_cast<float4*>(lds);
                int base = ch0 * SPW + slot;
                asm volatile("" : "+v"(base));      // not hoisted: see below
#pragma unroll
                for (int i = 0; i < 16; i++) {
                    const int idx = base + kofs(i) * SPW;
                    L[idx + (idx >> 4)] = make_float4(a[i].x, a[i].y, b[i].x, b[i].y);
                }
                __syncthreads();
                // Thread t stores entries t + i NT, i < 16: spectrum t mod SPW,
                // channel t / SPW + i NT / SPW (NT is a multiple of SPW).
                const bool wg_valid = first < n_items;
                const uint64_t af0 = first / S, si0 = first - af0 * S;
                const int tx = threadIdx.x;
                float4* const lb = wg_valid ? static_cast<float4*>(outv) + af0 * C * S + si0 +
                                                  (uint64_t)(tx / SPW) * S + (tx % SPW)
                                            : sink + tx;
                const uint64_t step = wg_valid ? (uint64_t)(NT / SPW) * S : 0;
                const int rb = tx + (tx >> 4);
#pragma unroll
                for (int i = 0; i < 16; i++) {
                    const float4 v = L[rb + i * (NT + NT / 16)];
                    float* const of = reinterpret_cast<float*>(lb + i * step);
                    __builtin_nontemporal_store(v.x, of);
                    __builtin_nontemporal_store(v.y, of + 1);
                    __builtin_nontemporal_store(v.z, of + 2);
                    __builtin_nontemporal_store(v.w, of + 3);
                }
            } else if constexpr (MODE == 0) {
                const uint64_t af = item / S, si = item - af * S;
                float4* const lb = valid ? static_cast<float4*>(outv) + (af * C * S + si) + (uint64_t)ch0 * S
                                         : sink + threadIdx.x;
                const uint64_t step = valid ? S : 0;
#pragma unroll
                for (int g = 0; g < G; g++)
#pragma unroll
                    for (int r = 0; r < R; r++) {
                        const int kb = kofs(g * R + r);
                        const float2 X = a[g * R + r], Y = b[g * R + r];
                        float4* const o = lb + (uint64_t)kb * step;
                        if constexpr (SM == 1) {
                            // Component-wise: merged into one nontemporal
                            // global_store_dwordx4 (a vector operand here is
                            // lowered through private memory).
                            float* const of = reinterpret_cast<float*>(o);
                            __builtin_nontemporal_store(X.x, of);
                            __builtin_nontemporal_store(X.y, of + 1);
                            __builtin_nontemporal_store(Y.x, of + 2);
                            __builtin_nontemporal_store(Y.y, of + 3);
                        } else {
                            *o = make_float4(X.x, X.y, Y.x, Y.y);
                        }
                    }
            } else {
#pragma unroll
                for (int i = 0; i < 16; i++) {
                    if constexpr (SW && MODE == 1) break;     // summed in the register pass
                    const float2 X = a[i], Y = b[i];
                    if constexpr (MODE == 1) {
                        acc[i][0] += X.x * X.x + X.y * X.y + Y.x * Y.x + Y.y * Y.y;
                    } else {
                        acc[i][0] += X.x * X.x + X.y * X.y;
                        acc[i][NPOL > 1 ? 1 : 0] += Y.x * Y.x + Y.y * Y.y;
                        acc[i][NPOL > 2 ? 2 : 0] += X.x * Y.x + X.y * Y.y;
                        acc[i][NPOL > 3 ? 3 : 0] += X.y * Y.x - X.x * Y.y;
                    }
                }
                const bool flush = (N <= (int)S) ? ((s + 1) % N == 0) : (s == (int)S - 1);
                if (flush) {
                    const int grp = (N <= (int)S) ? s / N : 0;
                    if (circular) {
#pragma unroll
                        for (int i = 0; i < 16; i++) circular_sums<NPOL>(acc[i]);
                    }
                    const unsigned gn = (unsigned)groups * NPOL;
                    float* const lb = valid ? static_cast<float*>(outv) + (item * C + ch0) * gn + grp * NPOL
                                            : reinterpret_cast<float*>(sink + threadIdx.x);
                    unsigned step = valid ? gn : 0;
                    // Opaque: keeps the 16 output addresses from being hoisted
                    // out of the spectrum loop (they would pin 32 registers).
                    asm volatile("" : "+v"(step));
                    // All earlier sums are read before any store (idle lanes'
                    // addresses coincide, so the compiler cannot interleave).
                    // The four full-Stokes sums of a channel are one 16-byte
                    // access (the output rows are 16-byte aligned).
                    if (accumulate) {
                        float old[16][NPOL];
#pragma unroll
                        for (int g = 0; g < G; g++)
#pragma unroll
                            for (int r = 0; r < R; r++) {
                                const int kb = kofs(g * R + r);
                                const float* const o = lb + (uint64_t)((unsigned)kb * step);
                                if constexpr (NPOL == 4) {
                                    const float4 v = *reinterpret_cast<const float4*>(o);
                                    old[g * R + r][0] = v.x;
                                    old[g * R + r][NPOL > 1 ? 1 : 0] = v.y;
                                    old[g * R + r][NPOL > 2 ? 2 : 0] = v.z;
                                    old[g * R + r][NPOL > 3 ? 3 : 0] = v.w;
                                } else {
                                    old[g * R + r][0] = o[0];
                                }
                            }
#pragma unroll
                        for (int i = 0; i < 16; i++)
#pragma unroll
                            for (int c = 0; c < NPOL; c++) acc[i][c] += old[i][c];
                    }
#pragma unroll
                    for (int g = 0; g < G; g++)
#pragma unroll
                        for (int r = 0; r < R; r++) {
                            const int kb = kofs(g * R + r);
                            float* const o = lb + (uint64_t)((unsigned)kb * step);
                            const int i = g * R + r;
                            if constexpr (NPOL == 4)
                                *reinterpret_cast<float4*>(o) = make_float4(acc[i][0], acc[i][NPOL > 1 ? 1 : 0],
                                                                            acc[i][NPOL > 2 ? 2 : 0],
                                                                            acc[i][NPOL > 3 ? 3 : 0]);
                            else
                                o[0] = acc[i][0];
                        }
#pragma unroll
                    for (int i = 0; i < 16; i++)
#pragma unroll
                        for (int c = 0; c < NPOL; c++) acc[i][c] = 0.f;
                }
            }
            __syncthreads();   // LDS is rewritten by the next spectrum's first pass
        }
    }
}

// Launch of fft16_kernel for C = 2^LC; NT threads use NT * 272 bytes of LDS
// (one 512-thread workgroup per CU).
template <int LC, int NT, int MODE, int SM>
static int launch_fft16_t(const float2* in, void* out, uint64_t S, uint64_t n_items, int N, int accumulate,
                          const float2* tw, hipStream_t s) {
    constexpr int C = 1 << LC;
    constexpr int SPW = NT / (C / 16);
    constexpr int NMID = LC / 4 - ((LC % 4) ? 1 : 2);
    constexpr size_t lds = sizeof(float2) * (2 * (size_t)(C + C / 16) * SPW + fft16_twl_offset(NMID));
    const void* fn = reinterpret_cast<const void*>(&fft16_kernel<LC, NT, MODE, SM>);
    if (lds > 65536) {
        hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) {
            set_error(std::string("channelizer: LDS attribute: ") + hipGetErrorString(e));
            return BLADE_HIP_ERROR;
        }
    }
    const uint64_t nb = ceil_div(n_items, SPW);
    const uint64_t resident = (uint64_t)device_cus() * (NT >= 512 ? 1 : 2);
    const uint64_t rounds = ceil_div(nb, resident);
    uint64_t blocks = ceil_div(nb, rounds);            // equal rounds for every workgroup
    if (MODE == 0 && SM == 0) blocks = ceil_div(blocks, 8) * 8;   // xcd_order needs G % 8 == 0
    float4* sink = reinterpret_cast<float4*>(const_cast<float2*>(tw) + C);
    hipLaunchKernelGGL((fft16_kernel<LC, NT, MODE, SM>), dim3((unsigned)blocks), dim3(NT), lds, s,
                       reinterpret_cast<const float4*>(in), out, S, n_items, N, accumulate, tw, sink);
    return check_launch("channelizer");
}

// fft16_kernel covers the power-of-two rates 256 <= C <= 8192.
static bool use_fft16(int C) { return C >= 256 && C <= 8192 && !(C & (C - 1)); }

template <int LC, int MODE>
static int launch_fft16_lc(const float2* in, void* out, uint64_t S, uint64_t n_items, int N, int accumulate,
                           const float2* tw, hipStream_t s) {
    // Nontemporal stores pay off only where the stores are whole lines: one
    // spectrum per (antenna, channel), or several staged through LDS when a
    // workgroup holds whole runs of at least four of them (S a multiple of
    // its SPW spectra, so each row of the staged tile is 64 contiguous
    // bytes); with two per row the LDS round trip costs more than the 32-byte
    // pieces save, and interleaved partial lines with nontemporal stores are
    // much slower than plain stores.
    if constexpr (MODE == 0) {
        constexpr uint64_t SPW = 512 / ((1 << LC) / 16);
        if (S == 1) return launch_fft16_t<LC, 512, MODE, 1>(in, out, S, n_items, N, accumulate, tw, s);
        if constexpr (SPW >= 4) {
            if (S % SPW == 0) return launch_fft16_t<LC, 512, MODE, 2>(in, out, S, n_items, N, accumulate, tw, s);
        }
    }
    return launch_fft16_t<LC, 512, MODE, 0>(in, out, S, n_items, N, accumulate, tw, s);
}

constexpr uint64_t NT_MAX_SPECTRA = 32;    // spectra per workgroup at C = 256, 512 threads

// Returns 1 and launches when fft16_kernel covers C, else 0.
template <int MODE>
static int launch_fft16(int C, const float2* in, void* out, uint64_t S, uint64_t n_items, int N, int accumulate,
                        const float2* tw, hipStream_t s, int* rc) {
    // Lane offsets within a workgroup's spectra are 32-bit.
    if ((uint64_t)(NT_MAX_SPECTRA + 2) * S * (uint64_t)C >= (1ull << 31)) return 0;
    switch (C) {
        case 1 << 8: *rc = launch_fft16_lc<8, MODE>(in, out, S, n_items, N, accumulate, tw, s); return 1;
        case 1 << 9: *rc = launch_fft16_lc<9, MODE>(in, out, S, n_items, N, accumulate, tw, s); return 1;
        case 1 << 10: *rc = launch_fft16_lc<10, MODE>(in, out, S, n_items, N, accumulate, tw, s); return 1;
        case 1 << 11: *rc = launch_fft16_lc<11, MODE>(in, out, S, n_items, N, accumulate, tw, s); return 1;
        case 1 << 12: *rc = launch_fft16_lc<12, MODE>(in, out, S, n_items, N, accumulate, tw, s); return 1;
        case 1 << 13: *rc = launch_fft16_lc<13, MODE>(in, out, S, n_items, N, accumulate, tw, s); return 1;
    }
    return 0;
}

// Direct O(C^2) DFT for rates with prime factors above 7.
template <int P>
__global__ __launch_bounds__(256) void chan_direct_kernel(const float2* __restrict__ in,
                                                          float2* __restrict__ out, int C,
                                                          uint64_t S, uint64_t n_spec,
                                                          const float2* __restrict__ tw) {
    extern __shared__ float2 data[];
    for (uint64_t sp = blockIdx.x; sp < n_spec; sp += gridDim.x) {
        const uint64_t af = sp / S, s = sp - af * S;
        const float2* src = in + sp * (uint64_t)C * P;
        for (int i = threadIdx.x; i < C * P; i += blockDim.x) {
            const int n = i / P, p = i - n * P;
            data[p * C + n] = src[i];
        }
        __syncthreads();
        const int half = C / 2;
        for (int i = threadIdx.x; i < C * P; i += blockDim.x) {
            const int k = i / P, p = i - k * P;
            float2 acc = make_float2(0.f, 0.f);
            int m = 0;
            for (int n = 0; n < C; n++) {
                const float2 x = data[p * C + n], w = tw[m];
                acc.x += x.x * w.x - x.y * w.y;
                acc.y += x.x * w.y + x.y * w.x;
                m += k;
                if (m >= C) m -= C;
            }
            int ko = k + half;
            if (ko >= C) ko -= C;
            out[((af * C + ko) * S + s) * P + p] = acc;
        }
        __syncthreads();
    }
}

static bool factorize(int C, Radices* rad) {
    rad->n = 0;
    int c = C;
    // Prefer radix 8, then 4, then one radix 2; then 3, 5, 7.
    while (c % 8 == 0 && rad->n < 16) { rad->r[rad->n++] = 8; c /= 8; }
    while (c % 4 == 0 && rad->n < 16) { rad->r[rad->n++] = 4; c /= 4; }
    while (c % 2 == 0 && rad->n < 16) { rad->r[rad->n++] = 2; c /= 2; }
    while (c % 3 == 0 && rad->n < 16) { rad->r[rad->n++] = 3; c /= 3; }
    while (c % 5 == 0 && rad->n < 16) { rad->r[rad->n++] = 5; c /= 5; }
    while (c % 7 == 0 && rad->n < 16) { rad->r[rad->n++] = 7; c /= 7; }
    return c == 1;
}

// Split C = N1 * N2 for the two-pass FFT: N2 near sqrt(C), N1 even, each
// tile of the passes within 16384 samples of LDS (with the padding sample);
// tiles of up to 16 consecutive columns that divide the other length.
struct BigPlan {
    int N1, N2, cols1, cols2;
    Radices r1, r2;
};

constexpr int kMaxBigRate = 1 << 20;     // the pass-1 twiddle table: 8 MiB

static int big_cols(int n_other, int P, int len) {
    int cols = 64;
    while (cols > 1 && (P * cols * (len + 1) > kMaxLdsComplex || n_other % cols)) cols /= 2;
    return P * cols * (len + 1) <= kMaxLdsComplex ? cols : 0;
}

static bool big_plan(int C, int P, BigPlan* pl) {
    if (C > kMaxBigRate || C % 2) return false;
    double best = 1e30;
    bool found = false;
    for (int n2 = 2; n2 < C; n2++) {
        if (C % n2) continue;
        const int n1 = C / n2;
        if (n1 % 2) continue;
        const int c1 = big_cols(n1, P, n2), c2 = big_cols(n2, P, n1);
        if (!c1 || !c2) continue;
        BigPlan t;
        if (!factorize(n1, &t.r1) || !factorize(n2, &t.r2)) continue;
        const double score = std::fabs(std::log((double)n2) - 0.5 * std::log((double)C)) - 0.01 * (c1 + c2);
        if (score < best) {
            best = score;
            t.N1 = n1; t.N2 = n2; t.cols1 = c1; t.cols2 = c2;
            *pl = t;
            found = true;
        }
    }
    return found;
}

template <int P>
static int launch_big(const BigPlan& pl, int C, const float2* in, float2* out, uint64_t S, uint64_t n_spec,
                      hipStream_t s) {
    const float2 *twC = nullptr, *tw1 = nullptr, *tw2 = nullptr;
    int rc = get_twiddles(C, s, &twC);
    if (!rc) rc = get_twiddles(pl.N1, s, &tw1);
    if (!rc) rc = get_twiddles(pl.N2, s, &tw2);
    if (rc) return rc;
    unsigned char* scratch = nullptr;
    rc = library_workspace(WS_CHANNELIZER_BIG, sizeof(float2) * (size_t)n_spec * C * P, s, &scratch);
    if (rc) return rc;
    float2* y = reinterpret_cast<float2*>(scratch);
    const size_t lds1 = sizeof(float2) * (size_t)P * pl.cols1 * (pl.N2 + 1);
    const size_t lds2 = sizeof(float2) * (size_t)P * pl.cols2 * (pl.N1 + 1);
    const void* f1 = reinterpret_cast<const void*>(&chan_big_pass1<P>);
    const void* f2 = reinterpret_cast<const void*>(&chan_big_pass2<P>);
    if (hipFuncSetAttribute(f1, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds1) != hipSuccess ||
        hipFuncSetAttribute(f2, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds2) != hipSuccess)
        return fail("channelizer: LDS attribute of the two-pass FFT");
    const uint64_t t1 = n_spec * (uint64_t)(pl.N1 / pl.cols1), t2 = n_spec * (uint64_t)(pl.N2 / pl.cols2);
    const unsigned g1 = (unsigned)std::min<uint64_t>(t1, 256 * 64), g2 = (unsigned)std::min<uint64_t>(t2, 256 * 64);
    hipLaunchKernelGGL((chan_big_pass1<P>), dim3(g1), dim3(1024), lds1, s, in, y, C, pl.N1, pl.N2, pl.cols1,
                       n_spec, tw2, twC, pl.r2);
    rc = check_launch("channelizer pass 1");
    if (rc) return rc;
    hipLaunchKernelGGL((chan_big_pass2<P>), dim3(g2), dim3(1024), lds2, s, y, out, C, pl.N1, pl.N2, pl.cols2,
                       n_spec, S, tw1, pl.r1);
    return check_launch("channelizer pass 2");
}

template <int C, int P>
static void launch_small(const float2* in, float2* out, uint64_t n_spec, uint64_t S, hipStream_t s) {
    SmallTw<C> tw;
    for (int m = 0; m < C; m++) {
        const double a = -2.0 * M_PI * (double)m / (double)C;
        tw.w[m] = make_float2((float)cos(a), (float)sin(a));
    }
    // One workgroup per 256 spectra (a compact front streams faster than a
    // capped grid-stride loop); the loop covers sizes past the cap.
    uint64_t blocks = ceil_div(n_spec, 256);
    if (blocks > 0x7fffffffull / 256) blocks = 0x7fffffffull / 256;
    hipLaunchKernelGGL((chan_small_kernel<C, P>), dim3((unsigned)blocks), dim3(256), 0, s, in, out,
                       n_spec, S, tw);
}

template <int P>
static int dispatch_small(int C, const float2* in, float2* out, uint64_t n_spec, uint64_t S,
                          hipStream_t s) {
    switch (C) {
        case 2: launch_small<2, P>(in, out, n_spec, S, s); return 1;
        case 4: launch_small<4, P>(in, out, n_spec, S, s); return 1;
        case 6: launch_small<6, P>(in, out, n_spec, S, s); return 1;
        case 8: launch_small<8, P>(in, out, n_spec, S, s); return 1;
        case 10: launch_small<10, P>(in, out, n_spec, S, s); return 1;
        case 12: launch_small<12, P>(in, out, n_spec, S, s); return 1;
        case 14: launch_small<14, P>(in, out, n_spec, S, s); return 1;
        case 16: launch_small<16, P>(in, out, n_spec, S, s); return 1;
    }
    return 0;
}

}  // namespace bl

using namespace bl;

extern "C" int blade_channelizer(const void* input, void* output, uint64_t A, uint64_t F,
                                 uint64_t T, uint64_t P, uint64_t rate, blade_stream_t stream) {
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    // Validation mirrors src/modules/channelizer/base.cc:17-42 except that any
    // rate dividing T is accepted (the oracle's general semantics).
    if (rate == 0 || T % rate) return fail("channelizer: the number of time samples should be divisible by the channelizer rate");
    if (rate % 2 && rate != 1) return fail("channelizer: the channelizer rate should be divisible by 2");
    if (P != 1 && P != 2) return fail("channelizer: number of polarizations of the input should be one or two");
    const uint64_t n = A * F * T * P;
    if (n == 0) return BLADE_SUCCESS;
    if (!input || !output) return fail("channelizer: null pointer");
    if ((uintptr_t)input % 16 || (uintptr_t)output % 16) return fail("channelizer: misaligned buffer");
    const int C = (int)rate;
    const uint64_t S = T / rate;
    const uint64_t n_spec = A * F * S;
    if (C == 1) {
        if (input == output) return BLADE_SUCCESS;
        hipError_t e = hipMemcpyAsync(output, input, n * 8, hipMemcpyDeviceToDevice, s);
        if (e != hipSuccess) {
            set_error(std::string("channelizer copy: ") + hipGetErrorString(e));
            return BLADE_HIP_ERROR;
        }
        return BLADE_SUCCESS;
    }
    const float2* in = static_cast<const float2*>(input);
    float2* out = static_cast<float2*>(output);
    if ((uint64_t)C * P > (uint64_t)kMaxLdsComplex) {
        // Past one workgroup's LDS: two passes through a scratch buffer (which
        // also makes any in-place call safe).
        BigPlan pl;
        if (!big_plan(C, (int)P, &pl))
            return fail("channelizer: rate must be even, at most 2^20, and factor into 2, 3, 5 and 7");
        return P == 2 ? launch_big<2>(pl, C, in, out, S, n_spec, s) : launch_big<1>(pl, C, in, out, S, n_spec, s);
    }
    if (input == output && S != 1)
        return fail("channelizer: in-place operation needs rate == number of time samples");
    if (C <= 16 && input != output) {
        const int ok = (P == 2) ? dispatch_small<2>(C, in, out, n_spec, S, s)
                                : dispatch_small<1>(C, in, out, n_spec, S, s);
        if (ok) return check_launch("channelizer");
    }
    const float2* tw = nullptr;
    int rc = get_twiddles(C, s, &tw);
    if (rc) return rc;
    if (P == 2 && use_fft16(C)) {
        int rc16 = 0;
        if (launch_fft16<0>(C, in, out, S, n_spec, 1, 0, tw, s, &rc16)) return rc16;
    }
    Fft2Plan pl;
    if (P == 2 && fft2_plan(C, &pl))
        return launch_fft2<0>(pl, in, out, C, S, n_spec, 1, 0, tw, s);
    Radices rad;
    const size_t lds = sizeof(float2) * (size_t)C * P;
    uint64_t blocks = n_spec;
    if (blocks > 256 * 64) blocks = 256 * 64;
    const void* fn = nullptr;
    int threads = 256;
    if (factorize(C, &rad)) {
        threads = 64;
        while (threads < 1024 && threads * 8 < C * (int)P) threads *= 2;
        fn = (P == 2) ? reinterpret_cast<const void*>(&chan_lds_kernel<2>)
                      : reinterpret_cast<const void*>(&chan_lds_kernel<1>);
    } else {
        fn = (P == 2) ? reinterpret_cast<const void*>(&chan_direct_kernel<2>)
                      : reinterpret_cast<const void*>(&chan_direct_kernel<1>);
    }
    if (lds > 65536) {
        hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) {
            set_error(std::string("channelizer: LDS attribute: ") + hipGetErrorString(e));
            return BLADE_HIP_ERROR;
        }
    }
    if (fn == reinterpret_cast<const void*>(&chan_lds_kernel<2>))
        hipLaunchKernelGGL((chan_lds_kernel<2>), dim3((unsigned)blocks), dim3(threads), lds, s,
                           in, out, C, S, n_spec, tw, rad);
    else if (fn == reinterpret_cast<const void*>(&chan_lds_kernel<1>))
        hipLaunchKernelGGL((chan_lds_kernel<1>), dim3((unsigned)blocks), dim3(threads), lds, s,
                           in, out, C, S, n_spec, tw, rad);
    else if (P == 2)
        hipLaunchKernelGGL((chan_direct_kernel<2>), dim3((unsigned)blocks), dim3(256), lds, s,
                           in, out, C, S, n_spec, tw);
    else
        hipLaunchKernelGGL((chan_direct_kernel<1>), dim3((unsigned)blocks), dim3(256), lds, s,
                           in, out, C, S, n_spec, tw);
    return check_launch("channelizer");
}

extern "C" int blade_channelizer_plan(uint64_t rate, blade_stream_t stream) {
    if (rate == 0 || (rate % 2 && rate != 1)) return fail("channelizer: invalid rate");
    if (rate > (uint64_t)kMaxBigRate) return fail("channelizer: rate above 2^20 is not supported");
    if (rate <= 16) return BLADE_SUCCESS;   // register kernels carry their twiddles as arguments
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    const float2* tw = nullptr;
    int rc = get_twiddles((int)rate, s, &tw);
    if (rc) return rc;
    BigPlan pl;   // the two-pass FFT's tables (for two polarizations)
    if (rate * 2 > (uint64_t)kMaxLdsComplex && big_plan((int)rate, 2, &pl)) {
        rc = get_twiddles(pl.N1, s, &tw);
        if (!rc) rc = get_twiddles(pl.N2, s, &tw);
    }
    return rc;
}

extern "C" int blade_channelizer_detector(const void* input, void* output, uint64_t A, uint64_t F,
                                          uint64_t T, uint64_t rate, uint64_t integration,
                                          uint64_t out_pols, int polarize, int accumulate,
                                          blade_stream_t stream) {
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    if (rate == 0 || T % rate) return fail("channelizer-detector: T must be divisible by the rate");
    if (rate % 2 && rate != 1) return fail("channelizer-detector: the rate should be divisible by 2");
    if (out_pols != 1 && out_pols != 4) return fail("channelizer-detector: output polarizations must be 1 or 4");
    const uint64_t S = T / rate;
    if (integration == 0) return fail("channelizer-detector: integration size should be more than zero");
    if (2 * rate > (uint64_t)kMaxLdsComplex) return fail("channelizer-detector: rate above 8192 is not supported");
    if (A * F * T == 0) return BLADE_SUCCESS;     // before S is used as a divisor
    if (integration <= S ? (S % integration != 0) : (integration % S != 0))
        return fail("channelizer-detector: integration must divide, or be a multiple of, the spectra per block");
    if (!input || !output) return fail("channelizer-detector: null pointer");
    if ((uintptr_t)input % 16 || (uintptr_t)output % 16) return fail("channelizer-detector: misaligned buffer");
    const int C = (int)rate;
    // bit 0: add into the output (stepped integration); bit 1: XY -> LR first
    const int flags = (accumulate ? 1 : 0) | (polarize ? 2 : 0);
    Radices rad;
    if (!factorize(C, &rad)) return fail("channelizer-detector: rate must factor into 2, 3, 5 and 7");
    const float2* tw = nullptr;
    int rc = get_twiddles(C, s, &tw);
    if (rc) return rc;
    if (use_fft16(C)) {
        int rc16 = 0;
        const float2* in16 = static_cast<const float2*>(input);
        if (out_pols == 1 ? launch_fft16<1>(C, in16, output, S, A * F, (int)integration, flags, tw, s, &rc16)
                          : launch_fft16<4>(C, in16, output, S, A * F, (int)integration, flags, tw, s, &rc16))
            return rc16;
    }
    Fft2Plan pl;
    if (fft2_plan(C, &pl)) {
        if (out_pols == 1)
            return launch_fft2<1>(pl, static_cast<const float2*>(input), output, C, S, A * F, (int)integration,
                                  flags, tw, s);
        return launch_fft2<4>(pl, static_cast<const float2*>(input), output, C, S, A * F, (int)integration,
                              flags, tw, s);
    }
    int threads = 64;
    while (threads < 1024 && threads * 8 < 2 * C) threads *= 2;
    const size_t lds = sizeof(float2) * 2 * (size_t)C;
    const uint64_t n_af = A * F;
    uint64_t blocks = n_af;
    if (blocks > 256 * 64) blocks = 256 * 64;
    const void* fn = out_pols == 1 ? reinterpret_cast<const void*>(&chan_detect_kernel<1, 8>)
                                   : reinterpret_cast<const void*>(&chan_detect_kernel<4, 8>);
    if (lds > 65536) {
        hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) {
            set_error(std::string("channelizer-detector: LDS attribute: ") + hipGetErrorString(e));
            return BLADE_HIP_ERROR;
        }
    }
    const float2* in = static_cast<const float2*>(input);
    float* out = static_cast<float*>(output);
    if (out_pols == 1)
        hipLaunchKernelGGL((chan_detect_kernel<1, 8>), dim3((unsigned)blocks), dim3(threads), lds, s, in, out, C,
                           (int)S, (int)integration, n_af, tw, rad, flags);
    else
        hipLaunchKernelGGL((chan_detect_kernel<4, 8>), dim3((unsigned)blocks), dim3(threads), lds, s, in, out, C,
                           (int)S, (int)integration, n_af, tw, rad, flags);
    return check_launch("channelizer-detector");
}
