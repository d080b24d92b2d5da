// FFT channelizer:  out[a, f*C + k, s, p] = fftshift(DFT_C(in[a, f, s*C:(s+1)*C, p]))[k]
// (oracle tests/modules/channelizer/advanced.py:49-65; forward, unnormalised,
// e^{-j 2pi k n / C}).  Written from the Stockham autosort formulation of the
// mixed-radix FFT (each pass reads stride C/R, twiddles by W_{Ns R}^{(j mod Ns) r},
// butterflies, writes to (j/Ns)*Ns*R + j mod Ns + r*Ns), so the output lands
// in natural order without a bit-reversal pass.  fftshift is an index remap on
// the store.  Twiddles come from a per-C table built in FP64 and rounded
// once to FP32.
#include "common.hh"

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
    hipError_t e = hipMalloc(&p, sizeof(float2) * C);
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

template <int R>
__device__ __forceinline__ void stockham_pass(float2* __restrict__ data, int C, int PC, int Ns,
                                              const float2* __restrict__ tw) {
    constexpr int G = groups_max<R>();
    const int nthr = blockDim.x;
    const int per_pol = C / R;
    const int total = PC / R;
    const int twstride = C / (Ns * R);
    float2 v[G][R];
#pragma unroll
    for (int g = 0; g < G; g++) {
        const int gi = threadIdx.x + g * nthr;
        if (gi < total) {
            const int p = gi / per_pol, j = gi - p * per_pol;
            const float2* base = data + p * C;
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
            float2* base = data + p * C;
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

// Channelizer with the detector fused on the output: one workgroup per
// (a, f) walks its S = T/C spectra, transforms both pols in LDS and
// integrates |X|^2+|Y|^2 (or XX*, YY*, Re XY*, Im XY*) of every fine channel
// over N consecutive spectra in registers, writing
//   out[a, f*C + k', s/N, q]   (N <= S, N | S)
//   out[a, f*C + k', 0, q]    (+=) when N spans several blocks (accumulate).
// The CF32 spectra never reach HBM.
template <int NPOL, int KPT>
__global__ __launch_bounds__(1024) void chan_detect_kernel(const float2* __restrict__ in,
                                                           float* __restrict__ out, int C,
                                                           int S, int N, uint64_t n_af,
                                                           const float2* __restrict__ tw,
                                                           Radices rad, int accumulate) {
    extern __shared__ float2 data[];
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
// Power-of-two rates 16 <= C <= 8192 with both polarizations.
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
                                                    int N, int accumulate,
                                                    const float2* __restrict__ tw) {
    extern __shared__ float2 data[];
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
    uint64_t blocks = ceil_div(n_items, pl.spw);
    const uint64_t cap = 256ull * (pl.threads >= 1024 ? 1 : (pl.threads >= 512 ? 2 : 4)) * 4;
    if (blocks > cap) blocks = cap;
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

template <int C, int P>
static void launch_small(const float2* in, float2* out, uint64_t n_spec, uint64_t S, hipStream_t s) {
    SmallTw<C> tw;
    for (int m = 0; m < C; m++) {
        const double a = -2.0 * M_PI * (double)m / (double)C;
        tw.w[m] = make_float2((float)cos(a), (float)sin(a));
    }
    uint64_t blocks = ceil_div(n_spec, 256);
    if (blocks > 256 * 32) blocks = 256 * 32;
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
    if (input == output && S != 1)
        return fail("channelizer: in-place operation needs rate == number of time samples");
    const float2* in = static_cast<const float2*>(input);
    float2* out = static_cast<float2*>(output);
    if (C <= 16 && input != output) {
        const int ok = (P == 2) ? dispatch_small<2>(C, in, out, n_spec, S, s)
                                : dispatch_small<1>(C, in, out, n_spec, S, s);
        if (ok) return check_launch("channelizer");
    }
    if ((uint64_t)C * P > (uint64_t)kMaxLdsComplex)
        return fail("channelizer: rate * polarizations above 16384 is not supported");
    const float2* tw = nullptr;
    int rc = get_twiddles(C, s, &tw);
    if (rc) return rc;
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
    if (rate > (uint64_t)kMaxLdsComplex) return fail("channelizer: rate above 16384 is not supported");
    if (rate <= 16) return BLADE_SUCCESS;   // register kernels carry their twiddles as arguments
    const float2* tw = nullptr;
    return get_twiddles((int)rate, reinterpret_cast<hipStream_t>(stream), &tw);
}

extern "C" int blade_channelizer_detector(const void* input, void* output, uint64_t A, uint64_t F,
                                          uint64_t T, uint64_t rate, uint64_t integration,
                                          uint64_t out_pols, int accumulate, blade_stream_t stream) {
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    if (rate == 0 || T % rate) return fail("channelizer-detector: T must be divisible by the rate");
    if (rate % 2 && rate != 1) return fail("channelizer-detector: the rate should be divisible by 2");
    if (out_pols != 1 && out_pols != 4) return fail("channelizer-detector: output polarizations must be 1 or 4");
    const uint64_t S = T / rate;
    if (integration == 0) return fail("channelizer-detector: integration size should be more than zero");
    if (integration <= S ? (S % integration != 0) : (integration % S != 0))
        return fail("channelizer-detector: integration must divide, or be a multiple of, the spectra per block");
    if (2 * rate > (uint64_t)kMaxLdsComplex) return fail("channelizer-detector: rate above 8192 is not supported");
    if (A * F * T == 0) return BLADE_SUCCESS;
    if (!input || !output) return fail("channelizer-detector: null pointer");
    if ((uintptr_t)input % 16 || (uintptr_t)output % 16) return fail("channelizer-detector: misaligned buffer");
    const int C = (int)rate;
    Radices rad;
    if (!factorize(C, &rad)) return fail("channelizer-detector: rate must factor into 2, 3, 5 and 7");
    const float2* tw = nullptr;
    int rc = get_twiddles(C, s, &tw);
    if (rc) return rc;
    Fft2Plan pl;
    if (fft2_plan(C, &pl)) {
        if (out_pols == 1)
            return launch_fft2<1>(pl, static_cast<const float2*>(input), output, C, S, A * F, (int)integration,
                                  accumulate, tw, s);
        return launch_fft2<4>(pl, static_cast<const float2*>(input), output, C, S, A * F, (int)integration,
                              accumulate, tw, s);
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
                           (int)S, (int)integration, n_af, tw, rad, accumulate);
    else
        hipLaunchKernelGGL((chan_detect_kernel<4, 8>), dim3((unsigned)blocks), dim3(threads), lds, s, in, out, C,
                           (int)S, (int)integration, n_af, tw, rad, accumulate);
    return check_launch("channelizer-detector");
}
