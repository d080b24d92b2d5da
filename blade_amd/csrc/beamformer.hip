// ATA beamformer on CDNA4 matrix cores, with an optional fused channelizer
// (rate 1/2/4) + polarizer epilogue.
//
//   y[b,f,t,p] = sum_a x[a,f,t,p] * w[b,a,f,0,p]          (no conjugate)
//   y[B,f,t,p] = (sum_a |x w[0]|^2, 0)  [sqrt'ed]         (incoherent beam)
// (reference src/modules/beamformer/beamformer.cu:13-79).
//
// The complex contraction per (f, p) is embedded as a real GEMM
//     [Re y; Im y] = [[Wr, -Wi], [Wi, Wr]] . [Re x; Im x]
// with M = 2B (rows: beam, re/im), K = 2A (antenna, re/im), N = T.  Phasors are
// the A operand (staged once per workgroup in LDS), voltages the B operand,
// read straight from HBM in their native [A,F,T,P] layout: each lane owns NT
// consecutive time samples of one antenna, i.e. one 16-byte load (CI8, NT=4)
// feeds 2 components x 2 pols x NT MFMAs.  The K index is ordered so that the
// lane halves (32x32 tiles) or quarters (16x16 tiles) of the wavefront read
// DIFFERENT antennas - no voltage is fetched twice.
//
// Two arithmetic variants:
//  * FP32: v_mfma_f32_{32x32x2,16x16x4}_f32, FP32 accumulate.
//  * split FP16 (CI8 input only): int8 -> fp16 is exact; each phasor
//    component is scaled by a per-(tile,pol) power of two, split into
//    hi = fp16(w) and lo = fp16(w - hi), and both halves are multiplied on
//    v_mfma_f32_{32x32x16,16x16x32}_f16 into one FP32 accumulator; the power
//    of two is removed exactly in the epilogue.
//
// The accumulator of a 32x32 tile holds, for lane l and register r,
// row 8(r/4) + 4(l/32) + r%4 and column l%32; rows are (beam, re/im) pairs so
// registers (r, r+1), r even, are one complex beam sample.  A 16x16 tile holds
// row 4(l/16) + r, column l%16.
#include "common.hh"

#include <algorithm>

namespace bl {

template <int TILE>
struct Acc;
template <>
struct Acc<32> {
    typedef floatx16 type;
    static constexpr int regs = 16;
};
template <>
struct Acc<16> {
    typedef floatx4 type;
    static constexpr int regs = 4;
};

template <int TILE>
__device__ __forceinline__ typename Acc<TILE>::type mfma_f32(float a, float b,
                                                             typename Acc<TILE>::type c) {
    if constexpr (TILE == 32) return __builtin_amdgcn_mfma_f32_32x32x2f32(a, b, c, 0, 0, 0);
    else return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

template <int TILE>
__device__ __forceinline__ typename Acc<TILE>::type mfma_f16(halfx8 a, halfx8 b,
                                                             typename Acc<TILE>::type c) {
    if constexpr (TILE == 32) return __builtin_amdgcn_mfma_f32_32x32x16_f16(a, b, c, 0, 0, 0);
    else return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
}

// Row index of accumulator register r for a lane in k-group g.
template <int TILE>
__device__ __forceinline__ int acc_row(int r, int g) {
    if constexpr (TILE == 32) return 8 * (r / 4) + 4 * g + (r % 4);
    else return 4 * g + r;
}

// Exact CI8 -> FP16 for one pol of a dword (re0, im0, re1, im1): the biased
// byte u = b + 128 is placed in the mantissa of 1024.0 (0x6400), giving
// 1024 + u exactly, and 1152 is subtracted.
__device__ __forceinline__ halfx2 ci8_pol_to_half2(uint32_t d, int p) {
    const uint32_t u = d ^ 0x80808080u;
    const uint32_t sh = 16u * p;
    const uint32_t bits = ((u >> sh) & 0xFFu) | (((u >> (sh + 8)) & 0xFFu) << 16) | 0x64006400u;
    halfx2 h = __builtin_bit_cast(halfx2, bits);
    const halfx2 bias = {(_Float16)1152.0f, (_Float16)1152.0f};
    return h - bias;
}

__device__ __forceinline__ float ci8_byte(uint32_t d, int idx) {
    return (float)(int8_t)((d >> (8 * idx)) & 0xFFu);
}

struct BfArgs {
    const void* x;          // [A,F,T,P] CI8 or CF32
    const float2* w;        // [B,A,F,1,P] CF32
    float2* y;              // output
    int A, F, T, B;
    int incoherent, inc_sqrt;
    int rate;               // epilogue channelizer rate (1, 2, 4)
    int polarize;
    int nsteps;             // K steps
};

// Epilogue: one complex beam sample for NT consecutive t, both pols, at
// t0 = first time sample of the lane; writes plain or channelized output.
template <int NT>
__device__ __forceinline__ void store_beam(const BfArgs& a, int b, int f, int t0,
                                           float2 (&v)[2][NT]) {
    if (a.rate == 1) {
        if (a.polarize) {
#pragma unroll
            for (int n = 0; n < NT; n++) {
                const float2 X = v[0][n], Y = v[1][n];
                v[0][n] = make_float2(X.x - Y.y, X.y + Y.x);
                v[1][n] = make_float2(X.x + Y.y, X.y - Y.x);
            }
        }
        float4* dst = reinterpret_cast<float4*>(a.y) + ((size_t)b * a.F + f) * a.T + t0;
#pragma unroll
        for (int n = 0; n < NT; n++)
            dst[n] = make_float4(v[0][n].x, v[0][n].y, v[1][n].x, v[1][n].y);
        return;
    }
    if constexpr (NT == 4) {
        const int C = a.rate;
        const int S = a.T / C;
        // out[b, f*C + k', s, p]  with k' = (k + C/2) mod C
        float4* base = reinterpret_cast<float4*>(a.y);
        if (C == 4) {
            float2 o[2][4];
#pragma unroll
            for (int p = 0; p < 2; p++) {
                const float2 s02 = cadd(v[p][0], v[p][2]), d02 = csub(v[p][0], v[p][2]);
                const float2 s13 = cadd(v[p][1], v[p][3]);
                const float2 dd = csub(v[p][1], v[p][3]);
                const float2 d13 = make_float2(dd.y, -dd.x);   // -j (x1 - x3)
                o[p][0] = cadd(s02, s13);
                o[p][1] = cadd(d02, d13);
                o[p][2] = csub(s02, s13);
                o[p][3] = csub(d02, d13);
            }
            const int s = t0 / 4;
#pragma unroll
            for (int k = 0; k < 4; k++) {
                float2 X = o[0][k], Y = o[1][k];
                if (a.polarize) {
                    const float2 L = make_float2(X.x - Y.y, X.y + Y.x);
                    const float2 R = make_float2(X.x + Y.y, X.y - Y.x);
                    X = L;
                    Y = R;
                }
                const int kp = (k + 2) & 3;
                base[(((size_t)b * a.F + f) * 4 + kp) * S + s] = make_float4(X.x, X.y, Y.x, Y.y);
            }
        } else {   // C == 2: two spectra per lane
#pragma unroll
            for (int h = 0; h < 2; h++) {
                const int s = t0 / 2 + h;
#pragma unroll
                for (int k = 0; k < 2; k++) {
                    float2 X = k == 0 ? cadd(v[0][2 * h], v[0][2 * h + 1]) : csub(v[0][2 * h], v[0][2 * h + 1]);
                    float2 Y = k == 0 ? cadd(v[1][2 * h], v[1][2 * h + 1]) : csub(v[1][2 * h], v[1][2 * h + 1]);
                    if (a.polarize) {
                        const float2 L = make_float2(X.x - Y.y, X.y + Y.x);
                        const float2 R = make_float2(X.x + Y.y, X.y - Y.x);
                        X = L;
                        Y = R;
                    }
                    const int kp = k ^ 1;
                    base[(((size_t)b * a.F + f) * 2 + kp) * S + s] = make_float4(X.x, X.y, Y.x, Y.y);
                }
            }
        }
    }
}

// LDS weight layout, FP32: wl[p][step][g][i] (i = row of the A operand).
// LDS weight layout, split FP16: wh[p][step][hl][lane] halfx8.
template <int TILE, int NT, bool CI8, bool F16X2>
__global__ __launch_bounds__(256, 2) void beamform_kernel(BfArgs a) {
    constexpr int KG = 64 / TILE;                 // k-groups of lanes
    constexpr int AK = F16X2 ? 4 * KG : KG;       // antennas per K step (pair of steps for FP32)
    constexpr int REGS = Acc<TILE>::regs;
    typedef typename Acc<TILE>::type AccT;

    extern __shared__ __align__(16) unsigned char lds_raw[];
    __shared__ unsigned int s_max[2];
    __shared__ float s_w0sq[2][128];

    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int i = lane % TILE;     // A-operand row / accumulator column
    const int g = lane / TILE;     // k-group
    const int f = blockIdx.y;
    const int mtile = blockIdx.z;
    const int b_base = mtile * (TILE / 2);
    const int nsteps = a.nsteps;
    const bool do_inc = a.incoherent && mtile == 0;

    // ---- Stage phasors for this (f, beam tile) into LDS. -------------------
    if (threadIdx.x < 2) s_max[threadIdx.x] = 0u;
    __syncthreads();
    if constexpr (F16X2) {
        // Per-pol power-of-two scale so that every scaled component is < 1.
        float m[2] = {0.f, 0.f};
        for (int e = threadIdx.x; e < (TILE / 2) * a.A; e += blockDim.x) {
            const int bl_ = e / a.A, an = e - bl_ * a.A, b = b_base + bl_;
            if (b < a.B) {
                const float4 w = reinterpret_cast<const float4*>(a.w)[((size_t)b * a.A + an) * a.F + f];
                m[0] = fmaxf(m[0], fmaxf(fabsf(w.x), fabsf(w.y)));
                m[1] = fmaxf(m[1], fmaxf(fabsf(w.z), fabsf(w.w)));
            }
        }
        atomicMax(&s_max[0], __float_as_uint(m[0]));
        atomicMax(&s_max[1], __float_as_uint(m[1]));
        __syncthreads();
        halfx8* wh = reinterpret_cast<halfx8*>(lds_raw);
        float scale[2];
#pragma unroll
        for (int p = 0; p < 2; p++) {
            int ex = 0;
            frexpf(__uint_as_float(s_max[p]), &ex);
            scale[p] = ldexpf(1.0f, -ex);
        }
        // One halfx8 (8 k values) per (p, step, hl, lane).
        const int total = 2 * nsteps * 2 * 64;
        for (int e = threadIdx.x; e < total; e += blockDim.x) {
            const int ln = e & 63;
            const int hl = (e >> 6) & 1;
            const int st = (e >> 7) % nsteps;
            const int p = (e >> 7) / nsteps;
            const int row = ln % TILE, kg = ln / TILE;
            const int b = b_base + row / 2, co = row & 1;
            halfx8 h;
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const int an = st * AK + 4 * kg + q;
                float wr = 0.f, wi = 0.f;
                if (b < a.B && an < a.A) {
                    const float2 w = a.w[(((size_t)b * a.A + an) * a.F + f) * 2 + p];
                    wr = w.x * scale[p];
                    wi = w.y * scale[p];
                }
                // k = (antenna, c): c = 0 multiplies Re x, c = 1 multiplies Im x.
                const float v0 = co == 0 ? wr : wi;      // coefficient of Re x
                const float v1 = co == 0 ? -wi : wr;     // coefficient of Im x
                const _Float16 h0 = (_Float16)v0, h1 = (_Float16)v1;
                if (hl == 0) {
                    h[2 * q] = h0;
                    h[2 * q + 1] = h1;
                } else {
                    h[2 * q] = (_Float16)(v0 - (float)h0);
                    h[2 * q + 1] = (_Float16)(v1 - (float)h1);
                }
            }
            wh[e] = h;
        }
    } else {
        float* wl = reinterpret_cast<float*>(lds_raw);
        // nsteps = 2 * ceil(A / KG); step st = 2q + c; entry [p][st][kg][row]
        const int total = 2 * nsteps * 64;
        for (int e = threadIdx.x; e < total; e += blockDim.x) {
            const int row = e % TILE;
            const int kg = (e / TILE) % KG;
            const int st = (e / 64) % nsteps;
            const int p = e / (64 * nsteps);
            const int an = (st / 2) * KG + kg, c = st & 1;
            const int b = b_base + row / 2, co = row & 1;
            float v = 0.f;
            if (b < a.B && an < a.A) {
                const float2 w = a.w[(((size_t)b * a.A + an) * a.F + f) * 2 + p];
                v = (co == 0) ? (c == 0 ? w.x : -w.y) : (c == 0 ? w.y : w.x);
            }
            wl[e] = v;
        }
    }
    if (a.incoherent && mtile == 0) {
        for (int e = threadIdx.x; e < 2 * a.A; e += blockDim.x) {
            const int an = e >> 1, p = e & 1;
            const float2 w = a.w[((size_t)an * a.F + f) * 2 + p];   // beam 0
            if (an < 128) s_w0sq[p][an] = w.x * w.x + w.y * w.y;
        }
    }
    __syncthreads();

    // ---- Time samples owned by this lane. ----------------------------------
    const int t_wave = (blockIdx.x * (blockDim.x / 64) + wave) * (TILE * NT);
    const int t0 = t_wave + NT * i;
    const bool t_ok = t0 < a.T;

    AccT acc[2][NT];
#pragma unroll
    for (int p = 0; p < 2; p++)
#pragma unroll
        for (int n = 0; n < NT; n++)
#pragma unroll
            for (int r = 0; r < REGS; r++) acc[p][n][r] = 0.f;
    float inc[2][NT];
#pragma unroll
    for (int p = 0; p < 2; p++)
#pragma unroll
        for (int n = 0; n < NT; n++) inc[p][n] = 0.f;

    const size_t ant_stride = (size_t)a.F * a.T;   // samples between antennas
    const size_t xoff = (size_t)f * a.T + t0;

    if constexpr (F16X2) {
        // K step st covers antennas st*AK + 4g + q, q = 0..3.
        const uint32_t* x = static_cast<const uint32_t*>(a.x);
        const halfx8* wh = reinterpret_cast<const halfx8*>(lds_raw);
        for (int st = 0; st < nsteps; st++) {
            uint32_t d[4][NT];
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const int an = st * AK + 4 * g + q;
                if (t_ok && an < a.A) {
                    const uint32_t* src = x + an * ant_stride + xoff;
                    if constexpr (NT == 4) {
                        const uint4 v = *reinterpret_cast<const uint4*>(src);
                        d[q][0] = v.x; d[q][1] = v.y; d[q][2] = v.z; d[q][3] = v.w;
                    } else {
#pragma unroll
                        for (int n = 0; n < NT; n++) d[q][n] = src[n];
                    }
                } else {
#pragma unroll
                    for (int n = 0; n < NT; n++) d[q][n] = 0x0u;
                }
            }
#pragma unroll
            for (int p = 0; p < 2; p++) {
                const halfx8 whi = wh[((p * nsteps + st) * 2 + 0) * 64 + lane];
                const halfx8 wlo = wh[((p * nsteps + st) * 2 + 1) * 64 + lane];
#pragma unroll
                for (int n = 0; n < NT; n++) {
                    halfx8 xb;
#pragma unroll
                    for (int q = 0; q < 4; q++) {
                        const halfx2 h = ci8_pol_to_half2(d[q][n], p);
                        xb[2 * q] = h[0];
                        xb[2 * q + 1] = h[1];
                    }
                    acc[p][n] = mfma_f16<TILE>(whi, xb, acc[p][n]);
                    acc[p][n] = mfma_f16<TILE>(wlo, xb, acc[p][n]);
                }
            }
            if (do_inc) {
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    const int an = st * AK + 4 * g + q;
                    if (an < a.A) {
#pragma unroll
                        for (int p = 0; p < 2; p++) {
                            const float w2 = s_w0sq[p][an];
#pragma unroll
                            for (int n = 0; n < NT; n++) {
                                const float xr = ci8_byte(d[q][n], 2 * p), xi = ci8_byte(d[q][n], 2 * p + 1);
                                inc[p][n] += (xr * xr + xi * xi) * w2;
                            }
                        }
                    }
                }
            }
        }
    } else {
        const float* wl = reinterpret_cast<const float*>(lds_raw);
        const int npairs = nsteps / 2;
        for (int q = 0; q < npairs; q++) {
            const int an = q * KG + g;
            float xv[2][2][NT];   // [p][c][n]
            if (t_ok && an < a.A) {
                if constexpr (CI8) {
                    const uint32_t* src = static_cast<const uint32_t*>(a.x) + an * ant_stride + xoff;
                    uint32_t d[NT];
                    if constexpr (NT == 4) {
                        const uint4 v = *reinterpret_cast<const uint4*>(src);
                        d[0] = v.x; d[1] = v.y; d[2] = v.z; d[3] = v.w;
                    } else {
#pragma unroll
                        for (int n = 0; n < NT; n++) d[n] = src[n];
                    }
#pragma unroll
                    for (int n = 0; n < NT; n++)
#pragma unroll
                        for (int p = 0; p < 2; p++)
#pragma unroll
                            for (int c = 0; c < 2; c++) xv[p][c][n] = ci8_byte(d[n], 2 * p + c);
                } else {
                    const float4* src = static_cast<const float4*>(a.x) + an * ant_stride + xoff;
#pragma unroll
                    for (int n = 0; n < NT; n++) {
                        const float4 v = src[n];
                        xv[0][0][n] = v.x; xv[0][1][n] = v.y; xv[1][0][n] = v.z; xv[1][1][n] = v.w;
                    }
                }
            } else {
#pragma unroll
                for (int n = 0; n < NT; n++)
#pragma unroll
                    for (int p = 0; p < 2; p++)
#pragma unroll
                        for (int c = 0; c < 2; c++) xv[p][c][n] = 0.f;
            }
#pragma unroll
            for (int p = 0; p < 2; p++) {
#pragma unroll
                for (int c = 0; c < 2; c++) {
                    const float wv = wl[((p * nsteps + 2 * q + c) * KG + g) * TILE + i];
#pragma unroll
                    for (int n = 0; n < NT; n++) acc[p][n] = mfma_f32<TILE>(wv, xv[p][c][n], acc[p][n]);
                }
            }
            if (do_inc && an < a.A) {
#pragma unroll
                for (int p = 0; p < 2; p++) {
                    const float w2 = s_w0sq[p][an];
#pragma unroll
                    for (int n = 0; n < NT; n++)
                        inc[p][n] += (xv[p][0][n] * xv[p][0][n] + xv[p][1][n] * xv[p][1][n]) * w2;
                }
            }
        }
    }

    // ---- Epilogue. ------------------------------------------------------------
    float unscale[2] = {1.f, 1.f};
    if constexpr (F16X2) {
#pragma unroll
        for (int p = 0; p < 2; p++) {
            int ex = 0;
            frexpf(__uint_as_float(s_max[p]), &ex);
            unscale[p] = ldexpf(1.0f, ex);
        }
    }
    if (t_ok) {
#pragma unroll
        for (int r = 0; r < REGS; r += 2) {
            const int row = acc_row<TILE>(r, g);
            const int b = b_base + row / 2;
            if (b < a.B) {
                float2 v[2][NT];
#pragma unroll
                for (int p = 0; p < 2; p++)
#pragma unroll
                    for (int n = 0; n < NT; n++)
                        v[p][n] = make_float2(acc[p][n][r] * unscale[p], acc[p][n][r + 1] * unscale[p]);
                store_beam<NT>(a, b, f, t0, v);
            }
        }
    }
    if (do_inc) {
#pragma unroll
        for (int p = 0; p < 2; p++)
#pragma unroll
            for (int n = 0; n < NT; n++) {
#pragma unroll
                for (int m = TILE; m < 64; m *= 2) inc[p][n] += __shfl_xor(inc[p][n], m, 64);
            }
        if (t_ok && g == 0) {
            float2 v[2][NT];
#pragma unroll
            for (int p = 0; p < 2; p++)
#pragma unroll
                for (int n = 0; n < NT; n++)
                    v[p][n] = make_float2(a.inc_sqrt ? sqrtf(inc[p][n]) : inc[p][n], 0.f);
            store_beam<NT>(a, a.B, f, t0, v);
        }
    }
}

template <int TILE, int NT, bool CI8, bool F16X2>
static int launch_bf(const BfArgs& args, hipStream_t s) {
    constexpr int KG = 64 / TILE;
    BfArgs a = args;
    if (F16X2) a.nsteps = (a.A + 4 * KG - 1) / (4 * KG);
    else a.nsteps = 2 * ((a.A + KG - 1) / KG);
    const size_t lds = F16X2 ? (size_t)2 * a.nsteps * 2 * 64 * 16 : (size_t)2 * a.nsteps * 64 * 4;
    const int waves = 4;
    const int span = waves * TILE * NT;
    const int mtiles = (a.B + TILE / 2 - 1) / (TILE / 2);
    dim3 grid((unsigned)((a.T + span - 1) / span), (unsigned)a.F, (unsigned)mtiles);
    const void* fn = reinterpret_cast<const void*>(&beamform_kernel<TILE, NT, CI8, F16X2>);
    if (lds > 150 * 1024) return fail("beamformer: too many antennas for the LDS phasor stage");
    if (lds > 65536) {
        hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) {
            set_error(std::string("beamformer: LDS attribute: ") + hipGetErrorString(e));
            return BLADE_HIP_ERROR;
        }
    }
    hipLaunchKernelGGL((beamform_kernel<TILE, NT, CI8, F16X2>), grid, dim3(64 * waves), lds, s, a);
    return check_launch("beamformer");
}

static int run_beamformer(int in_type, int precision, const void* input, const void* phasors,
                          void* output, uint64_t A, uint64_t F, uint64_t T, uint64_t P,
                          uint64_t B, int incoherent, int incoherent_sqrt, uint64_t rate,
                          int polarize, hipStream_t s) {
    if (P != 2) return fail("beamformer: number of polarizations must be 2");
    if (in_type != BLADE_CF32 && in_type != BLADE_CI8) return fail("beamformer: input must be CF32 or CI8");
    if (precision != BLADE_PREC_F32 && precision != BLADE_PREC_F16X2) return fail("beamformer: unknown precision");
    if (precision == BLADE_PREC_F16X2 && in_type != BLADE_CI8)
        return fail("beamformer: split-FP16 needs CI8 input (exact widening)");
    if (rate != 1 && rate != 2 && rate != 4) return fail("beamformer chain: rate must be 1, 2 or 4");
    if (T % rate) return fail("beamformer chain: T must be divisible by the rate");
    if (incoherent && A > 128) return fail("beamformer: incoherent beam supports up to 128 antennas");
    if (A == 0 || B == 0) return fail("beamformer: need at least one antenna and one beam");
    if (A > 4096 || F > 65535 || T > (1u << 30) || B > 4096) return fail("beamformer: shape too large");
    if (F == 0 || T == 0) return BLADE_SUCCESS;
    if (!input || !phasors || !output) return fail("beamformer: null pointer");
    if ((uintptr_t)input % 16 || (uintptr_t)phasors % 16 || (uintptr_t)output % 16)
        return fail("beamformer: buffers must be 16-byte aligned");
    BfArgs a;
    a.x = input;
    a.w = static_cast<const float2*>(phasors);
    a.y = static_cast<float2*>(output);
    a.A = (int)A; a.F = (int)F; a.T = (int)T; a.B = (int)B;
    a.incoherent = incoherent;
    a.inc_sqrt = incoherent_sqrt;
    a.rate = (int)rate;
    a.polarize = polarize;
    a.nsteps = 0;
    const bool nt4 = (T % 4 == 0);
    if (rate > 1 && !nt4) return fail("beamformer chain: T must be divisible by 4");
    const bool small = B <= 8;   // M = 2B fits a 16-row tile
    const bool ci8 = in_type == BLADE_CI8;
    if (precision == BLADE_PREC_F16X2) {
        if (small) return nt4 ? launch_bf<16, 4, true, true>(a, s) : launch_bf<16, 1, true, true>(a, s);
        return nt4 ? launch_bf<32, 4, true, true>(a, s) : launch_bf<32, 1, true, true>(a, s);
    }
    if (ci8) {
        if (small) return nt4 ? launch_bf<16, 4, true, false>(a, s) : launch_bf<16, 1, true, false>(a, s);
        return nt4 ? launch_bf<32, 4, true, false>(a, s) : launch_bf<32, 1, true, false>(a, s);
    }
    if (small) return nt4 ? launch_bf<16, 4, false, false>(a, s) : launch_bf<16, 1, false, false>(a, s);
    return nt4 ? launch_bf<32, 4, false, false>(a, s) : launch_bf<32, 1, false, false>(a, s);
}

}  // namespace bl

using namespace bl;

extern "C" int blade_beamformer_ata(int in_type, int precision, const void* input,
                                    const void* phasors, void* output, uint64_t A, uint64_t F,
                                    uint64_t T, uint64_t P, uint64_t B, int incoherent,
                                    int incoherent_sqrt, blade_stream_t stream) {
    return run_beamformer(in_type, precision, input, phasors, output, A, F, T, P, B, incoherent,
                          incoherent_sqrt, 1, 0, reinterpret_cast<hipStream_t>(stream));
}

extern "C" int blade_beamform_chain(int in_type, int precision, const void* input,
                                    const void* phasors, void* output, uint64_t A, uint64_t F,
                                    uint64_t T, uint64_t P, uint64_t B, int incoherent,
                                    int incoherent_sqrt, uint64_t rate, int polarize,
                                    blade_stream_t stream) {
    return run_beamformer(in_type, precision, input, phasors, output, A, F, T, P, B, incoherent,
                          incoherent_sqrt, rate, polarize, reinterpret_cast<hipStream_t>(stream));
}
