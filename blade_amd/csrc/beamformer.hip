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
// time samples of one antenna in runs of 4 consecutive samples (one 16-byte
// CI8 load feeds 2 components x 2 pols x 4 x MW MFMAs), and the 16 lanes of a
// group read 256 contiguous bytes per load.  The K index is ordered so that the
// four 16-lane groups of the wavefront read DIFFERENT antennas: no voltage is
// fetched twice.
//
// 16x16 tiles are used throughout: on gfx950 v_mfma_f32_16x16x4_f32 and
// v_mfma_f32_16x16x32_f16 issue at the full FP32 / FP16 matrix rate.  A
// wavefront holds MW beam tiles (8 beams each) so that every voltage it loads
// is multiplied by up to 16 beams.
//
// Two arithmetic variants:
//  * FP32: v_mfma_f32_16x16x4_f32, FP32 accumulate.  With 16 beams per
//    wavefront (MW = 2) the tile rows are 16 beams and the complex product
//    takes three real contractions instead of four (Gauss):
//        P1 = Wr.Xr,  P2 = Wi.Xi,  P3 = (Wr + Wi).(Xr + Xi),
//        Re y = P1 - P2,  Im y = P3 - P1 - P2,
//    with Xr + Xi exact for CI8 input and Wr + Wi formed once in the pack
//    kernel; with 8 beams (MW = 1) the rows are (beam, re/im) pairs of the
//    embedding below.
//  * split FP16 (CI8 input only): int8 -> fp16 is exact; each phasor
//    component is scaled by a per-(workgroup, pol) power of two, split into
//    hi = fp16(w) and lo = fp16(w - hi), and both halves are multiplied on
//    v_mfma_f32_16x16x32_f16 into one FP32 accumulator; the power of two is
//    removed exactly in the epilogue.
//
// The accumulator of a 16x16 tile holds, for lane l and register r, row
// 4(l/16) + r and column l%16; rows are (beam, re/im) pairs, so registers
// (r, r+1), r even, hold one complex beam sample of time sample l%16.
#include <mutex>

#include "common.hh"

#include <algorithm>
#include <cstdlib>
#include <cstring>

namespace bl {

constexpr int TILE = 16;      // rows and columns of one MFMA tile
constexpr int KG = 4;         // lane groups (l / 16) along K
constexpr int WAVES = 4;      // wavefronts per workgroup

__device__ __forceinline__ floatx4 mfma_f32(float a, float b, floatx4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
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
    int nsteps;             // K steps (FP32: 2 per antenna group; FP16: 1 per 16 antennas)
    int n_tchunks;          // chunks of WAVES*TILE*NT samples along T
    int chunks_per_wg;      // chunks walked by one workgroup
    size_t wbytes;          // bytes of the phasor A operands in LDS
    float4* sink;           // 64 x 16 B: target of the stores of lanes with nothing to store
};

// Packs the phasors of one (f, beam group) into the A-operand layout the
// beamformer reads from LDS, followed by |w0|^2[p][a] (256 floats) and the
// split-FP16 unscale factors (2 floats).
//   FP32       [m][p][step][g][row] float, step = 2 * antenna group + component
//   split FP16 [m][p][step][hl][lane] halfx8 (8 k values = 4 antennas x re/im),
//              components scaled by a per-pol power of two so that all are < 1,
//              hi = fp16(w), lo = fp16(w - hi).
// G3 selects the three-product FP32 layout (16-beam tiles, see bf_body).
//
// bf_pack_fill writes the (beam tile m, pol p) part from staged phasors `sw`
// ([SB][A] float2 of pol p: the 8 beams of tile m for FP32, all MW * 8 beams
// of the group for split FP16, whose scale is shared by the group's pol).
template <int MW, bool F16X2, bool G3>
__device__ __forceinline__ void bf_pack_fill(const BfArgs& a, const int m, const int p, const float2* sw,
                                             const float scale, unsigned char* dst) {
    constexpr int AK = F16X2 ? 4 * KG : KG;
    const int A = a.A;
    const int nsteps = a.nsteps;
    if constexpr (F16X2) {
        // beam (within the staged rows) of A-operand row `row` of tile m
        const int row_beam0 = m * (TILE / 2);
        // [m][p][step][hl][lane] halfx8, the (m, p) block
        halfx8* wh = reinterpret_cast<halfx8*>(dst) + (size_t)((m * 2 + p) * nsteps) * 128;
        for (int e = threadIdx.x; e < nsteps * 128; e += blockDim.x) {
            const int st = e >> 7, hl = (e >> 6) & 1, ln = e & 63;
            const int row = ln & (TILE - 1), kg = ln >> 4;
            const int bl_ = row_beam0 + (row >> 1), co = row & 1;
            halfx8 h;
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const int an = st * AK + 4 * kg + q;
                float wr = 0.f, wi = 0.f;
                if (an < A) {
                    const float2 w = sw[bl_ * A + an];
                    wr = w.x * scale;
                    wi = w.y * scale;
                }
                // k = (antenna, c): c = 0 multiplies Re x, c = 1 multiplies Im x.
                const float v0 = co == 0 ? wr : wi;
                const float v1 = co == 0 ? -wi : wr;
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
    } else if constexpr (G3) {
        // Three-product layout [p][unit][q][g][row] float, q = (Wr, Wi, Wr + Wi),
        // row = beam of the 16-beam tile; this part is rows 8m .. 8m + 7.
        const int nunits = nsteps / 2;
        float* wl = reinterpret_cast<float*>(dst) + (size_t)p * nunits * 3 * 64;
        for (int e = threadIdx.x; e < nunits * 3 * 32; e += blockDim.x) {
            const int rl = e & 7, kg = (e >> 3) & 3, uq = e >> 5;
            const int un = uq / 3, q = uq - 3 * un;
            const int an = un * KG + kg;
            float v = 0.f;
            if (an < A) {
                const float2 w = sw[rl * A + an];
                v = q == 0 ? w.x : (q == 1 ? w.y : w.x + w.y);
            }
            wl[(un * 3 + q) * 64 + kg * TILE + m * (TILE / 2) + rl] = v;
        }
    } else {
        // [m][p][step][g][row] float, step = 2 * antenna group + component
        float* wl = reinterpret_cast<float*>(dst) + (size_t)((m * 2 + p) * nsteps) * 64;
        for (int e = threadIdx.x; e < nsteps * 64; e += blockDim.x) {
            const int st = e >> 6, kg = (e >> 4) & 3, row = e & (TILE - 1);
            const int an = (st >> 1) * KG + kg, c = st & 1;
            const int bl_ = row >> 1, co = row & 1;
            float v = 0.f;
            if (an < A) {
                const float2 w = sw[bl_ * A + an];
                v = (co == 0) ? (c == 0 ? w.x : -w.y) : (c == 0 ? w.y : w.x);
            }
            wl[e] = v;
        }
    }
}

// |w0|^2 of pol p over the antennas (row 0 of the staged beams), read only
// by the beam-group-0 workgroups, whose staged row 0 is beam 0.
__device__ __forceinline__ void bf_pack_w0sq(const BfArgs& a, const float2* sw, float* w0sq) {
    for (int an = threadIdx.x; an < 128; an += blockDim.x) {
        float v = 0.f;
        if (an < a.A) {
            const float2 w = sw[an];
            v = w.x * w.x + w.y * w.y;
        }
        w0sq[an] = v;
    }
}

// The whole (f, beam group) straight into the workgroup's LDS: one round of
// phasor loads for both pols and all MW * 8 beams (s_w: [2][MW * 8][A]),
// then every (m, p) part.  Ends with a barrier.
template <int MW, bool F16X2, bool G3>
__device__ __forceinline__ void bf_pack_lds(const BfArgs& a, const int f, const int gz, float2* s_w,
                                            unsigned int* s_max, unsigned char* dst, float* w0sq,
                                            float* unscale) {
    constexpr int NB = MW * (TILE / 2);
    const int A = a.A;
    if (threadIdx.x < 2) s_max[threadIdx.x] = 0u;
    float mx[2] = {0.f, 0.f};
    if (((uintptr_t)a.w & 15) == 0) {
        // Both pols of a (beam, antenna) are one 16-byte load: half the
        // scattered phasor loads (FP32 124.6 -> 123.6 us, split FP16 114.7
        // -> 113.0 us per Mode-B call).
        for (int r = threadIdx.x; r < NB * A; r += blockDim.x) {
            const int bl_ = r / A, an = r - bl_ * A, b = gz * NB + bl_;
            float4 w = make_float4(0.f, 0.f, 0.f, 0.f);
            if (b < a.B) w = reinterpret_cast<const float4*>(a.w)[((size_t)b * A + an) * a.F + f];
            s_w[r] = make_float2(w.x, w.y);
            s_w[NB * A + r] = make_float2(w.z, w.w);
            mx[0] = fmaxf(mx[0], fmaxf(fabsf(w.x), fabsf(w.y)));
            mx[1] = fmaxf(mx[1], fmaxf(fabsf(w.z), fabsf(w.w)));
        }
    } else {
        for (int e = threadIdx.x; e < 2 * NB * A; e += blockDim.x) {
            const int p = e / (NB * A), r = e - p * (NB * A);
            const int bl_ = r / A, an = r - bl_ * A, b = gz * NB + bl_;
            float2 w = make_float2(0.f, 0.f);
            if (b < a.B) w = a.w[(((size_t)b * A + an) * a.F + f) * 2 + p];
            s_w[e] = w;
            const float v = fmaxf(fabsf(w.x), fabsf(w.y));
            mx[0] = p == 0 ? fmaxf(mx[0], v) : mx[0];
            mx[1] = p == 1 ? fmaxf(mx[1], v) : mx[1];
        }
    }
    __syncthreads();
    float scale[2] = {1.f, 1.f}, uns[2] = {1.f, 1.f};
    if constexpr (F16X2) {
        atomicMax(&s_max[0], __float_as_uint(mx[0]));
        atomicMax(&s_max[1], __float_as_uint(mx[1]));
        __syncthreads();
#pragma unroll
        for (int p = 0; p < 2; p++) {
            int ex = 0;
            frexpf(__uint_as_float(s_max[p]), &ex);
            scale[p] = ldexpf(1.0f, -ex);
            uns[p] = ldexpf(1.0f, ex);
        }
    }
#pragma unroll
    for (int mp = 0; mp < 2 * MW; mp++) {
        const int p = mp & 1, m = mp >> 1;
        const float2* sw = s_w + (size_t)p * NB * A + (F16X2 ? 0 : m * (TILE / 2) * A);
        bf_pack_fill<MW, F16X2, G3>(a, m, p, sw, scale[p], dst);
    }
    if (gz == 0) {
        bf_pack_w0sq(a, s_w, w0sq);
        bf_pack_w0sq(a, s_w + (size_t)NB * A, w0sq + 128);
    }
    if (threadIdx.x < 2) unscale[threadIdx.x] = uns[threadIdx.x];
    __syncthreads();
}

__device__ __forceinline__ void polarize2(float2& X, float2& Y) {
    const float2 L = make_float2(X.x - Y.y, X.y + Y.x);
    const float2 R = make_float2(X.x + Y.y, X.y - Y.x);
    X = L;
    Y = R;
}

// Time sample of a lane's n-th column: lanes own runs of RUN consecutive
// samples, the runs of one lane TILE * RUN samples apart, so a 16-lane group
// reads and writes TILE * RUN contiguous samples per access.  RUN = 4 gives
// 16-byte voltage loads and keeps the 4 samples of a DFT-4 in one lane; the
// plain (rate 1) epilogues use RUN = 1, where each store instruction writes
// 16 lanes x 16 contiguous bytes instead of 16 bytes of every 64.
template <int NT, int RUN_ = (NT < 4 ? NT : 4)>
struct TimeMap {
    static constexpr int RUN = RUN_;
    // offset of column n from the lane's first sample (chunk base + RUN * i)
    static __device__ __forceinline__ constexpr int off(int n) {
        return (n / RUN) * (TILE * RUN) + n % RUN;
    }
};

// One 16-byte output store (both pols of one sample); nontemporal on request.
template <bool NTS>
__device__ __forceinline__ void store4(float4* p, float2 X, float2 Y) {
    if constexpr (NTS) {
        // Component-wise: merged into one global_store_dwordx4 ... nt (a
        // vector operand here would be lowered through private memory).
        float* q = reinterpret_cast<float*>(p);
        __builtin_nontemporal_store(X.x, q);
        __builtin_nontemporal_store(X.y, q + 1);
        __builtin_nontemporal_store(Y.x, q + 2);
        __builtin_nontemporal_store(Y.y, q + 3);
    } else {
        *p = make_float4(X.x, X.y, Y.x, Y.y);
    }
}

// Epilogue: beam b, both pols, the NT samples of a lane whose first sample is
// t0 (TimeMap); writes the plain beam, or spectra of `rate` samples
// (fftshifted) into [B', F*C, T/C, P], optionally polarized.  Every rate
// issues NT stores per call whatever `ok` says (lanes with nothing to store
// write to a.sink): the store count of an epilogue is the same on every path,
// which keeps the wait counts of the following loads exact.
// Epilogue variant: EPI = 0 reads rate / polarize / incoherent from the
// arguments; otherwise EPI = rate | 8 * polarize | 16 * incoherent, fixed at
// compile time (a straight-line epilogue).
__device__ __forceinline__ constexpr int epi_code(int rate, bool pol, bool inc) {
    return rate | (pol ? 8 : 0) | (inc ? 16 : 0);
}
template <int EPI> __device__ __forceinline__ int epi_rate(const BfArgs& a) { return EPI ? (EPI & 7) : a.rate; }
template <int NT, int EPI> constexpr int epi_run() { return (EPI && (EPI & 7) == 1) ? 1 : (NT < 4 ? NT : 4); }
template <int EPI> __device__ __forceinline__ bool epi_pol(const BfArgs& a) { return EPI ? (EPI & 8) != 0 : a.polarize != 0; }
template <int EPI> __device__ __forceinline__ bool epi_inc(const BfArgs& a) { return EPI ? (EPI & 16) != 0 : a.incoherent != 0; }

template <int NT, bool NTS, int EPI>
__device__ __forceinline__ void store_beam(const BfArgs& a, int b, int f, int t0,
                                           float2 (&v)[2][NT], bool ok) {
    typedef TimeMap<NT, epi_run<NT, EPI>()> M;
    float4* const sink = a.sink + (threadIdx.x & 63);
    const int rate = epi_rate<EPI>(a);
    const bool pol = epi_pol<EPI>(a);
    if (rate == 1) {
        if (pol) {
#pragma unroll
            for (int n = 0; n < NT; n++) polarize2(v[0][n], v[1][n]);
        }
        float4* dst = reinterpret_cast<float4*>(a.y) + ((size_t)b * a.F + f) * a.T + t0;
#pragma unroll
        for (int n = 0; n < NT; n++) store4<NTS>(ok ? dst + M::off(n) : sink, v[0][n], v[1][n]);
        return;
    }
    if constexpr (NT % 4 == 0) {
        float4* base = reinterpret_cast<float4*>(a.y);
        if (rate == 4) {
            const int S = a.T / 4;
#pragma unroll
            for (int h = 0; h < NT / 4; h++) {
                float2 o[2][4];
#pragma unroll
                for (int p = 0; p < 2; p++) {
                    const float2* x = &v[p][4 * h];
                    const float2 s02 = cadd(x[0], x[2]), d02 = csub(x[0], x[2]);
                    const float2 s13 = cadd(x[1], x[3]), dd = csub(x[1], x[3]);
                    const float2 d13 = make_float2(dd.y, -dd.x);   // -j (x1 - x3)
                    o[p][0] = cadd(s02, s13);
                    o[p][1] = cadd(d02, d13);
                    o[p][2] = csub(s02, s13);
                    o[p][3] = csub(d02, d13);
                }
                const int s = (t0 + M::off(4 * h)) / 4;
#pragma unroll
                for (int k = 0; k < 4; k++) {
                    float2 X = o[0][k], Y = o[1][k];
                    if (pol) polarize2(X, Y);
                    const int kp = (k + 2) & 3;   // fftshift
                    store4<NTS>(ok ? base + (((size_t)b * a.F + f) * 4 + kp) * S + s : sink, X, Y);
                }
            }
        } else {   // rate 2
            const int S = a.T / 2;
#pragma unroll
            for (int h = 0; h < NT / 2; h++) {
                const int s = (t0 + M::off(2 * h)) / 2;
#pragma unroll
                for (int k = 0; k < 2; k++) {
                    float2 X = k == 0 ? cadd(v[0][2 * h], v[0][2 * h + 1]) : csub(v[0][2 * h], v[0][2 * h + 1]);
                    float2 Y = k == 0 ? cadd(v[1][2 * h], v[1][2 * h + 1]) : csub(v[1][2 * h], v[1][2 * h + 1]);
                    if (pol) polarize2(X, Y);
                    store4<NTS>(ok ? base + (((size_t)b * a.F + f) * 2 + (k ^ 1)) * S + s : sink, X, Y);
                }
            }
        }
    }
}

constexpr int EPI_MODEB = 4 | 8;          // rate 4, polarize, no incoherent beam
constexpr int EPI_PLAIN = 1;              // rate 1
constexpr int EPI_PLAIN_INC = 1 | 16;     // rate 1, incoherent beam

// Mode-B epilogue through LDS (split FP16, 16-beam tiles): the four
// wavefronts' spectra of one chunk are exchanged so that each store
// instruction writes 1 KiB of one output row (64 consecutive spectra of one
// (beam, fine channel)) instead of 256 B pieces of four rows.  Half of the
// group's 16 beams at a time: [8 beams][4 fine channels][64 + 1] float4 in
// LDS (the staged phasors' space, free once they are packed).
template <int NT, bool F16X2, int MW, int EPI>
constexpr bool staged_epilogue() { return F16X2 && MW == 2 && NT == 4 && EPI == EPI_MODEB; }
constexpr int kStageRow = 65;                                   // float4 per staged row, padded
constexpr size_t kStageBytes = (size_t)8 * 4 * kStageRow * 16;  // one half of the beams

// One beam's 4 consecutive samples (one lane): DFT-4, fftshift, XY -> LR,
// into the staged rows of local beam lb at column col (the spectrum's index
// within the chunk).
__device__ __forceinline__ void stage_beam(float4* stg, int lb, int col, float2 (&v)[2][4]) {
    float2 o[2][4];
#pragma unroll
    for (int p = 0; p < 2; p++) {
        const float2* x = &v[p][0];
        const float2 s02 = cadd(x[0], x[2]), d02 = csub(x[0], x[2]);
        const float2 s13 = cadd(x[1], x[3]), dd = csub(x[1], x[3]);
        const float2 d13 = make_float2(dd.y, -dd.x);   // -j (x1 - x3)
        o[p][0] = cadd(s02, s13);
        o[p][1] = cadd(d02, d13);
        o[p][2] = csub(s02, s13);
        o[p][3] = csub(d02, d13);
    }
#pragma unroll
    for (int k = 0; k < 4; k++) {
        float2 X = o[0][k], Y = o[1][k];
        polarize2(X, Y);
        stg[(lb * 4 + ((k + 2) & 3)) * kStageRow + col] = make_float4(X.x, X.y, Y.x, Y.y);
    }
}

// One "unit" of voltage loads: what a lane needs for one K block.
//  FP32, CI8  : one antenna x NT samples          -> NT words
//  FP32, CF32 : one antenna x NT samples x 16 B   -> 4 NT words
//  split FP16 : four antennas x NT samples        -> 4 NT words
template <int NT, bool CI8, bool F16X2>
struct Unit {
    static constexpr int WORDS = (F16X2 || !CI8) ? 4 * NT : NT;
    uint32_t w[WORDS];
};

template <int NW>
__device__ __forceinline__ void load_words(uint32_t* dst, const uint32_t* src) {
    if constexpr (NW % 4 == 0) {
#pragma unroll
        for (int k = 0; k < NW / 4; k++) {
            const uint4 v = reinterpret_cast<const uint4*>(src)[k];
            dst[4 * k] = v.x; dst[4 * k + 1] = v.y; dst[4 * k + 2] = v.z; dst[4 * k + 3] = v.w;
        }
    } else {
#pragma unroll
        for (int n = 0; n < NW; n++) dst[n] = src[n];
    }
}

// LDS phasor layouts (A operands):
//   FP32       wl[m][p][step][g][row]        float   (step = 2 * antenna group + component)
//   split FP16 wh[m][p][step][hl][lane]      halfx8  (8 k values = 4 antennas x re/im)
//
// Work split: blockIdx = (t-range, f, beam group of MW tiles).  A workgroup
// walks up to `chunks_per_wg` chunks of WAVES*TILE*NT samples with the
// phasors staged once; voltage loads run D units ahead of the MFMAs through
// a register ring that crosses chunk boundaries.
// One work item: channel f, beam group gz (of `groups`), chunks cw0,
// cw0 + cstride, cw0 + 2 cstride, ... of the block described by `a`.  The
// single-block kernel interleaves (cw0 = blockIdx.x, cstride = gridDim.x),
// so the workgroups of one channel read and write adjacent pieces of each
// antenna and beam row at the same time; with voltages coming from HBM
// rather than the last-level cache that is ~3% faster than consecutive
// chunk ranges.  The work item packs its phasors straight into LDS
// (bf_pack_lds): lds_raw holds the A operands, then the staged phasors
// [2][MW * 8][A] float2; s_unscale / s_w0sq the split-FP16 unscale and |w0|^2.
template <int NT, int MW, bool CI8, bool F16X2, int D, bool NTS, int EPI, bool G3 = (!F16X2 && MW == 2)>
__device__ __forceinline__ void bf_body(const BfArgs& a, const int f, const int gz, const int cw0,
                                        unsigned char* lds_raw, float* s_unscale, float (*s_w0sq)[128],
                                        unsigned int* s_max, const int cstride = 1) {
    constexpr int AK = F16X2 ? 4 * KG : KG;       // antennas per unit
    typedef Unit<NT, CI8, F16X2> U_t;

    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int i = lane % TILE;     // A-operand row / accumulator column
    const int g = lane / TILE;     // k-group
    const int b_wg = gz * MW * (TILE / 2);
    const int nsteps = a.nsteps;
    const bool do_inc = epi_inc<EPI>(a) && gz == 0;

    const int n_units = F16X2 ? nsteps : nsteps / 2;
    const int upad = ((n_units + D - 1) / D) * D;
    const int nch = min(a.chunks_per_wg, (a.n_tchunks - cw0 + cstride - 1) / cstride);
    const int J = nch * upad;
    const size_t ant_stride = (size_t)a.F * a.T;   // complex samples between antennas

    typedef TimeMap<NT, epi_run<NT, EPI>()> M;
    // First sample of this lane in chunk c (runs of M::RUN samples, TimeMap).
    auto chunk_t0 = [&](int c) { return ((cw0 + c * cstride) * WAVES + wave) * (TILE * NT) + M::RUN * i; };

    // Loads are unconditional so that the wait-count analysis keeps them in
    // flight across the loop: indices past the end re-load the last unit,
    // padded antennas re-load antenna A-1 (their phasors are zero), and lanes
    // past T re-load t = 0 (their results are not stored).
    auto load_unit = [&](int j, U_t& u) {
        j = min(j, J - 1);
        const int c = j / upad;
        const int un = min(j - c * upad, n_units - 1);
        int t0 = chunk_t0(c);
        t0 = (t0 < a.T) ? t0 : 0;
        const size_t off = (size_t)f * a.T + t0;
        constexpr int R = M::RUN;
        if constexpr (F16X2) {
            const uint32_t* x = static_cast<const uint32_t*>(a.x);
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const int an = min(un * AK + 4 * g + q, a.A - 1);
#pragma unroll
                for (int h = 0; h < NT / R; h++)
                    load_words<R>(u.w + q * NT + h * R, x + an * ant_stride + off + M::off(h * R));
            }
        } else {
            const int an = min(un * KG + g, a.A - 1);
#pragma unroll
            for (int h = 0; h < NT / R; h++) {
                if constexpr (CI8) {
                    load_words<R>(u.w + h * R, static_cast<const uint32_t*>(a.x) + an * ant_stride + off + M::off(h * R));
                } else {
                    load_words<4 * R>(u.w + 4 * h * R,
                                      static_cast<const uint32_t*>(a.x) + 4 * (an * ant_stride + off + M::off(h * R)));
                }
            }
        }
    };

    // ---- Prime the voltage ring before staging phasors (overlaps latencies).
    U_t ring[D];
#pragma unroll
    for (int d = 0; d < D; d++) load_unit(d, ring[d]);

    // ---- This (f, beam group)'s packed phasors into LDS. ---------------------
    // A operands, |w0|^2 per antenna and pol, and the power-of-two unscale of
    // the split-FP16 path.
    bf_pack_lds<MW, F16X2, G3>(a, f, gz, reinterpret_cast<float2*>(lds_raw + a.wbytes), s_max, lds_raw,
                               &s_w0sq[0][0], s_unscale);
    const float unscale[2] = {s_unscale[0], s_unscale[1]};

    // Three-product FP32 tiles (16 beams): accumulators P1, P2, P3.
    static_assert(!G3 || (!F16X2 && MW == 2), "three products need FP32 and 16-beam tiles");
    constexpr int NACC = G3 ? 3 : MW;
    floatx4 acc[NACC][2][NT];
    float inc[2][NT];
    auto zero_acc = [&]() {
#pragma unroll
        for (int m = 0; m < NACC; m++)
#pragma unroll
            for (int p = 0; p < 2; p++)
#pragma unroll
                for (int n = 0; n < NT; n++)
#pragma unroll
                    for (int r = 0; r < 4; r++) acc[m][p][n][r] = 0.f;
#pragma unroll
        for (int p = 0; p < 2; p++)
#pragma unroll
            for (int n = 0; n < NT; n++) inc[p][n] = 0.f;
    };
    zero_acc();

    // MFMAs (and incoherent power) of one unit.
    auto compute = [&](const U_t& u, int un) {
        if constexpr (F16X2) {
            const halfx8* wh = reinterpret_cast<const halfx8*>(lds_raw);
            halfx8 whi[2][MW], wlo[2][MW];
#pragma unroll
            for (int p = 0; p < 2; p++)
#pragma unroll
                for (int m = 0; m < MW; m++) {
                    whi[p][m] = wh[(((m * 2 + p) * nsteps + un) * 2 + 0) * 64 + lane];
                    wlo[p][m] = wh[(((m * 2 + p) * nsteps + un) * 2 + 1) * 64 + lane];
                }
#pragma unroll
            for (int n = 0; n < NT; n++) {
                halfx8 xb[2];   // B operands of both pols from the same four dwords
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    halfx2 h0, h1;
                    ci8_to_half2x2(u.w[q * NT + n], h0, h1);
                    xb[0][2 * q] = h0[0];
                    xb[0][2 * q + 1] = h0[1];
                    xb[1][2 * q] = h1[0];
                    xb[1][2 * q + 1] = h1[1];
                }
#pragma unroll
                for (int p = 0; p < 2; p++)
#pragma unroll
                    for (int m = 0; m < MW; m++) {
                        acc[m][p][n] = mfma_f16(whi[p][m], xb[p], acc[m][p][n]);
                        acc[m][p][n] = mfma_f16(wlo[p][m], xb[p], acc[m][p][n]);
                    }
            }
            if (do_inc) {
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    const int an = un * AK + 4 * g + q;
                    const int ac = min(an, a.A - 1);
#pragma unroll
                    for (int p = 0; p < 2; p++) {
                        const float w2 = an < a.A ? s_w0sq[p][ac] : 0.f;
#pragma unroll
                        for (int n = 0; n < NT; n++) {
                            const float xr = ci8_byte(u.w[q * NT + n], 2 * p);
                            const float xi = ci8_byte(u.w[q * NT + n], 2 * p + 1);
                            inc[p][n] += (xr * xr + xi * xi) * w2;
                        }
                    }
                }
            }
        } else {
            const float* wl = reinterpret_cast<const float*>(lds_raw);
            float xv[2][3][NT];   // [p][c][n]: Re x, Im x and (three-product tiles) Re x + Im x
#pragma unroll
            for (int n = 0; n < NT; n++) {
                if constexpr (CI8) {
#pragma unroll
                    for (int p = 0; p < 2; p++)
#pragma unroll
                        for (int c = 0; c < 2; c++) xv[p][c][n] = ci8_byte(u.w[n], 2 * p + c);
                } else {
#pragma unroll
                    for (int p = 0; p < 2; p++)
#pragma unroll
                        for (int c = 0; c < 2; c++) xv[p][c][n] = __uint_as_float(u.w[4 * n + 2 * p + c]);
                }
#pragma unroll
                for (int p = 0; p < 2; p++) xv[p][2][n] = xv[p][0][n] + xv[p][1][n];
            }
            if constexpr (G3) {
                const int nunits = nsteps / 2;
#pragma unroll
                for (int p = 0; p < 2; p++)
#pragma unroll
                    for (int q = 0; q < 3; q++) {
                        const float wv = wl[((p * nunits + un) * 3 + q) * 64 + g * TILE + i];
#pragma unroll
                        for (int n = 0; n < NT; n++) acc[q][p][n] = mfma_f32(wv, xv[p][q][n], acc[q][p][n]);
                    }
            } else {
#pragma unroll
            for (int p = 0; p < 2; p++) {
#pragma unroll
                for (int c = 0; c < 2; c++) {
                    float wv[MW];
#pragma unroll
                    for (int m = 0; m < MW; m++)
                        wv[m] = wl[((((m * 2 + p) * nsteps) + 2 * un + c) * KG + g) * TILE + i];
#pragma unroll
                    for (int n = 0; n < NT; n++)
#pragma unroll
                        for (int m = 0; m < MW; m++) acc[m][p][n] = mfma_f32(wv[m], xv[p][c][n], acc[m][p][n]);
                }
            }
            }
            const int an = un * KG + g;
            if (do_inc) {
#pragma unroll
                for (int p = 0; p < 2; p++) {
                    const float w2 = an < a.A ? s_w0sq[p][min(an, a.A - 1)] : 0.f;
#pragma unroll
                    for (int n = 0; n < NT; n++)
                        inc[p][n] += (xv[p][0][n] * xv[p][0][n] + xv[p][1][n] * xv[p][1][n]) * w2;
                }
            }
        }
    };

    auto epilogue = [&](int c) {
        const int t0 = chunk_t0(c);
        const bool t_ok = t0 < a.T;
        // Opaque copy of f: keeps the output addresses from being hoisted out
        // of the chunk loop (they would pin dozens of registers).
        int fo = f;
        asm volatile("" : "+s"(fo));
        if constexpr (staged_epilogue<NT, F16X2, MW, EPI>()) {
            // Every wavefront of the workgroup runs the same chunks, so the
            // barriers below are reached by all four.
            float4* stg = reinterpret_cast<float4*>(lds_raw + a.wbytes);
            const int col = wave * TILE + i;                       // spectrum within the chunk
            const int S = a.T / 4;
            const int s = (cw0 + c * cstride) * (WAVES * TILE) + lane;
            float4* const base = reinterpret_cast<float4*>(a.y);
            float4* const sink = a.sink + lane;
#pragma unroll
            for (int h = 0; h < 2; h++) {
                __syncthreads();   // the previous half's (or chunk's) rows are stored
#pragma unroll
                for (int r = 0; r < 4; r += 2) {
                    float2 v[2][4];
#pragma unroll
                    for (int p = 0; p < 2; p++)
#pragma unroll
                        for (int n = 0; n < 4; n++)
                            v[p][n] = make_float2(acc[h][p][n][r] * unscale[p], acc[h][p][n][r + 1] * unscale[p]);
                    stage_beam(stg, (4 * g + r) / 2, col, v);
                }
                __syncthreads();
#pragma unroll
                for (int j = 0; j < 8; j++) {
                    const int row = wave * 8 + j;
                    const int b = b_wg + 8 * h + (row >> 2);
                    const float4 q = stg[row * kStageRow + lane];
                    float4* dst = (s < S && b < a.B)
                                      ? base + (((size_t)b * a.F + fo) * 4 + (row & 3)) * S + s
                                      : sink;
                    store4<NTS>(dst, make_float2(q.x, q.y), make_float2(q.z, q.w));
                }
            }
            return;
        }
        if constexpr (G3) {
            // Rows are beams: lane group g holds beams 4g .. 4g + 3.
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int b = b_wg + 4 * g + r;
                float2 v[2][NT];
#pragma unroll
                for (int p = 0; p < 2; p++)
#pragma unroll
                    for (int n = 0; n < NT; n++) {
                        const float p1 = acc[0][p][n][r], p2 = acc[1][p][n][r], p3 = acc[2][p][n][r];
                        v[p][n] = make_float2(p1 - p2, p3 - p1 - p2);
                    }
                store_beam<NT, NTS, EPI>(a, min(b, a.B - 1), fo, t0, v, t_ok && b < a.B);
            }
        } else
#pragma unroll
        for (int m = 0; m < MW; m++) {
#pragma unroll
            for (int r = 0; r < 4; r += 2) {
                const int b = b_wg + m * (TILE / 2) + (4 * g + r) / 2;
                float2 v[2][NT];
#pragma unroll
                for (int p = 0; p < 2; p++)
#pragma unroll
                    for (int n = 0; n < NT; n++) {
                        if constexpr (F16X2)
                            v[p][n] = make_float2(acc[m][p][n][r] * unscale[p], acc[m][p][n][r + 1] * unscale[p]);
                        else
                            v[p][n] = make_float2(acc[m][p][n][r], acc[m][p][n][r + 1]);
                    }
                store_beam<NT, NTS, EPI>(a, min(b, a.B - 1), fo, t0, v, t_ok && b < a.B);
            }
        }
        // The incoherent beam's stores are issued by every workgroup (to the
        // sink where there is none) with plain stores: the sink stays in L2.
        if (epi_inc<EPI>(a)) {
            if (do_inc) {
#pragma unroll
                for (int p = 0; p < 2; p++)
#pragma unroll
                    for (int n = 0; n < NT; n++) {
                        inc[p][n] += __shfl_xor(inc[p][n], 16, 64);
                        inc[p][n] += __shfl_xor(inc[p][n], 32, 64);
                    }
            }
            float2 v[2][NT];
#pragma unroll
            for (int p = 0; p < 2; p++)
#pragma unroll
                for (int n = 0; n < NT; n++)
                    v[p][n] = make_float2(a.inc_sqrt ? sqrtf(inc[p][n]) : inc[p][n], 0.f);
            store_beam<NT, false, EPI>(a, a.B, fo, t0, v, do_inc && t_ok && g == 0);
        }
    };

    // ---- Main loop: D units per step, loads run D units ahead. -------------
    // A ring slot is consumed before it is reloaded, so every slot keeps its
    // registers across iterations (no copies, no drain of the loads in
    // flight at the loop latch).
    auto group = [&](int ub, int j) {   // units ub..ub+D-1; j = global unit index
#pragma unroll
        for (int d = 0; d < D; d++) {
            if (ub + d < n_units) compute(ring[d], ub + d);
            load_unit(j + d + D, ring[d]);
        }
    };
    // Chunk 0, then every later chunk with its first group peeled: the peeled
    // group follows the previous epilogue's stores in straight-line code, so
    // its waits count those stores (in flight after the awaited loads) instead
    // of draining them, and the steady-state loop only sees its own loads.
    for (int ub = 0; ub < upad; ub += D) group(ub, ub);
    epilogue(0);
    zero_acc();
    for (int c = 1; c < nch; c++) {
        group(0, c * upad);
        for (int ub = D; ub < upad; ub += D) group(ub, c * upad + ub);
        epilogue(c);
        zero_acc();
    }
}

// One block: blockIdx = (first chunk, f, beam group); chunks interleave.
template <int NT, int MW, bool CI8, bool F16X2, int D, bool NTS, int EPI>
__global__ __launch_bounds__(256, (staged_epilogue<NT, F16X2, MW, EPI>() ? 3 : 2)) void beamform_kernel(BfArgs a) {
    extern __shared__ __align__(16) unsigned char lds_raw[];
    __shared__ float s_unscale[2];
    __shared__ float s_w0sq[2][128];
    __shared__ unsigned int s_max[2];
    bf_body<NT, MW, CI8, F16X2, D, NTS, EPI, (!F16X2 && MW == 2)>(a, blockIdx.y, blockIdx.z, blockIdx.x, lds_raw,
                                                                s_unscale, s_w0sq, s_max, gridDim.x);
}

static int env_int(const char* name, int dflt) {
    const char* v = getenv(name);
    return v ? atoi(v) : dflt;
}

// Bytes of the packed phasor A operands in LDS.  FP32 with two beam tiles
// stores three products per antenna group (Wr, Wi, Wr + Wi over 16 beams)
// instead of the 2 x 2 embedding rows.
static size_t operand_bytes(bool f16x2, int mw, int nsteps, bool g3) {
    return f16x2 ? (size_t)mw * 2 * nsteps * 2 * 64 * 16
                 : (g3 ? (size_t)3 * nsteps * 64 * 4 : (size_t)mw * 2 * nsteps * 64 * 4);
}

constexpr size_t kSinkBytes = 64 * sizeof(float4);

static int nsteps_for(bool f16x2, int A) {
    return f16x2 ? (A + 4 * KG - 1) / (4 * KG) : 2 * ((A + KG - 1) / KG);
}

static int groups_for(int B, int mw) { return (B + mw * (TILE / 2) - 1) / (mw * (TILE / 2)); }

// Chunks per workgroup.  The resident workgroups (occupancy x CUs) run the
// grid in rounds, and a round takes cpw chunk times, so the estimated time is
// ceil(workgroups / resident) * cpw chunk times.  The chosen cpw is the
// largest one within 2% of the best estimate: fewer, longer workgroups pack
// their phasors fewer times.  Measured at the production shapes: the Mode-B
// kernels (3 per CU) get one round of 768 workgroups, cpw = 8; the Mode-A
// kernel (4 per CU) gets three rounds of 1024, cpw = 4 (146 against 150 us
// with one round of 768).
static int pick_chunks_per_wg(const void* fn, size_t lds, int n_tchunks, long per_chunk) {
    // Occupancy per (kernel, LDS bytes), cached in a fixed table (no heap
    // allocation on the launch path; past 64 entries it is recomputed).
    struct Occ { const void* fn; size_t lds; int occ; };
    static std::mutex mu;
    static Occ cache[64];
    static int n_cached = 0;
    int occ = 0;
    {
        std::lock_guard<std::mutex> lk(mu);
        for (int k = 0; k < n_cached; k++)
            if (cache[k].fn == fn && cache[k].lds == lds) occ = cache[k].occ;
        if (!occ) {
            if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, fn, 64 * WAVES, lds) != hipSuccess || occ < 1)
                occ = 3;
            if (n_cached < 64) cache[n_cached++] = {fn, lds, occ};
        }
    }
    const long resident = (long)occ * device_cus();
    const int cmax = std::min(n_tchunks, 4096);
    auto est = [&](int c) {
        const long wgs = (long)((n_tchunks + c - 1) / c) * per_chunk;
        return ((wgs + resident - 1) / resident) * c;
    };
    long best_t = est(1);
    for (int c = 2; c <= cmax; c++) best_t = std::min(best_t, est(c));
    int best = 1;
    for (int c = 1; c <= cmax; c++)
        if (est(c) * 100 <= best_t * 102) best = c;
    return best;
}

template <int NT, int MW, bool CI8, bool F16X2, int D, bool NTS, int EPI = 0>
static int launch_bf(const BfArgs& args, unsigned char* ws, size_t ws_bytes, hipStream_t s) {
    BfArgs a = args;
    a.nsteps = nsteps_for(F16X2, a.A);
    a.wbytes = operand_bytes(F16X2, MW, a.nsteps, !F16X2 && MW == 2);
    // A operands, then the staged phasors of both pols (packed in the kernel).
    size_t lds = a.wbytes + (size_t)2 * MW * (TILE / 2) * a.A * sizeof(float2);
    if (staged_epilogue<NT, F16X2, MW, EPI>()) lds = std::max(lds, a.wbytes + kStageBytes);
    const int span = WAVES * TILE * NT;
    const int groups = groups_for(a.B, MW);
    // Device scratch: only the store sink of idle lanes (the phasors are
    // packed into LDS by the kernel).
    const size_t need = kSinkBytes;
    unsigned char* sink_mem = nullptr;
    if (ws) {
        if (ws_bytes < need) return fail("beamformer: workspace too small");
        if ((uintptr_t)ws % 256) return fail("beamformer: workspace must be 256-byte aligned");
        sink_mem = ws;
    } else {
        const int rc = library_workspace(WS_BEAMFORMER_SINK, need, s, &sink_mem);
        if (rc) return rc;
    }
    a.sink = reinterpret_cast<float4*>(sink_mem);
    a.n_tchunks = (a.T + span - 1) / span;
    const void* fn = reinterpret_cast<const void*>(&beamform_kernel<NT, MW, CI8, F16X2, D, NTS, EPI>);
    if (lds > 150 * 1024) return fail("beamformer: too many antennas for the LDS phasor stage");
    if (lds > 65536) {
        hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) {
            set_error(std::string("beamformer: LDS attribute: ") + hipGetErrorString(e));
            return BLADE_HIP_ERROR;
        }
    }
    int cpw = env_int("BLADE_BF_CHUNKS_PER_WG", 0);
    if (cpw <= 0) cpw = pick_chunks_per_wg(fn, lds, a.n_tchunks, (long)a.F * groups);
    cpw = std::min(cpw, a.n_tchunks);
    a.chunks_per_wg = cpw;
    dim3 grid((unsigned)((a.n_tchunks + cpw - 1) / cpw), (unsigned)a.F, (unsigned)groups);
    hipLaunchKernelGGL((beamform_kernel<NT, MW, CI8, F16X2, D, NTS, EPI>), grid, dim3(64 * WAVES), lds, s, a);
    return check_launch("beamformer");
}

// Variant table: (time samples per lane, beam tiles per wave) by precision.
// The production epilogues (FFT-4 + polarizer; plain beams with or without
// the incoherent beam) of the NT = 4 variants are compiled as fixed
// straight-line code: with every path issuing the same stores, the loads that
// follow an epilogue wait for themselves only, not for the beam stores.
// Stores: nontemporal where the 16 lanes of a group write contiguous 16-byte
// records (the fixed epilogues); in the generic epilogue a store instruction
// writes 16 bytes of every 32 or 64 and ordinary stores let L2 merge the lines
// before they go out.

template <int NT, int MW, bool CI8, bool F16X2, int D>
static int launch_epi(const BfArgs& a, int epi, unsigned char* ws, size_t wsb, hipStream_t s) {
    if constexpr (NT == 4) {
        if (epi == EPI_MODEB) return launch_bf<NT, MW, CI8, F16X2, D, true, EPI_MODEB>(a, ws, wsb, s);
        // The rate-1 epilogues map a wave onto TILE * NT consecutive samples
        // (runs of one): T must not cut a wave.
        if (a.T % (TILE * NT) == 0) {
            if (epi == EPI_PLAIN) return launch_bf<NT, MW, CI8, F16X2, D, true, EPI_PLAIN>(a, ws, wsb, s);
            if (epi == EPI_PLAIN_INC)
                return launch_bf<NT, MW, CI8, F16X2, D, true, EPI_PLAIN_INC>(a, ws, wsb, s);
        }
    }
    return launch_bf<NT, MW, CI8, F16X2, D, false, 0>(a, ws, wsb, s);
}

template <bool CI8, bool F16X2>
static int dispatch(const BfArgs& a, int nt, int mw, int epi, unsigned char* ws, size_t wsb, hipStream_t s) {
    constexpr int D = (F16X2 || !CI8) ? 2 : 4;
    if (nt == 4)
        return mw == 2 ? launch_epi<4, 2, CI8, F16X2, D>(a, epi, ws, wsb, s)
                       : launch_epi<4, 1, CI8, F16X2, D>(a, epi, ws, wsb, s);
    return mw == 2 ? launch_epi<1, 2, CI8, F16X2, D>(a, epi, ws, wsb, s)
                   : launch_epi<1, 1, CI8, F16X2, D>(a, epi, ws, wsb, s);
}

// ---------------------------------------------------------------------------
// Batched blocks of uneven shapes in one persistent launch.
//
// Every block (its own A, F, T, B and buffers) is cut into work items of
// (chunk range, f, beam group); a fixed grid of workgroups walks the
// concatenated item list with a static stride, so every workgroup exits after
// its last item.  The block table travels in the kernel arguments (no copy,
// capturable in a hipGraph).  Rate-1 beams, no incoherent beam.
// ---------------------------------------------------------------------------
constexpr int kBatchMax = 32;

struct BatchBlock {
    const void* x;
    const float2* w;
    float2* y;
    size_t wbytes;          // bytes of the phasor A operands in LDS
    int A, F, T, B;
    int nsteps, n_tchunks, cpw, groups, mw;
    int item_base;          // first work item of the block
};

struct BatchArgs {
    int n_blocks, total_items;
    float4* sink;
    BatchBlock blk[kBatchMax];
};

__device__ __forceinline__ BfArgs batch_block_args(const BatchArgs& ba, int k) {
    const BatchBlock& b = ba.blk[k];
    BfArgs a;
    a.x = b.x; a.w = b.w; a.y = b.y;
    a.A = b.A; a.F = b.F; a.T = b.T; a.B = b.B;
    a.incoherent = 0; a.inc_sqrt = 0; a.rate = 1; a.polarize = 0;
    a.nsteps = b.nsteps; a.n_tchunks = b.n_tchunks; a.chunks_per_wg = b.cpw;
    a.wbytes = b.wbytes;
    a.sink = ba.sink;
    return a;
}

template <int NT, bool F16X2, int D>
__global__ __launch_bounds__(256, 2) void beamform_batched_kernel(BatchArgs ba) {
    extern __shared__ __align__(16) unsigned char lds_raw[];
    __shared__ float s_unscale[2];
    __shared__ float s_w0sq[2][128];
    __shared__ unsigned int s_max[2];
    // A work item is a (channel, chunk range) of one block.  FP32: its beam
    // groups run back to back in the item's workgroup (fewer, longer items:
    // 5.3 against 5.6 ms on the seeded 32-block batch).  Split FP16: one
    // item per beam group (4.5 against 4.6 ms).
    constexpr bool GROUPS_IN_ITEM = !F16X2;
    for (int item = blockIdx.x; item < ba.total_items; item += gridDim.x) {
        int k = 0;
        while (k + 1 < ba.n_blocks && item >= ba.blk[k + 1].item_base) k++;
        const BatchBlock& b = ba.blk[k];
        int local = item - b.item_base;
        int gz0 = 0, gz1 = b.groups;
        if constexpr (!GROUPS_IN_ITEM) {
            gz0 = local % b.groups;
            gz1 = gz0 + 1;
            local /= b.groups;
        }
        const int f = local % b.F;
        const int tr = local / b.F;
        const BfArgs a = batch_block_args(ba, k);
        for (int gz = gz0; gz < gz1; gz++) {
            __syncthreads();   // the previous group's (or item's) readers of LDS are done
            // FP32 16-beam tiles take the three-product form, as in the
            // single-block kernel.  Each group packs its (channel, beam
            // group)'s phasors into LDS itself, as the single-block kernel
            // does.
            if (b.mw == 2)
                bf_body<NT, 2, true, F16X2, D, true, EPI_PLAIN, !F16X2>(a, f, gz, tr * b.cpw, lds_raw, s_unscale,
                                                                        s_w0sq, s_max);
            else
                bf_body<NT, 1, true, F16X2, D, true, EPI_PLAIN>(a, f, gz, tr * b.cpw, lds_raw, s_unscale,
                                                                s_w0sq, s_max);
        }
    }
}

struct BatchPlan {
    BatchArgs args;
    size_t ws_bytes;        // the store sink
    size_t lds;
    int grid;
};

// Plans up to kBatchMax blocks; the workspace pointer may be null (sizing).
static int plan_batch(bool f16x2, const blade_beamform_block* blocks, int n,
                      const blade_batch_tuning* tuning, unsigned char* ws, BatchPlan* plan) {
    constexpr int NT = 4;
    const int span = WAVES * TILE * NT;
    const int mw_force = tuning ? tuning->beam_tiles : 0;
    const int cpw_req = tuning ? tuning->chunks_per_item : 0;
    const int wpc = (tuning && tuning->workgroups_per_cu > 0) ? tuning->workgroups_per_cu : 3;
    if (tuning && tuning->time_samples_per_lane != 0 && tuning->time_samples_per_lane != NT)
        return fail("batched beamformer: time_samples_per_lane must be 0 or 4");
    if (mw_force < 0 || mw_force > 2) return fail("batched beamformer: beam_tiles must be 0, 1 or 2");
    BatchArgs& ba = plan->args;
    std::memset(&ba, 0, sizeof(ba));
    ba.n_blocks = n;
    size_t lds = 0;
    long items = 0;
    // Auto chunk count: about wpc * CUs * 2 items in total (two rounds).
    long all_chunks = 0;
    for (int k = 0; k < n; k++) {
        const blade_beamform_block& bk = blocks[k];
        // items: (channel, chunk range), and the beam group for split FP16
        all_chunks += (long)ceil_div(bk.T, span) * bk.F *
                      (f16x2 ? groups_for((int)bk.B, mw_force ? mw_force : (bk.B > 8 ? 2 : 1)) : 1);
    }
    const long target = 2L * wpc * device_cus();
    const int cpw_auto = (int)std::max(1L, (all_chunks + target - 1) / target);
    for (int k = 0; k < n; k++) {
        const blade_beamform_block& bk = blocks[k];
        if (bk.A == 0 || bk.B == 0 || bk.F == 0 || bk.T == 0)
            return fail("batched beamformer: every block needs A, F, T, B > 0");
        if (bk.A > 4096 || bk.F > 65535 || bk.T > (1u << 30) || bk.B > 4096)
            return fail("batched beamformer: block shape too large");
        if (bk.T % (TILE * NT)) return fail("batched beamformer: T must be a multiple of 64");
        if (!bk.input || !bk.phasors || !bk.output) return fail("batched beamformer: null pointer");
        if ((uintptr_t)bk.input % 16 || (uintptr_t)bk.phasors % 16 || (uintptr_t)bk.output % 16)
            return fail("batched beamformer: buffers must be 16-byte aligned");
        BatchBlock& b = ba.blk[k];
        b.mw = mw_force ? mw_force : (bk.B > 8 ? 2 : 1);
        b.x = bk.input;
        b.w = static_cast<const float2*>(bk.phasors);
        b.y = static_cast<float2*>(bk.output);
        b.A = (int)bk.A; b.F = (int)bk.F; b.T = (int)bk.T; b.B = (int)bk.B;
        b.nsteps = nsteps_for(f16x2, b.A);
        b.wbytes = operand_bytes(f16x2, b.mw, b.nsteps, !f16x2 && b.mw == 2);
        b.groups = groups_for(b.B, b.mw);
        b.n_tchunks = (int)ceil_div(bk.T, span);
        b.cpw = std::min(b.n_tchunks, cpw_req > 0 ? cpw_req : cpw_auto);
        b.item_base = (int)items;
        items += (long)ceil_div(b.n_tchunks, b.cpw) * b.F * (f16x2 ? b.groups : 1);
        // A operands, then the staged phasors of both pols and the group's beams
        lds = std::max(lds, b.wbytes + (size_t)2 * b.mw * (TILE / 2) * b.A * sizeof(float2));
    }
    if (items > (1L << 30)) return fail("batched beamformer: too many work items");
    if (lds > 150 * 1024) return fail("batched beamformer: too many antennas");
    ba.total_items = (int)items;
    ba.sink = ws ? reinterpret_cast<float4*>(ws) : nullptr;
    plan->ws_bytes = kSinkBytes;
    plan->lds = lds;
    plan->grid = (int)std::min<long>(items, (long)wpc * device_cus());
    return BLADE_SUCCESS;
}

template <bool F16X2>
static int launch_batch(BatchPlan& plan, hipStream_t s) {
    constexpr int D = F16X2 ? 2 : 4;
    const void* fn = reinterpret_cast<const void*>(&beamform_batched_kernel<4, F16X2, D>);
    if (plan.lds > 65536) {
        hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)plan.lds);
        if (e != hipSuccess) return fail(std::string("batched beamformer: LDS attribute: ") + hipGetErrorString(e));
    }
    hipLaunchKernelGGL((beamform_batched_kernel<4, F16X2, D>), dim3((unsigned)plan.grid), dim3(64 * WAVES),
                       plan.lds, s, plan.args);
    return check_launch("batched beamformer");
}

static int run_batched(int precision, const blade_beamform_block* blocks, uint64_t n,
                       const blade_batch_tuning* tuning, unsigned char* ws, size_t ws_bytes,
                       uint64_t* size_out, hipStream_t s) {
    if (precision != BLADE_PREC_F32 && precision != BLADE_PREC_F16X2)
        return fail("batched beamformer: unknown precision");
    if (n > 0 && !blocks) return fail("batched beamformer: null block table");
    const bool f16 = precision == BLADE_PREC_F16X2;
    size_t total = 0;
    for (uint64_t k0 = 0; k0 < n; k0 += kBatchMax) {
        const int cnt = (int)std::min<uint64_t>(kBatchMax, n - k0);
        BatchPlan plan;
        unsigned char* base = ws ? ws + total : nullptr;
        if (base && (uintptr_t)base % 256) return fail("batched beamformer: workspace must be 256-byte aligned");
        int rc = plan_batch(f16, blocks + k0, cnt, tuning, base, &plan);
        if (rc) return rc;
        const size_t need = (plan.ws_bytes + 255) / 256 * 256;
        if (!size_out) {
            if (!ws || ws_bytes < total + need) return fail("batched beamformer: workspace too small");
            rc = f16 ? launch_batch<true>(plan, s) : launch_batch<false>(plan, s);
            if (rc) return rc;
        }
        total += need;
    }
    if (size_out) *size_out = total;
    return BLADE_SUCCESS;
}

static int run_beamformer(int in_type, int precision, const void* input, const void* phasors,
                          void* output, uint64_t A, uint64_t F, uint64_t T, uint64_t P,
                          uint64_t B, int incoherent, int incoherent_sqrt, uint64_t rate,
                          int polarize, unsigned char* ws, size_t wsb, hipStream_t s) {
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
    if (rate > 1 && T % 4) return fail("beamformer chain: T must be divisible by 4");
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
    a.n_tchunks = 0;
    a.chunks_per_wg = 1;
    a.wbytes = 0;
    a.sink = nullptr;
    const int mw = B > 8 ? 2 : 1;
    const int epi = (int)rate | (polarize ? 8 : 0) | (incoherent ? 16 : 0);
    if (precision == BLADE_PREC_F16X2) return dispatch<true, true>(a, T % 4 == 0 ? 4 : 1, mw, epi, ws, wsb, s);
    if (in_type == BLADE_CI8) {
        return dispatch<true, false>(a, T % 4 == 0 ? 4 : 1, mw, epi, ws, wsb, s);
    }
    return dispatch<false, false>(a, T % 4 == 0 ? 4 : 1, mw, epi, ws, wsb, s);
}

// Scratch of blade_beamform_chain_noalloc: the store sink of idle lanes as
// a file-scope device array, kSinkBytes (1 KiB) on every device for the life
// of the process.  In host code `scratch` names a host-side stand-in; its
// device address, which differs per device, is looked up once per device
// with that device current, and only that address reaches the kernels.
// Calls that overlap on one device write the same array at the same time;
// it is a write-only sink whose contents nothing reads, so their results do
// not depend on each other.
static __device__ __align__(256) unsigned char scratch[kSinkBytes];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mutex;
static unsigned char* g_scratch_ptr[kMaxDevices] = {};

static int static_scratch(unsigned char** out) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) {
        set_error(std::string("beamformer scratch: hipGetDevice: ") + hipGetErrorString(e));
        return BLADE_HIP_ERROR;
    }
    if (dev < 0 || dev >= kMaxDevices) return fail("beamformer scratch: device index above 63");
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            set_error(std::string("beamformer scratch: hipGetSymbolAddress: ") +
                      (e != hipSuccess ? hipGetErrorString(e) : "null address"));
            return BLADE_HIP_ERROR;
        }
        g_scratch_ptr[dev] = static_cast<unsigned char*>(p);
    }
    *out = g_scratch_ptr[dev];
    return BLADE_SUCCESS;
}

}  // namespace bl

using namespace bl;

extern "C" int blade_beamform_chain_noalloc(int in_type, int precision, const void* input,
                                            const void* phasors, void* output, uint64_t A,
                                            uint64_t F, uint64_t T, uint64_t P, uint64_t B,
                                            int incoherent, int incoherent_sqrt, uint64_t rate,
                                            int polarize, blade_stream_t stream) {
    const uint64_t need = blade_beamformer_workspace_size(precision, A, F, B);
    if (need > kSinkBytes)
        return fail("beamformer (no allocation): the call needs " + std::to_string(need) +
                    " bytes of scratch, above the static " + std::to_string(kSinkBytes));
    unsigned char* ws = nullptr;
    const int rc = static_scratch(&ws);
    if (rc) return rc;
    return run_beamformer(in_type, precision, input, phasors, output, A, F, T, P, B, incoherent,
                          incoherent_sqrt, rate, polarize, ws, kSinkBytes,
                          reinterpret_cast<hipStream_t>(stream));
}

extern "C" int blade_beamformer_ata(int in_type, int precision, const void* input,
                                    const void* phasors, void* output, uint64_t A, uint64_t F,
                                    uint64_t T, uint64_t P, uint64_t B, int incoherent,
                                    int incoherent_sqrt, blade_stream_t stream) {
    return run_beamformer(in_type, precision, input, phasors, output, A, F, T, P, B, incoherent,
                          incoherent_sqrt, 1, 0, nullptr, 0, reinterpret_cast<hipStream_t>(stream));
}

extern "C" int blade_beamform_chain(int in_type, int precision, const void* input,
                                    const void* phasors, void* output, uint64_t A, uint64_t F,
                                    uint64_t T, uint64_t P, uint64_t B, int incoherent,
                                    int incoherent_sqrt, uint64_t rate, int polarize,
                                    blade_stream_t stream) {
    return run_beamformer(in_type, precision, input, phasors, output, A, F, T, P, B, incoherent,
                          incoherent_sqrt, rate, polarize, nullptr, 0,
                          reinterpret_cast<hipStream_t>(stream));
}

extern "C" uint64_t blade_beamformer_workspace_size(int precision, uint64_t A, uint64_t F,
                                                    uint64_t B) {
    (void)precision; (void)A; (void)F; (void)B;   // the same for every shape
    return kSinkBytes;
}

extern "C" int blade_beamform_chain_ws(int in_type, int precision, const void* input,
                                       const void* phasors, void* output, uint64_t A, uint64_t F,
                                       uint64_t T, uint64_t P, uint64_t B, int incoherent,
                                       int incoherent_sqrt, uint64_t rate, int polarize,
                                       void* workspace, uint64_t workspace_bytes,
                                       blade_stream_t stream) {
    if (!workspace) return fail("beamformer: null workspace");
    return run_beamformer(in_type, precision, input, phasors, output, A, F, T, P, B, incoherent,
                          incoherent_sqrt, rate, polarize, static_cast<unsigned char*>(workspace),
                          (size_t)workspace_bytes, reinterpret_cast<hipStream_t>(stream));
}

extern "C" uint64_t blade_beamform_batched_workspace_size(int precision,
                                                         const blade_beamform_block* blocks,
                                                         uint64_t n_blocks,
                                                         const blade_batch_tuning* tuning) {
    uint64_t size = 0;
    if (run_batched(precision, blocks, n_blocks, tuning, nullptr, 0, &size, nullptr)) return 0;
    return size;
}

extern "C" int blade_beamform_batched(int precision, const blade_beamform_block* blocks,
                                      uint64_t n_blocks, const blade_batch_tuning* tuning,
                                      void* workspace, uint64_t workspace_bytes,
                                      blade_stream_t stream) {
    return run_batched(precision, blocks, n_blocks, tuning, static_cast<unsigned char*>(workspace),
                       (size_t)workspace_bytes, nullptr, reinterpret_cast<hipStream_t>(stream));
}
