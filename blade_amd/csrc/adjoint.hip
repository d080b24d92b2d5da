// Adjoint of the ATA beamformer on CDNA4 matrix cores, and the fused
// Stokes-I calibration gradient.
//
// For y[b,f,t,p] = sum_a w[b,a,f,p] x[a,f,t,p] (the beamformer of
// reference src/modules/beamformer/beamformer.cu:13-79) and an upstream
// gradient gy = dL/dconj(y), the phasor gradient is the contraction over time
//     gw[b,a,f,p] = dL/dconj(w[b,a,f,p]) = sum_t gy[b,f,t,p] conj(x[a,f,t,p]).
// Per (f, p) this is a (B x T)(T x A) product, embedded as the real GEMM
//     [Re gw; Im gw] = [[gy_r, gy_i], [gy_i, -gy_r]] . [x_r; x_i]
// with rows (beam, re/im), K = (time, component) and columns = antennas, on
// v_mfma_f32_16x16x4f32 (one 16-row tile = 8 beams).
//
// Fused Stokes-I mode (stokes_grad_kernel): L = -sum_b lambda_b
// sum_{f,t,p} |y_b|^2, so gy = -lambda_b y_b.  Each workgroup recomputes its
// tile of y from the CI8 voltages staged in LDS and runs the adjoint on the
// same voltages: the beams are never written to memory.  Both contractions
// are taken transposed, with rows = time for the forward product
//     y^T[t][(b,c)] = sum_(a,c') x[a][t] . Wemb[(a,c')][(b,c)]
// and rows = antennas for the adjoint
//     gw^T[a][(b,c)] = sum_(t,c') x[a][t] . G[(t,c')][(b,c)],
// so the forward accumulator a lane holds (times 4g..4g+3 of column i) is
// exactly the adjoint's B operand for K = those four times: gy never passes
// through LDS.  G's second component block (c' = imaginary part of x) is the
// neighbouring column's value, one lane swap away.  With at most 4 beams the
// FP32 form runs stokes_grad_b4_kernel instead, on 4x4-block matrix
// instructions whose columns are exactly the beams.
//
// Plain adjoint (adjoint_kernel, gy from memory): rows (beam, re/im),
// K = (time, component), columns = antennas.
//
// Work split: blockIdx = (time split, f, beam tile of 8).  Partial gradients
// are summed across waves in LDS and across time splits by
// adjoint_reduce_kernel, in a fixed order (deterministic, no global atomics).
#include "common.hh"

#include <algorithm>
#include <cstring>

namespace bl {

constexpr int ADJ_TC = 128;        // time samples per LDS tile
constexpr int ADJ_WAVES = 4;
constexpr int ADJ_MAXAT = 8;       // antenna tiles of 16: A <= 128
constexpr int ADJ_MAXB = 64;       // beams of the fused loss
constexpr int ADJ_GYS = 17;        // s_gy row stride: 16 values + 1 word of padding
constexpr int ADJ_BATCH = 8;       // global loads in flight per thread while staging (16 costs a wave of occupancy)
constexpr int SG_TC = 64;          // stokes_grad_kernel: time samples per tile, 16 per wave
constexpr int SG_XPAD = 4;         // its LDS voltage row stride is apad + 4 words (see sg_xs)

// Row stride (words) of the LDS voltage tile s_x[t][antenna]: the padded
// antenna count plus one, so the 16 lanes of a group, which read 16
// consecutive times of one antenna, hit 16 different banks (a stride of a
// multiple of 16 words would put them in two).
__host__ __device__ constexpr int adj_xs(int apad) { return apad + 1; }

struct AdjArgs {
    const uint32_t* x;       // [A,F,T] dwords: CI8 (re0, im0, re1, im1)
    const float2* w;         // [B,A,F,1,P] phasors (fused mode)
    const float2* gy;        // [B,F,T,P] upstream gradient (adjoint mode)
    float* partial;          // [nsplit][F][P][btiles][16][apad]
    float* loss_part;        // [nsplit][F][btiles]
    int A, F, T, B;
    int apad, n_at, nsteps;  // padded antennas, antenna tiles, forward K steps
    int tsplit;              // time samples per workgroup (multiple of ADJ_TC)
    int btiles;
    float lambda[ADJ_MAXB];  // beam weights of the fused loss
};

__device__ __forceinline__ floatx4 adj_mfma(float a, float b, floatx4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

// The value of lane l ^ 1 (DPP quad permutation [1, 0, 3, 2]).
__device__ __forceinline__ float swap_pairs(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0xB1, 0xF, 0xF, false));
}

__device__ __forceinline__ float ci8_component(uint32_t d, int idx) {
    return (float)(int8_t)((d >> (8 * idx)) & 0xFFu);
}

__global__ __launch_bounds__(256) void adjoint_kernel(AdjArgs a) {
    extern __shared__ __align__(16) unsigned char lds_raw[];
    const int apad = a.apad;
    const int xs = adj_xs(apad);
    uint32_t* s_x = reinterpret_cast<uint32_t*>(lds_raw);              // [TC][xs]
    float* s_gy = reinterpret_cast<float*>(s_x + ADJ_TC * xs);          // [TC][ADJ_GYS]
    float* s_G = s_gy + ADJ_TC * ADJ_GYS;                               // [2][16][apad]

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int i = lane & 15;
    const int g = lane >> 4;
    const int sp = blockIdx.x;
    const int f = blockIdx.y;
    const int bt = blockIdx.z;
    const int b0 = bt * 8;
    const int A = a.A, F = a.F, T = a.T, B = a.B;
    const int t_begin = sp * a.tsplit;
    const int t_end = min(T, t_begin + a.tsplit);

    for (int e = tid; e < 2 * 16 * apad; e += blockDim.x) s_G[e] = 0.f;

    floatx4 acc2[2][ADJ_MAXAT];
#pragma unroll
    for (int p = 0; p < 2; p++)
#pragma unroll
        for (int at = 0; at < ADJ_MAXAT; at++) acc2[p][at] = floatx4{0.f, 0.f, 0.f, 0.f};

    for (int t0 = t_begin; t0 < t_end; t0 += ADJ_TC) {
        __syncthreads();   // the previous tile's readers are done
        // Voltages of every antenna for this tile, s_x[t][a] (zero-padded).
        // Loads in batches of ADJ_BATCH per thread, all issued before any is
        // used, so a batch costs one memory round trip.
        const int n_x = ADJ_TC * apad;
        for (int e0 = tid; e0 < n_x; e0 += ADJ_BATCH * ADJ_WAVES * 64) {
            uint32_t v[ADJ_BATCH];
#pragma unroll
            for (int k = 0; k < ADJ_BATCH; k++) {
                const int e = e0 + k * ADJ_WAVES * 64;
                const int an = e / ADJ_TC, tt = t0 + (e - an * ADJ_TC);
                v[k] = (e < n_x && an < A && tt < t_end) ? a.x[((size_t)an * F + f) * T + tt] : 0u;
            }
#pragma unroll
            for (int k = 0; k < ADJ_BATCH; k++) {
                const int e = e0 + k * ADJ_WAVES * 64;
                const int an = e / ADJ_TC, t = e - an * ADJ_TC;
                if (e < n_x) s_x[t * xs + an] = v[k];
            }
        }
        __syncthreads();

#pragma unroll
        for (int p = 0; p < 2; p++) {
#pragma unroll
            for (int e = tid; e < ADJ_TC * 8; e += ADJ_WAVES * 64) {
                const int bb = e / ADJ_TC, t = e - bb * ADJ_TC;
                const int b = b0 + bb, tt = t0 + t;
                float2 v = make_float2(0.f, 0.f);
                if (b < B && tt < t_end) v = a.gy[(((size_t)b * F + f) * T + tt) * 2 + p];
                s_gy[t * ADJ_GYS + 2 * bb] = v.x;
                s_gy[t * ADJ_GYS + 2 * bb + 1] = v.y;
            }
            __syncthreads();

            // Adjoint: K = (time, component), 4 samples per MFMA pair.
#pragma unroll 2
            for (int t4 = 0; t4 < 8; t4++) {
                const int t = wave * 32 + t4 * 4 + g;
                const float gr = s_gy[t * ADJ_GYS + 2 * (i >> 1)];
                const float gi = s_gy[t * ADJ_GYS + 2 * (i >> 1) + 1];
                const float a0 = (i & 1) == 0 ? gr : gi;     // multiplies x_r
                const float a1 = (i & 1) == 0 ? gi : -gr;    // multiplies x_i
#pragma unroll
                for (int at = 0; at < ADJ_MAXAT; at++) {
                    if (at < a.n_at) {
                        const uint32_t d = s_x[t * xs + at * 16 + i];
                        acc2[p][at] = adj_mfma(a0, ci8_component(d, 2 * p), acc2[p][at]);
                        acc2[p][at] = adj_mfma(a1, ci8_component(d, 2 * p + 1), acc2[p][at]);
                    }
                }
            }
            __syncthreads();   // s_gy is rewritten by the next pol
        }
    }

    // Sum the four waves' partial gradients in LDS, then write this
    // workgroup's [2][16][apad] block.
#pragma unroll
    for (int p = 0; p < 2; p++)
#pragma unroll
        for (int at = 0; at < ADJ_MAXAT; at++) {
            if (at < a.n_at) {
#pragma unroll
                for (int r = 0; r < 4; r++)
                    atomicAdd(&s_G[(p * 16 + 4 * g + r) * apad + at * 16 + i], acc2[p][at][r]);
            }
        }
    __syncthreads();
    float* out = a.partial + ((((size_t)sp * F + f) * 2) * a.btiles + bt) * 16 * apad;
    const size_t pstride = (size_t)a.btiles * 16 * apad;   // between pols
    for (int e = tid; e < 2 * 16 * apad; e += blockDim.x) {
        const int p = e / (16 * apad), r = e - p * 16 * apad;
        out[p * pstride + r] = s_G[e];
    }
}

// LDS row stride (words) of stokes_grad_kernel's voltage tile s_x[t][antenna]:
// apad + 4 puts both read patterns on distinct banks, the forward's (16 lanes
// read 16 consecutive times of one antenna; lane groups take adjacent
// antennas) and the adjoint's (16 lanes read 16 consecutive antennas at one
// time; lane groups take times 4 apart).
__host__ __device__ constexpr int sg_xs(int apad) { return apad + SG_XPAD; }

// Fused Stokes-I loss and phasor gradient for A <= 16 NAT antennas.  Each
// wave takes 16 time rows of every SG_TC-sample tile; the next tile's
// voltages are loaded into registers while this one is computed and written
// to the other half of a double-buffered LDS stage: one barrier per tile.
template <int NAT>
__global__ __launch_bounds__(256, (NAT <= 4 ? 4 : 2)) void stokes_grad_kernel(AdjArgs a) {
    extern __shared__ __align__(16) unsigned char lds_raw[];
    constexpr int APAD = 16 * NAT;
    constexpr int XS = sg_xs(APAD);
    constexpr int PER = SG_TC * APAD / (64 * ADJ_WAVES);     // staged dwords per thread
    uint32_t* s_x = reinterpret_cast<uint32_t*>(lds_raw);     // [2][SG_TC][XS]
    float* s_w = reinterpret_cast<float*>(s_x + 2 * SG_TC * XS);   // [2][nsteps][64]
    __shared__ float s_loss[ADJ_WAVES];

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int i = lane & 15;
    const int g = lane >> 4;
    const int sp = blockIdx.x;
    const int f = blockIdx.y;
    const int bt = blockIdx.z;
    const int b0 = bt * 8;
    const int A = a.A, F = a.F, T = a.T, B = a.B;
    const int nsteps = a.nsteps;
    const int n_ag = nsteps / 2;                              // antenna groups of 4
    const int t_begin = sp * a.tsplit;
    const int t_end = min(T, t_begin + a.tsplit);
    const int ntiles = (t_end - t_begin + SG_TC - 1) / SG_TC;

    // Voltage staging: thread (wave, lane) takes antennas wave + 4k at time
    // lane of the tile (one coalesced 256-byte row per wave and k).
    uint32_t pre[PER];
    bool pre_t_ok = true;
    // Unconditional loads from clamped addresses (no branch per antenna row);
    // padded antennas and samples past the split are zeroed when the tile is
    // written to LDS, so the loads stay in flight until then.
    auto load_tile = [&](int t0) {
        const int tt = t0 + lane;
        pre_t_ok = tt < t_end;
        const int ttc = min(tt, t_end - 1);
#pragma unroll
        for (int k = 0; k < PER; k++)
            pre[k] = a.x[((size_t)min(wave + ADJ_WAVES * k, A - 1) * F + f) * T + ttc];
    };
    auto store_tile = [&](uint32_t* dst) {
#pragma unroll
        for (int k = 0; k < PER; k++) {
            const int an = wave + ADJ_WAVES * k;
            dst[lane * XS + an] = (pre_t_ok && an < A) ? pre[k] : 0u;
        }
    };
    load_tile(t_begin);

    // Forward B operand of beams b0..b0+7, both pols: [p][step][g][col],
    // step = 2 * antenna group + component of x, col = (beam, re/im).
    {
        const int n_w = 2 * nsteps * 64;
        for (int e0 = tid; e0 < n_w; e0 += ADJ_BATCH * ADJ_WAVES * 64) {
            float2 wv[ADJ_BATCH];
#pragma unroll
            for (int k = 0; k < ADJ_BATCH; k++) {
                const int e = e0 + k * ADJ_WAVES * 64;
                const int p = e / (nsteps * 64);
                const int r = e - p * nsteps * 64;
                const int st = r >> 6, kg = (r >> 4) & 3, col = r & 15;
                const int an = (st >> 1) * 4 + kg, b = b0 + (col >> 1);
                wv[k] = (e < n_w && b < B && an < A) ? a.w[(((size_t)b * A + an) * F + f) * 2 + p]
                                                      : make_float2(0.f, 0.f);
            }
#pragma unroll
            for (int k = 0; k < ADJ_BATCH; k++) {
                const int e = e0 + k * ADJ_WAVES * 64;
                if (e < n_w) {
                    const int r = e % (nsteps * 64);
                    const int c = (r >> 6) & 1, co = r & 1;
                    const float2 w = wv[k];
                    // y_r = x_r w_r - x_i w_i, y_i = x_r w_i + x_i w_r
                    s_w[e] = (co == 0) ? (c == 0 ? w.x : -w.y) : (c == 0 ? w.y : w.x);
                }
            }
        }
    }
    store_tile(s_x);
    if (ntiles > 1) load_tile(t_begin + SG_TC);
    __syncthreads();

    const int bcol = b0 + (i >> 1);
    const float lam = bcol < B ? a.lambda[bcol] : 0.f;
    const float swap_sign = (i & 1) ? -1.f : 1.f;
    floatx4 acc2[2][NAT];
#pragma unroll
    for (int p = 0; p < 2; p++)
#pragma unroll
        for (int at = 0; at < NAT; at++) acc2[p][at] = floatx4{0.f, 0.f, 0.f, 0.f};
    float loss = 0.f;

    for (int k = 0; k < ntiles; k++) {
        const uint32_t* sx = s_x + (k & 1) * SG_TC * XS;
        const int trow = wave * 16;
        // Forward: y^T rows = this wave's 16 times, both pols from one dword.
        floatx4 y[2] = {floatx4{0.f, 0.f, 0.f, 0.f}, floatx4{0.f, 0.f, 0.f, 0.f}};
#pragma unroll
        for (int ag = 0; ag < 4 * NAT; ag++) {
            if (ag < n_ag) {                       // uniform
                const uint32_t d = sx[(trow + i) * XS + ag * 4 + g];
#pragma unroll
                for (int p = 0; p < 2; p++) {
                    const float w0 = s_w[(p * nsteps + 2 * ag) * 64 + lane];
                    const float w1 = s_w[(p * nsteps + 2 * ag + 1) * 64 + lane];
                    y[p] = adj_mfma(ci8_component(d, 2 * p), w0, y[p]);
                    y[p] = adj_mfma(ci8_component(d, 2 * p + 1), w1, y[p]);
                }
            }
        }
        // Lane (g, i), register r: y at time trow + 4g + r, column i.  Its
        // gradient is gy = -lambda y; the imaginary-x block of G takes the
        // neighbouring column (re <-> im) with the sign of conj(x).
        float gyo[2][4], gys[2][4];
#pragma unroll
        for (int p = 0; p < 2; p++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const float v = y[p][r];
                loss += lam * v * v;
                gyo[p][r] = -lam * v;
                gys[p][r] = swap_sign * swap_pairs(gyo[p][r]);
            }
        // Adjoint: K = times trow + 4g + r (g = k index), both x components.
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const uint32_t* row = sx + (trow + 4 * g + r) * XS + i;
#pragma unroll
            for (int at = 0; at < NAT; at++) {
                const uint32_t d = row[at * 16];
#pragma unroll
                for (int p = 0; p < 2; p++) {
                    acc2[p][at] = adj_mfma(ci8_component(d, 2 * p), gyo[p][r], acc2[p][at]);
                    acc2[p][at] = adj_mfma(ci8_component(d, 2 * p + 1), gys[p][r], acc2[p][at]);
                }
            }
        }
        if (k + 1 < ntiles) {
            // Tile k + 1 into the other buffer (last read in iteration k - 1,
            // before the barrier that ended it), then tile k + 2's loads.
            store_tile(s_x + ((k + 1) & 1) * SG_TC * XS);
            if (k + 2 < ntiles) load_tile(t_begin + (k + 2) * SG_TC);
        }
        __syncthreads();
    }

    // Partials of the four waves into LDS (the voltage stage is free now),
    // [wave][p][col][antenna], then summed in wave order.
    float* s_part = reinterpret_cast<float*>(lds_raw);
#pragma unroll
    for (int p = 0; p < 2; p++)
#pragma unroll
        for (int at = 0; at < NAT; at++)
#pragma unroll
            for (int r = 0; r < 4; r++)
                s_part[((wave * 2 + p) * 16 + i) * APAD + at * 16 + 4 * g + r] = acc2[p][at][r];
    for (int off = 32; off > 0; off >>= 1) loss += __shfl_xor(loss, off, 64);
    if (lane == 0) s_loss[wave] = loss;
    __syncthreads();
    constexpr int NPART = 2 * 16 * APAD;
    float* out = a.partial + ((((size_t)sp * F + f) * 2) * a.btiles + bt) * 16 * APAD;
    const size_t pstride = (size_t)a.btiles * 16 * APAD;   // between pols
    for (int e = tid; e < NPART; e += blockDim.x) {
        const float v = ((s_part[e] + s_part[NPART + e]) + s_part[2 * NPART + e]) + s_part[3 * NPART + e];
        const int p = e / (16 * APAD), r = e - p * 16 * APAD;
        out[p * pstride + r] = v;
    }
    if (tid == 0)
        a.loss_part[((size_t)sp * F + f) * a.btiles + bt] = ((s_loss[0] + s_loss[1]) + s_loss[2]) + s_loss[3];
}

// ---------------------------------------------------------------------------
// Stokes-I loss and gradient for at most 4 beams (FP32), on
// v_mfma_f32_4x4x1_16b_f32: 16 independent 4x4 blocks with K = 1.  Lane
// 4 blk + c holds row c of block blk's A operand, column c of its B operand,
// and column c of its result (register r = row r).  With 4 beams the columns
// are exactly the beams, so no matrix work goes to absent beams (the 16x16
// tiles of stokes_grad_kernel are half empty there).  Complex products take
// three real products (P3 - P1 - P2 form).
//   Forward, blocks = groups of 4 times, K = one antenna per instruction:
//     lane (blk, c) feeds x[a][t0 + 4 blk + c] and w[a][beam c]; register r
//     of the result is y at time t0 + 4 blk + r, beam c.
//   Adjoint, blocks = groups of 4 antennas, K = one time per instruction:
//     lane (blk, c) feeds conj(x)[4 blk + c][t] and gy[beam c][t]; register r
//     is gw of antenna 4 blk + r, beam c.  gy passes through a per-wave LDS
//     tile [p][product][beam][time], read as broadcasts (all blocks read the
//     same four addresses).
// Each wave takes 64-time tiles; partial sums are added across waves in LDS
// in wave order and across time splits by adjoint_reduce_kernel, whose
// partial layout this kernel writes (rows 2 b and 2 b + 1, beam tile 0).
// Needs A <= 64, T % 4 == 0, A F T < 2^31 and 16-byte aligned voltages.
// ---------------------------------------------------------------------------
constexpr int SG4_TC = 64;            // times per wave tile
constexpr int SG4_WS = 8;             // LDS floats per (antenna, beam): (wr, wi, wr + wi) per pol, padded
// gy tile rows ([p][product][beam], SG4_TC times each) are padded by 4
// floats, so the four beams' rows that a read instruction broadcasts start
// on different LDS banks (unpadded, their 256-byte stride put all four on the
// same banks: 7 conflict cycles per LDS instruction).
constexpr int SG4_GS = SG4_TC + 4;
constexpr int SG4_GY = 2 * 3 * 4 * SG4_GS;   // floats of one wave's gy tile
constexpr int SG4_WGS = 3;            // resident workgroups per CU (launch bounds)

__device__ __forceinline__ floatx4 mfma4(float a, float b, floatx4 c) {
    return __builtin_amdgcn_mfma_f32_4x4x1f32(a, b, c, 0, 0, 0);
}

template <int NAT>
__global__ __launch_bounds__(256, SG4_WGS) void stokes_grad_b4_kernel(AdjArgs a) {
    extern __shared__ __align__(16) unsigned char lds_raw[];
    constexpr int AMAX = 16 * NAT;
    float* const s_w = reinterpret_cast<float*>(lds_raw);            // [A][4][SG4_WS]
    float* const s_gy = s_w + (size_t)a.A * 4 * SG4_WS;               // [wave][SG4_GY]
    __shared__ float s_loss[ADJ_WAVES];

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int blk = lane >> 2;
    const int c = lane & 3;
    const int sp = blockIdx.x;
    const int f = blockIdx.y;
    const int A = a.A, F = a.F, T = a.T, B = a.B;
    const int t_begin = sp * a.tsplit;
    const int t_end = min(T, t_begin + a.tsplit);
    const int ntiles = (t_end - t_begin + SG4_TC - 1) / SG4_TC;
    const unsigned ant_stride = (unsigned)F * (unsigned)T;   // A F T < 2^31 (host check)

    // Phasors of (antenna, beam): wr, wi, wr + wi of pol 0, then of pol 1.
    for (int e = tid; e < A * 4; e += blockDim.x) {
        const int an = e >> 2, b = e & 3;
        float v[6];
#pragma unroll
        for (int p = 0; p < 2; p++) {
            const float2 w = b < B ? a.w[(((size_t)b * A + an) * F + f) * 2 + p] : make_float2(0.f, 0.f);
            v[3 * p] = w.x;
            v[3 * p + 1] = w.y;
            v[3 * p + 2] = w.x + w.y;
        }
        float4* dst = reinterpret_cast<float4*>(s_w + (size_t)e * SG4_WS);
        dst[0] = make_float4(v[0], v[1], v[2], v[3]);
        dst[1] = make_float4(v[4], v[5], 0.f, 0.f);
    }
    __syncthreads();

    const float lam = c < B ? a.lambda[c] : 0.f;
    // Adjoint sums per pol: Q1 = sum gr xr, Q2 = sum gi xi, Q3 = sum (gr + gi)(xr - xi);
    // gw = (Q1 + Q2) + j (Q3 - Q1 + Q2).
    floatx4 q[2][3];
#pragma unroll
    for (int p = 0; p < 2; p++)
#pragma unroll
        for (int k = 0; k < 3; k++) q[p][k] = floatx4{0.f, 0.f, 0.f, 0.f};
    float loss = 0.f;
    float* const gyw = s_gy + wave * SG4_GY;
    // This lane's adjoint antenna row (clamped: rows past A are never stored).
    const uint4* const xa = reinterpret_cast<const uint4*>(a.x + min(4 * blk + c, A - 1) * ant_stride + f * T);

    // Forward voltages of tile kk at this lane's time, antennas 8 g .. 8 g + 7
    // (group g + 1 is in flight while group g is computed: 3% faster than
    // all antennas at once, with 136 instead of 168 VGPRs).  The stride and
    // count are opaque to the compiler, which would otherwise keep one
    // offset and one guard mask per antenna in scalar registers across tiles
    // (and spill them).
    uint32_t xd[2][8];
    auto load_fwd = [&](int kk, int g, uint32_t (&dst)[8]) {
        const int tl = t_begin + kk * SG4_TC + 4 * blk + c;
        const uint32_t* const xf = a.x + (f * T + min(tl, t_end - 1));
        unsigned st = ant_stride;
        int na = A;
        asm volatile("" : "+s"(st), "+s"(na));
#pragma unroll
        for (int j = 0; j < 8; j++)
            if (8 * g + j < na) dst[j] = xf[(8 * g + j) * st];     // uniform
    };
    // Adjoint voltages of 16 times from t: this lane's antenna row (times
    // past tn re-read the last 4 in bounds; their gy is not used).
    auto load_adj = [&](int t0, int tn, int t, uint4 (&xv)[4]) {
#pragma unroll
        for (int s4 = 0; s4 < 4; s4++) xv[s4] = xa[(t0 + min(t + 4 * s4, tn - 4)) >> 2];
    };

    for (int k = wave; k < ntiles; k += ADJ_WAVES) {
        const int t0 = t_begin + k * SG4_TC;
        const int tn = min(SG4_TC, t_end - t0);            // a multiple of 4
        uint4 xv[2][4];
        load_adj(t0, tn, 0, xv[0]);                        // in flight during the forward
        load_fwd(k, 0, xd[0]);
        // ---- forward: this lane feeds time t0 + 4 blk + c (clamped to the
        // split; results for times past it are zeroed below).  The antenna
        // count is opaque here too: 48 hoisted guard masks would not fit in
        // scalar registers.
        int na = A;
        asm volatile("" : "+s"(na));
        floatx4 y[2][3];
#pragma unroll
        for (int p = 0; p < 2; p++)
#pragma unroll
            for (int m = 0; m < 3; m++) y[p][m] = floatx4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int an = 0; an < AMAX; an++) {
            if (an % 8 == 0 && an + 8 < AMAX) load_fwd(k, an / 8 + 1, xd[(an / 8 + 1) & 1]);
            if (an < na) {                                        // uniform
                const uint32_t d = xd[(an / 8) & 1][an % 8];
                const float* wv = s_w + (an * 4 + c) * SG4_WS;
                const float4 w0 = *reinterpret_cast<const float4*>(wv);
                const float2 w1 = *reinterpret_cast<const float2*>(wv + 4);
                const float wp[2][3] = {{w0.x, w0.y, w0.z}, {w0.w, w1.x, w1.y}};
#pragma unroll
                for (int p = 0; p < 2; p++) {
                    const float xr = ci8_component(d, 2 * p), xi = ci8_component(d, 2 * p + 1);
                    y[p][0] = mfma4(xr, wp[p][0], y[p][0]);
                    y[p][1] = mfma4(xi, wp[p][1], y[p][1]);
                    y[p][2] = mfma4(xr + xi, wp[p][2], y[p][2]);
                }
            }
        }
        // ---- gy = -lambda y at (time t0 + 4 blk + r, beam c): gr, gi, gr + gi
        // into this wave's LDS tile [p][product][beam][time].  Times past the
        // split read clamped voltages: their y, and so gy and the loss, are
        // set to 0.
#pragma unroll
        for (int p = 0; p < 2; p++) {
            float g[3][4];
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const bool ok = 4 * blk + r < tn;
                const float yr = ok ? y[p][0][r] - y[p][1][r] : 0.f;
                const float yi = ok ? (y[p][2][r] - y[p][0][r]) - y[p][1][r] : 0.f;
                loss += lam * (yr * yr + yi * yi);
                const float gr = -lam * yr, gi = -lam * yi;
                g[0][r] = gr;
                g[1][r] = gi;
                g[2][r] = gr + gi;
            }
#pragma unroll
            for (int m = 0; m < 3; m++)
                *reinterpret_cast<float4*>(gyw + ((p * 3 + m) * 4 + c) * SG4_GS + 4 * blk) =
                    make_float4(g[m][0], g[m][1], g[m][2], g[m][3]);
        }
        // ---- adjoint: this lane feeds antenna 4 blk + c and gy of beam c;
        // one instruction per (time, pol, product).  The wave reads back its
        // own LDS writes (LDS operations of a wave complete in order).  The
        // next 16 times' voltages load while these are computed.
#pragma unroll
        for (int h = 0; h < SG4_TC / 16; h++) {
            if (h + 1 < SG4_TC / 16) load_adj(t0, tn, 16 * (h + 1), xv[(h + 1) & 1]);
            if (16 * h < tn) {                                    // uniform
#pragma unroll
                for (int s4 = 0; s4 < 4; s4++) {
                    const int tt = 16 * h + 4 * s4;
                    if (tt < tn) {                                // uniform
                        float4 gv[2][3];
#pragma unroll
                        for (int p = 0; p < 2; p++)
#pragma unroll
                            for (int m = 0; m < 3; m++)
                                gv[p][m] = *reinterpret_cast<const float4*>(gyw + ((p * 3 + m) * 4 + c) * SG4_GS + tt);
                        const uint4 xq = xv[h & 1][s4];
                        const uint32_t dw[4] = {xq.x, xq.y, xq.z, xq.w};
#pragma unroll
                        for (int s = 0; s < 4; s++) {
#pragma unroll
                            for (int p = 0; p < 2; p++) {
                                const float xr = ci8_component(dw[s], 2 * p), xi = ci8_component(dw[s], 2 * p + 1);
                                q[p][0] = mfma4(xr, gv[p][0][s], q[p][0]);
                                q[p][1] = mfma4(xi, gv[p][1][s], q[p][1]);
                                q[p][2] = mfma4(xr - xi, gv[p][2][s], q[p][2]);
                            }
                        }
                    }
                }
            }
        }
    }

    // This wave's sums into its own gy tile (its last reads of it are done):
    // [p][row = 2 beam + re/im][antenna 64], then summed over waves in order.
    float* const pw = gyw;
#pragma unroll
    for (int p = 0; p < 2; p++)
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int an = 4 * blk + r;
            pw[(p * 8 + 2 * c) * 64 + an] = q[p][0][r] + q[p][1][r];
            pw[(p * 8 + 2 * c + 1) * 64 + an] = (q[p][2][r] - q[p][0][r]) + q[p][1][r];
        }
    for (int off = 32; off > 0; off >>= 1) loss += __shfl_xor(loss, off, 64);
    if (lane == 0) s_loss[wave] = loss;
    __syncthreads();
    const int apad = a.apad;
    float* out = a.partial + ((size_t)sp * F + f) * 2 * a.btiles * 16 * apad;
    const size_t pstride = (size_t)a.btiles * 16 * apad;   // between pols
    for (int e = tid; e < 2 * 8 * A; e += blockDim.x) {
        const int an = e % A, pr = e / A;                   // pr = p * 8 + row
        const float* src = s_gy + pr * 64 + an;
        const float v = ((src[0] + src[SG4_GY]) + src[2 * SG4_GY]) + src[3 * SG4_GY];
        out[(pr >> 3) * pstride + (pr & 7) * apad + an] = v;
    }
    if (tid == 0) a.loss_part[((size_t)sp * F + f) * a.btiles] = ((s_loss[0] + s_loss[1]) + s_loss[2]) + s_loss[3];
}

// Power-of-two scale s with m * s in [0.5, 1) for m > 0 (1 for m = 0), and 1/s.
__device__ __forceinline__ void pow2_scale(float m, float& s, float& inv) {
    int ex = 0;
    frexpf(m, &ex);
    s = m > 0.f ? ldexpf(1.0f, -ex) : 1.f;
    inv = m > 0.f ? ldexpf(1.0f, ex) : 1.f;
}

// v (|v| < 1) as hi + lo FP16 halves, into element j of two vectors.
__device__ __forceinline__ void split_f16(float v, halfx8& hi, halfx8& lo, int j) {
    const _Float16 h = (_Float16)v;
    hi[j] = h;
    lo[j] = (_Float16)(v - (float)h);
}

// Split-FP16 form of stokes_grad_kernel: the same transposed contractions and
// tiles on v_mfma_f32_16x16x32_f16 (K = 32 per instruction), FP32
// accumulation.  CI8 voltages are exact in FP16.
//   * Forward B operand: the phasor embedding times a power-of-two scale per
//     pol (every component below 1), as hi + lo FP16 halves; K = 16
//     antennas x (re, im) of x per instruction.
//   * Adjoint B operand: the wave's gy of the tile, K = 16 times x (re, im)
//     -- exactly the forward outputs the lane holds (times 4g..4g+3 and the
//     neighbouring column) -- scaled by a power of two from the wave's
//     largest |gy| and split the same way; each tile's product is unscaled
//     into the FP32 sums.
// Error against FP64 is that of FP32 arithmetic (hi + lo carry ~22 bits).
template <int NAT>
__global__ __launch_bounds__(256, (NAT <= 3 ? 3 : 2)) void stokes_grad_f16_kernel(AdjArgs a) {
    extern __shared__ __align__(16) unsigned char lds_raw[];
    constexpr int APAD = 16 * NAT;
    constexpr int XS = sg_xs(APAD);
    constexpr int PER = SG_TC * APAD / (64 * ADJ_WAVES);
    uint32_t* s_x = reinterpret_cast<uint32_t*>(lds_raw);                 // [2][SG_TC][XS]
    halfx8* s_wh = reinterpret_cast<halfx8*>(s_x + 2 * SG_TC * XS);       // [p][blk][hl][64]
    __shared__ float s_loss[ADJ_WAVES];
    __shared__ unsigned int s_wmax[2];

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int i = lane & 15;
    const int g = lane >> 4;
    const int sp = blockIdx.x;
    const int f = blockIdx.y;
    const int bt = blockIdx.z;
    const int b0 = bt * 8;
    const int A = a.A, F = a.F, T = a.T, B = a.B;
    const int t_begin = sp * a.tsplit;
    const int t_end = min(T, t_begin + a.tsplit);
    const int ntiles = (t_end - t_begin + SG_TC - 1) / SG_TC;

    uint32_t pre[PER];
    bool pre_t_ok = true;
    // Unconditional loads from clamped addresses (no branch per antenna row);
    // padded antennas and samples past the split are zeroed when the tile is
    // written to LDS, so the loads stay in flight until then.
    auto load_tile = [&](int t0) {
        const int tt = t0 + lane;
        pre_t_ok = tt < t_end;
        const int ttc = min(tt, t_end - 1);
#pragma unroll
        for (int k = 0; k < PER; k++)
            pre[k] = a.x[((size_t)min(wave + ADJ_WAVES * k, A - 1) * F + f) * T + ttc];
    };
    auto store_tile = [&](uint32_t* dst) {
#pragma unroll
        for (int k = 0; k < PER; k++) {
            const int an = wave + ADJ_WAVES * k;
            dst[lane * XS + an] = (pre_t_ok && an < A) ? pre[k] : 0u;
        }
    };
    load_tile(t_begin);

    // Phasor scale per pol: the largest component magnitude of this
    // workgroup's beams (an exact maximum, so the order of the atomics does
    // not matter).
    if (tid < 2) s_wmax[tid] = 0u;
    __syncthreads();
    {
        float mx[2] = {0.f, 0.f};
        for (int e = tid; e < 2 * 8 * A; e += blockDim.x) {
            const int p = e / (8 * A), r = e - p * 8 * A;
            const int bb = r / A, an = r - bb * A, b = b0 + bb;
            if (b < B) {
                const float2 w = a.w[(((size_t)b * A + an) * F + f) * 2 + p];
                mx[p] = fmaxf(mx[p], fmaxf(fabsf(w.x), fabsf(w.y)));
            }
        }
        atomicMax(&s_wmax[0], __float_as_uint(mx[0]));
        atomicMax(&s_wmax[1], __float_as_uint(mx[1]));
    }
    __syncthreads();
    float wsc[2], wuns[2];
    pow2_scale(__uint_as_float(s_wmax[0]), wsc[0], wuns[0]);
    pow2_scale(__uint_as_float(s_wmax[1]), wsc[1], wuns[1]);
    // Forward B operand: lane (g, i) of block blk holds k = 8g + 2q + c ->
    // (antenna 16 blk + 4g + q, component c of x), column i -> (beam
    // b0 + i/2, re/im).  y_r = x_r w_r - x_i w_i, y_i = x_r w_i + x_i w_r.
    for (int e = tid; e < 2 * NAT * 64; e += blockDim.x) {
        const int ln = e & 63, pb = e >> 6;
        const int p = pb / NAT, blk = pb - p * NAT;
        const int li = ln & 15, lg = ln >> 4;
        const int b = b0 + (li >> 1), co = li & 1;
        halfx8 hi, lo;
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int an = blk * 16 + 4 * lg + q;
            float2 w = make_float2(0.f, 0.f);
            if (b < B && an < A) w = a.w[(((size_t)b * A + an) * F + f) * 2 + p];
            const float v0 = (co == 0 ? w.x : w.y) * wsc[p];     // multiplies Re x
            const float v1 = (co == 0 ? -w.y : w.x) * wsc[p];    // multiplies Im x
            split_f16(v0, hi, lo, 2 * q);
            split_f16(v1, hi, lo, 2 * q + 1);
        }
        s_wh[(pb * 2 + 0) * 64 + ln] = hi;
        s_wh[(pb * 2 + 1) * 64 + ln] = lo;
    }
    store_tile(s_x);
    if (ntiles > 1) load_tile(t_begin + SG_TC);
    __syncthreads();

    const int bcol = b0 + (i >> 1);
    const float lam = bcol < B ? a.lambda[bcol] : 0.f;
    const float swap_sign = (i & 1) ? -1.f : 1.f;
    floatx4 acc2[2][NAT];
#pragma unroll
    for (int p = 0; p < 2; p++)
#pragma unroll
        for (int at = 0; at < NAT; at++) acc2[p][at] = floatx4{0.f, 0.f, 0.f, 0.f};
    float loss = 0.f;
    const floatx4 zero4 = {0.f, 0.f, 0.f, 0.f};

    for (int k = 0; k < ntiles; k++) {
        const uint32_t* sx = s_x + (k & 1) * SG_TC * XS;
        const int trow = wave * 16;
        // Forward: rows = this wave's 16 times; one 16-byte LDS read gives
        // 4 antennas x both pols.  The phasor operands are re-read from LDS
        // every tile (an opaque offset keeps the compiler from holding all
        // 4 NAT of them in registers across the loop).
        int woff = lane;
        asm volatile("" : "+v"(woff));
        const halfx8* wh = s_wh + woff;
        floatx4 y[2] = {zero4, zero4};
#pragma unroll
        for (int blk = 0; blk < NAT; blk++) {
            const uint4 d = *reinterpret_cast<const uint4*>(sx + (trow + i) * XS + blk * 16 + 4 * g);
            halfx8 xa[2];
            const uint32_t dw[4] = {d.x, d.y, d.z, d.w};
#pragma unroll
            for (int q = 0; q < 4; q++) {
                halfx2 h0, h1;
                ci8_to_half2x2(dw[q], h0, h1);
                xa[0][2 * q] = h0[0]; xa[0][2 * q + 1] = h0[1];
                xa[1][2 * q] = h1[0]; xa[1][2 * q + 1] = h1[1];
            }
#pragma unroll
            for (int p = 0; p < 2; p++) {
                y[p] = mfma_f16(xa[p], wh[((p * NAT + blk) * 2 + 0) * 64], y[p]);
                y[p] = mfma_f16(xa[p], wh[((p * NAT + blk) * 2 + 1) * 64], y[p]);
            }
        }
        // gy = -lambda y (unscaled), its wave maximum, and the B operands.
        halfx8 gh[2], gl[2];
        float gsc_inv[2];
#pragma unroll
        for (int p = 0; p < 2; p++) {
            float gyo[4], gys[4], m = 0.f;
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const float v = y[p][r] * wuns[p];
                loss += lam * v * v;
                gyo[r] = -lam * v;
                gys[r] = swap_sign * swap_pairs(gyo[r]);
                m = fmaxf(m, fabsf(gyo[r]));
            }
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off, 64));
            float sc;
            pow2_scale(m, sc, gsc_inv[p]);
#pragma unroll
            for (int r = 0; r < 4; r++) {
                split_f16(gyo[r] * sc, gh[p], gl[p], 2 * r);
                split_f16(gys[r] * sc, gh[p], gl[p], 2 * r + 1);
            }
        }
        // Adjoint: rows = antennas 16 at + i, K = times trow + 4g + q (x re, im).
#pragma unroll
        for (int at = 0; at < NAT; at++) {
            halfx8 xa[2];
#pragma unroll
            for (int q = 0; q < 4; q++) {
                halfx2 h0, h1;
                ci8_to_half2x2(sx[(trow + 4 * g + q) * XS + at * 16 + i], h0, h1);
                xa[0][2 * q] = h0[0]; xa[0][2 * q + 1] = h0[1];
                xa[1][2 * q] = h1[0]; xa[1][2 * q + 1] = h1[1];
            }
#pragma unroll
            for (int p = 0; p < 2; p++) {
                floatx4 t = mfma_f16(xa[p], gh[p], zero4);
                t = mfma_f16(xa[p], gl[p], t);
#pragma unroll
                for (int r = 0; r < 4; r++) acc2[p][at][r] = fmaf(t[r], gsc_inv[p], acc2[p][at][r]);
            }
        }
        if (k + 1 < ntiles) {
            store_tile(s_x + ((k + 1) & 1) * SG_TC * XS);
            if (k + 2 < ntiles) load_tile(t_begin + (k + 2) * SG_TC);
        }
        __syncthreads();
    }

    float* s_part = reinterpret_cast<float*>(lds_raw);
#pragma unroll
    for (int p = 0; p < 2; p++)
#pragma unroll
        for (int at = 0; at < NAT; at++)
#pragma unroll
            for (int r = 0; r < 4; r++)
                s_part[((wave * 2 + p) * 16 + i) * APAD + at * 16 + 4 * g + r] = acc2[p][at][r];
    for (int off = 32; off > 0; off >>= 1) loss += __shfl_xor(loss, off, 64);
    if (lane == 0) s_loss[wave] = loss;
    __syncthreads();
    constexpr int NPART = 2 * 16 * APAD;
    float* out = a.partial + ((((size_t)sp * F + f) * 2) * a.btiles + bt) * 16 * APAD;
    const size_t pstride = (size_t)a.btiles * 16 * APAD;
    for (int e = tid; e < NPART; e += blockDim.x) {
        const float v = ((s_part[e] + s_part[NPART + e]) + s_part[2 * NPART + e]) + s_part[3 * NPART + e];
        const int p = e / (16 * APAD), r = e - p * 16 * APAD;
        out[p * pstride + r] = v;
    }
    if (tid == 0)
        a.loss_part[((size_t)sp * F + f) * a.btiles + bt] = ((s_loss[0] + s_loss[1]) + s_loss[2]) + s_loss[3];
}

// gw[b,a,f,p] = sum over time splits of the partial rows (b, re) and (b, im).
// Threads follow the partial layout (antenna fastest), so the nsplit x 2
// reads per output coalesce; the writes, one float2 per thread into
// [B,A,F,1,P], are the strided side (8 B per thread, one pass).
__global__ __launch_bounds__(256) void adjoint_reduce_kernel(AdjArgs a, int nsplit, float2* gw) {
    const size_t n = (size_t)a.F * 2 * a.btiles * 8 * a.A;
    const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= n) return;
    size_t r = idx;
    const int an = (int)(r % a.A);
    r /= a.A;
    const int bb = (int)(r % 8);
    r /= 8;
    const int bt = (int)(r % a.btiles);
    r /= a.btiles;
    const int p = (int)(r & 1);
    const int f = (int)(r >> 1);
    const int b = bt * 8 + bb;
    if (b >= a.B) return;
    float re = 0.f, im = 0.f;
    for (int sp = 0; sp < nsplit; sp++) {
        const float* blk = a.partial + ((((size_t)sp * a.F + f) * 2 + p) * a.btiles + bt) * 16 * a.apad;
        re += blk[(2 * bb) * a.apad + an];
        im += blk[(2 * bb + 1) * a.apad + an];
    }
    gw[(((size_t)b * a.A + an) * a.F + f) * 2 + p] = make_float2(re, im);
}

// loss = -sum of the workgroup partials, accumulated in FP64 in a fixed order.
__global__ __launch_bounds__(256) void adjoint_loss_kernel(const float* part, size_t n, double* loss) {
    __shared__ double s[256];
    double v = 0.0;
    for (size_t e = threadIdx.x; e < n; e += blockDim.x) v += (double)part[e];
    s[threadIdx.x] = v;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) s[threadIdx.x] += s[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) *loss = -s[0];
}

static int run_adjoint(bool fused, bool f16, const void* input, const void* phasors, const void* grad_out,
                       const float* beam_weights, double* loss, void* grad_phasors, uint64_t A,
                       uint64_t F, uint64_t T, uint64_t P, uint64_t B, hipStream_t s) {
    const char* who = fused ? "stokes-i gradient" : "beamformer adjoint";
    if (P != 2) return fail(std::string(who) + ": number of polarizations must be 2");
    if (A == 0 || B == 0) return fail(std::string(who) + ": need at least one antenna and one beam");
    if (A > 16 * ADJ_MAXAT) return fail(std::string(who) + ": at most 128 antennas");
    if (fused && B > ADJ_MAXB) return fail(std::string(who) + ": at most 64 beams");
    if (B > 4096 || F > 65535 || T > (1u << 30)) return fail(std::string(who) + ": shape too large");
    if (!input || !grad_phasors || (fused ? (!phasors || !loss || !beam_weights) : !grad_out))
        return fail(std::string(who) + ": null pointer");
    if ((uintptr_t)input % 4 || (uintptr_t)grad_phasors % 8 || (phasors && (uintptr_t)phasors % 16) ||
        (grad_out && (uintptr_t)grad_out % 8))
        return fail(std::string(who) + ": misaligned buffer");
    AdjArgs a;
    std::memset(&a, 0, sizeof(a));
    a.x = static_cast<const uint32_t*>(input);
    a.w = static_cast<const float2*>(phasors);
    a.gy = static_cast<const float2*>(grad_out);
    a.A = (int)A; a.F = (int)F; a.T = (int)T; a.B = (int)B;
    a.n_at = (int)ceil_div(A, 16);
    a.apad = 16 * a.n_at;
    a.nsteps = 2 * (int)ceil_div(A, 4);
    a.btiles = (int)ceil_div(B, 8);
    if (fused)
        for (uint64_t b = 0; b < B; b++) a.lambda[b] = beam_weights[b];
    // FP32 loss + gradient with at most 4 beams: the 4x4-block kernel.
    const bool b4 = fused && !f16 && B <= 4 && A <= 64 && T % 4 == 0 && (uintptr_t)input % 16 == 0 &&
                    A * F * T < (1ull << 31);
    // Time splits, each a whole number of tiles.  Plain adjoint: about 2048
    // workgroups.  Fused: the split length that minimises rounds of resident
    // workgroups (per CU: 4 FP32 or 3 split FP16 up to 64 or 48 antennas,
    // else 2; the 4x4-block kernel SG4_WGS) x tiles per wave (longest on
    // ties, which keeps the partial sums few).
    const int tc = b4 ? SG4_TC : fused ? SG_TC : ADJ_TC;
    const long tiles = (long)ceil_div(std::max<uint64_t>(T, 1), (uint64_t)tc);
    const long per_split = (long)F * a.btiles;
    long tiles_per;
    if (fused) {
        const long slots = (long)device_cus() * (b4 ? SG4_WGS : f16 ? (a.n_at <= 3 ? 3 : 2) : (a.n_at <= 4 ? 4 : 2));
        long best = -1;
        tiles_per = tiles;
        for (long tp = tiles; tp >= 1; tp--) {
            const long ns = (tiles + tp - 1) / tp;
            const long per_wave = b4 ? (tp + ADJ_WAVES - 1) / ADJ_WAVES : tp;   // b4: waves take tiles in turn
            const long cost = ((ns * per_split + slots - 1) / slots) * per_wave;
            if (best < 0 || cost < best) { best = cost; tiles_per = tp; }
        }
    } else {
        const long ns = std::max(1L, std::min(tiles, (2048 + per_split - 1) / per_split));
        tiles_per = (tiles + ns - 1) / ns;
    }
    const long nsplit = (tiles + tiles_per - 1) / tiles_per;
    a.tsplit = (int)(tiles_per * tc);

    const size_t part_floats = (size_t)nsplit * F * 2 * a.btiles * 16 * a.apad;
    const size_t loss_floats = (size_t)nsplit * F * a.btiles;
    unsigned char* ws = nullptr;
    const size_t ws_bytes = (part_floats + loss_floats) * sizeof(float);
    if (F == 0) return BLADE_SUCCESS;
    int rc = library_workspace(WS_ADJOINT, ws_bytes, s, &ws);
    if (rc) return rc;
    a.partial = reinterpret_cast<float*>(ws);
    a.loss_part = a.partial + part_floats;

    const size_t lds = !fused ? ((size_t)ADJ_TC * adj_xs(a.apad) + ADJ_TC * ADJ_GYS + 2 * 16 * a.apad) * 4
                       : b4    ? ((size_t)a.A * 4 * SG4_WS + (size_t)ADJ_WAVES * SG4_GY) * 4
                       : f16   ? (size_t)2 * SG_TC * sg_xs(a.apad) * 4 + (size_t)2 * a.n_at * 2 * 64 * 16
                               : ((size_t)2 * SG_TC * sg_xs(a.apad) + 2 * (size_t)a.nsteps * 64) * 4;
    const void* fn = nullptr;
    if (b4) {
        switch (a.n_at) {
            case 1: fn = reinterpret_cast<const void*>(&stokes_grad_b4_kernel<1>); break;
            case 2: fn = reinterpret_cast<const void*>(&stokes_grad_b4_kernel<2>); break;
            case 3: fn = reinterpret_cast<const void*>(&stokes_grad_b4_kernel<3>); break;
            default: fn = reinterpret_cast<const void*>(&stokes_grad_b4_kernel<4>); break;
        }
    } else if (fused && f16) {
        switch (a.n_at) {
            case 1: fn = reinterpret_cast<const void*>(&stokes_grad_f16_kernel<1>); break;
            case 2: fn = reinterpret_cast<const void*>(&stokes_grad_f16_kernel<2>); break;
            case 3: fn = reinterpret_cast<const void*>(&stokes_grad_f16_kernel<3>); break;
            case 4: fn = reinterpret_cast<const void*>(&stokes_grad_f16_kernel<4>); break;
            case 5: fn = reinterpret_cast<const void*>(&stokes_grad_f16_kernel<5>); break;
            case 6: fn = reinterpret_cast<const void*>(&stokes_grad_f16_kernel<6>); break;
            case 7: fn = reinterpret_cast<const void*>(&stokes_grad_f16_kernel<7>); break;
            default: fn = reinterpret_cast<const void*>(&stokes_grad_f16_kernel<8>); break;
        }
    } else if (fused) {
        switch (a.n_at) {
            case 1: fn = reinterpret_cast<const void*>(&stokes_grad_kernel<1>); break;
            case 2: fn = reinterpret_cast<const void*>(&stokes_grad_kernel<2>); break;
            case 3: fn = reinterpret_cast<const void*>(&stokes_grad_kernel<3>); break;
            case 4: fn = reinterpret_cast<const void*>(&stokes_grad_kernel<4>); break;
            case 5: fn = reinterpret_cast<const void*>(&stokes_grad_kernel<5>); break;
            case 6: fn = reinterpret_cast<const void*>(&stokes_grad_kernel<6>); break;
            case 7: fn = reinterpret_cast<const void*>(&stokes_grad_kernel<7>); break;
            default: fn = reinterpret_cast<const void*>(&stokes_grad_kernel<8>); break;
        }
    } else {
        fn = reinterpret_cast<const void*>(&adjoint_kernel);
    }
    if (lds > 65536) {
        hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) {
            set_error(std::string(who) + ": LDS attribute: " + hipGetErrorString(e));
            return BLADE_HIP_ERROR;
        }
    }
    dim3 grid((unsigned)nsplit, (unsigned)F, (unsigned)a.btiles);
    const dim3 blk(64 * ADJ_WAVES);
    if (b4) {
        switch (a.n_at) {
            case 1: hipLaunchKernelGGL(stokes_grad_b4_kernel<1>, grid, blk, lds, s, a); break;
            case 2: hipLaunchKernelGGL(stokes_grad_b4_kernel<2>, grid, blk, lds, s, a); break;
            case 3: hipLaunchKernelGGL(stokes_grad_b4_kernel<3>, grid, blk, lds, s, a); break;
            default: hipLaunchKernelGGL(stokes_grad_b4_kernel<4>, grid, blk, lds, s, a); break;
        }
    } else if (fused && f16) {
        switch (a.n_at) {
            case 1: hipLaunchKernelGGL(stokes_grad_f16_kernel<1>, grid, blk, lds, s, a); break;
            case 2: hipLaunchKernelGGL(stokes_grad_f16_kernel<2>, grid, blk, lds, s, a); break;
            case 3: hipLaunchKernelGGL(stokes_grad_f16_kernel<3>, grid, blk, lds, s, a); break;
            case 4: hipLaunchKernelGGL(stokes_grad_f16_kernel<4>, grid, blk, lds, s, a); break;
            case 5: hipLaunchKernelGGL(stokes_grad_f16_kernel<5>, grid, blk, lds, s, a); break;
            case 6: hipLaunchKernelGGL(stokes_grad_f16_kernel<6>, grid, blk, lds, s, a); break;
            case 7: hipLaunchKernelGGL(stokes_grad_f16_kernel<7>, grid, blk, lds, s, a); break;
            default: hipLaunchKernelGGL(stokes_grad_f16_kernel<8>, grid, blk, lds, s, a); break;
        }
    } else if (fused) {
        switch (a.n_at) {
            case 1: hipLaunchKernelGGL(stokes_grad_kernel<1>, grid, dim3(64 * ADJ_WAVES), lds, s, a); break;
            case 2: hipLaunchKernelGGL(stokes_grad_kernel<2>, grid, dim3(64 * ADJ_WAVES), lds, s, a); break;
            case 3: hipLaunchKernelGGL(stokes_grad_kernel<3>, grid, dim3(64 * ADJ_WAVES), lds, s, a); break;
            case 4: hipLaunchKernelGGL(stokes_grad_kernel<4>, grid, dim3(64 * ADJ_WAVES), lds, s, a); break;
            case 5: hipLaunchKernelGGL(stokes_grad_kernel<5>, grid, dim3(64 * ADJ_WAVES), lds, s, a); break;
            case 6: hipLaunchKernelGGL(stokes_grad_kernel<6>, grid, dim3(64 * ADJ_WAVES), lds, s, a); break;
            case 7: hipLaunchKernelGGL(stokes_grad_kernel<7>, grid, dim3(64 * ADJ_WAVES), lds, s, a); break;
            default: hipLaunchKernelGGL(stokes_grad_kernel<8>, grid, dim3(64 * ADJ_WAVES), lds, s, a); break;
        }
    } else {
        hipLaunchKernelGGL(adjoint_kernel, grid, dim3(64 * ADJ_WAVES), lds, s, a);
    }
    const size_t n = (size_t)F * 2 * a.btiles * 8 * A;
    hipLaunchKernelGGL(adjoint_reduce_kernel, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0, s, a,
                       (int)nsplit, static_cast<float2*>(grad_phasors));
    if (fused)
        hipLaunchKernelGGL(adjoint_loss_kernel, dim3(1), dim3(256), 0, s, a.loss_part, loss_floats, loss);
    return check_launch(who);
}

}  // namespace bl

using namespace bl;

extern "C" int blade_beamformer_adjoint(const void* input, const void* grad_output,
                                        void* grad_phasors, uint64_t A, uint64_t F, uint64_t T,
                                        uint64_t P, uint64_t B, blade_stream_t stream) {
    return run_adjoint(false, false, input, nullptr, grad_output, nullptr, nullptr, grad_phasors, A, F, T,
                       P, B, reinterpret_cast<hipStream_t>(stream));
}

extern "C" int blade_beamformer_stokes_i_grad(int precision, const void* input, const void* phasors,
                                              const float* beam_weights, double* loss,
                                              void* grad_phasors, uint64_t A, uint64_t F,
                                              uint64_t T, uint64_t P, uint64_t B,
                                              blade_stream_t stream) {
    if (precision != BLADE_PREC_F32 && precision != BLADE_PREC_F16X2)
        return fail("stokes-i gradient: unknown precision");
    return run_adjoint(true, precision == BLADE_PREC_F16X2, input, phasors, nullptr, beam_weights, loss, grad_phasors, A, F, T, P,
                       B, reinterpret_cast<hipStream_t>(stream));
}
