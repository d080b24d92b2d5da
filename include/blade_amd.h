/*
 * blade_amd.h — C-ABI of the MI355X (gfx950) ATA beamforming path.
 *
 * Plain pointers, sizes and a hipStream_t; no C++ or torch types.  Every
 * buffer argument is DEVICE memory owned by the caller, and every call is
 * stream-ordered and asynchronous.  Entry points return a blade_result
 * (the reference's Result codes, include/blade/macros.hh:31-48);
 * blade_amd_last_error() gives the calling thread's last message.
 *
 * Layouts are the reference's row-major tensors
 * (include/blade/memory/custom.hh:10-163):
 *   voltages / beams  [A|B, F, T, P]       (ArrayShape)
 *   phasors           [B, A, F, 1, P]      (PhasorShape)
 *   delays            [B, A]               (DelayShape)
 * Complex types are interleaved {re, im}; CI8 is two int8.
 */
#ifndef BLADE_AMD_H
#define BLADE_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ihipStream_t* blade_stream_t;   /* == hipStream_t */

/* Result codes: values of the reference's Blade::Result (macros.hh:31-48). */
enum blade_result {
    BLADE_SUCCESS = 0,
    BLADE_ERROR = 1,
    BLADE_HIP_ERROR = 2,          /* the reference's CUDA_ERROR */
    BLADE_ASSERTION_ERROR = 3,
};

/* Element types (the reference's TypeInfo names, include/blade/memory/types.hh). */
enum blade_type {
    BLADE_I8 = 0, BLADE_F16 = 1, BLADE_F32 = 2, BLADE_F64 = 3, BLADE_I32 = 4,
    BLADE_CI8 = 5, BLADE_CF16 = 6, BLADE_CF32 = 7, BLADE_CF64 = 8,
};

/* Beamformer arithmetic. */
enum blade_precision {
    BLADE_PREC_F32 = 0,     /* FP32 MFMA (v_mfma_f32_*_f32), FP32 accumulate */
    BLADE_PREC_F16X2 = 1,   /* split FP16: phasor = hi + lo, 2 FP16 MFMAs, FP32 accumulate (CI8 input only) */
};

/* Library identification; never fails. */
const char* blade_amd_version(void);
/* Thread-local message of the last failing call ("" if none). */
const char* blade_amd_last_error(void);

/*
 * Cast — component-wise static_cast over n_elements tensor elements.
 * Replaces Modules::Cast<IT,OT>::process (src/modules/cast/base.cc:60-66,
 * kernel src/modules/cast/cast.cu:5-12).  Complex<->real pairs are rejected
 * like cast/base.cc:39-43.  in == out is allowed only when the types match.
 */
int blade_cast(int in_type, int out_type, const void* in, void* out,
               uint64_t n_elements, blade_stream_t stream);

/*
 * ATA beamformer.
 * Replaces Modules::Beamformer::ATA<CF32,CF32> (src/modules/beamformer/ata.cc:7-72,
 * generic.cc:31-33; kernel src/modules/beamformer/beamformer.cu:13-79):
 *   y[b,f,t,p] = sum_a x[a,f,t,p] * w[b,a,f,0,p]
 *   y[B,f,t,p] = (sum_a |x w[0]|^2, 0) when incoherent (sqrt'ed when incoherent_sqrt)
 * input:  in_type CF32 or CI8, [A,F,T,P];  phasors: CF32 [B,A,F,1,P];
 * output: CF32 [B + incoherent, F, T, P].  P must be 2 (as the reference).
 * precision BLADE_PREC_F16X2 requires in_type == BLADE_CI8.
 */
int blade_beamformer_ata(int in_type, int precision, const void* input, const void* phasors,
                         void* output, uint64_t A, uint64_t F, uint64_t T, uint64_t P,
                         uint64_t B, int incoherent, int incoherent_sqrt,
                         blade_stream_t stream);

/*
 * Fused ATA Mode-B chain on one block: beamform -> channelize (rate C in {1,2,4})
 * -> optional polarizer (XY -> LR), one HBM pass.
 * Equivalent to Beamformer::ATA (beamformer.cu:13-79) followed by
 * Channelizer (oracle tests/modules/channelizer/advanced.py:49-65) and
 * Polarizer (polarizer.cu:5-25); output CF32 [B + incoherent, F*C, T/C, P].
 */
int blade_beamform_chain(int in_type, int precision, const void* input, const void* phasors,
                         void* output, uint64_t A, uint64_t F, uint64_t T, uint64_t P,
                         uint64_t B, int incoherent, int incoherent_sqrt, uint64_t rate,
                         int polarize, blade_stream_t stream);

/*
 * The two calls above pack the phasors into the matrix-core operand layout
 * inside the kernel, in LDS, on every call; their only device scratch (a
 * 1 KiB sink for the stores of idle lanes) is a library-owned buffer per
 * (device, stream).  Callers that own their memory (modules, hipGraph
 * capture) size a device workspace with blade_beamformer_workspace_size
 * and pass it to blade_beamform_chain_ws (same semantics as
 * blade_beamform_chain).  The reference's beamformer needs no scratch
 * (src/modules/beamformer/generic.cc:31-33); this pair has no counterpart.
 */
uint64_t blade_beamformer_workspace_size(int precision, uint64_t A, uint64_t F, uint64_t B);
int blade_beamform_chain_ws(int in_type, int precision, const void* input, const void* phasors,
                            void* output, uint64_t A, uint64_t F, uint64_t T, uint64_t P,
                            uint64_t B, int incoherent, int incoherent_sqrt, uint64_t rate,
                            int polarize, void* workspace, uint64_t workspace_bytes,
                            blade_stream_t stream);

/*
 * blade_beamform_chain without any allocation: no reference counterpart.
 * For callers that cannot allocate (code under hipGraph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  Same arguments and
 * results as blade_beamform_chain; its only device scratch (the 1 KiB store
 * sink of idle lanes, blade_beamformer_workspace_size) is a static device
 * array of the library, one per device, held until the process ends.  Calls
 * needing more scratch than that are rejected.  The array is write-only, so
 * calls that overlap on one device (different streams) share it without
 * changing each other's results.
 */
int blade_beamform_chain_noalloc(int in_type, int precision, const void* input, const void* phasors,
                                 void* output, uint64_t A, uint64_t F, uint64_t T, uint64_t P,
                                 uint64_t B, int incoherent, int incoherent_sqrt, uint64_t rate,
                                 int polarize, blade_stream_t stream);

/*
 * Many beamformer blocks of different shapes in one persistent launch (no
 * reference counterpart; each block is what one Beamformer::ATA module call
 * computes, src/modules/beamformer/beamformer.cu:13-79, without the
 * incoherent beam).  Each block: CI8 input [A,F,T,2], CF32 phasors
 * [B,A,F,1,2], CF32 output [B,F,T,2]; T a multiple of 64.  The call is
 * asynchronous on `stream`, allocates nothing, and can be captured in a
 * hipGraph: size the caller-owned workspace with
 * blade_beamform_batched_workspace_size (same blocks, same tuning).
 */
typedef struct blade_beamform_block {
    const void* input;
    const void* phasors;
    void* output;
    uint64_t A, F, T, B;
} blade_beamform_block;

/*
 * Tile shape and launch shape of the batched launch; 0 = automatic.
 *   time_samples_per_lane: 4
 *   beam_tiles: 16-row beam tiles per wavefront, 1 (8 beams) or 2 (16;
 *               FP32 then forms three real products per complex one)
 *   chunks_per_item: time chunks of 256 samples per work item
 *   workgroups_per_cu: workgroups per compute unit in the grid (grid =
 *               min(work items, this x CUs); items beyond it are walked)
 */
typedef struct blade_batch_tuning {
    int32_t time_samples_per_lane;
    int32_t beam_tiles;
    int32_t chunks_per_item;
    int32_t workgroups_per_cu;
} blade_batch_tuning;

uint64_t blade_beamform_batched_workspace_size(int precision, const blade_beamform_block* blocks,
                                               uint64_t n_blocks, const blade_batch_tuning* tuning);
int blade_beamform_batched(int precision, const blade_beamform_block* blocks, uint64_t n_blocks,
                           const blade_batch_tuning* tuning, void* workspace,
                           uint64_t workspace_bytes, blade_stream_t stream);

/*
 * Adjoint of the beamformer (no reference counterpart: the reference has no
 * gradients; the forward it differentiates is beamformer.cu:13-79).
 * For y[b,f,t,p] = sum_a w[b,a,f,p] x[a,f,t,p] and grad_output = dL/dconj(y)
 * CF32 [B,F,T,P], writes grad_phasors = dL/dconj(w) CF32 [B,A,F,1,P]:
 *     grad_phasors[b,a,f,p] = sum_t grad_output[b,f,t,p] * conj(x[a,f,t,p]).
 * input: CI8 [A,F,T,P]; P = 2, A <= 128.  This is also the gradient in
 * PyTorch's convention when grad_output is PyTorch's gradient of y.
 */
int blade_beamformer_adjoint(const void* input, const void* grad_output, void* grad_phasors,
                             uint64_t A, uint64_t F, uint64_t T, uint64_t P, uint64_t B,
                             blade_stream_t stream);

/*
 * Loss and phasor gradient of the detected beam power used for gain/delay
 * calibration (no reference counterpart):
 *     L = -sum_b beam_weights[b] sum_{f,t,p} |y[b,f,t,p]|^2   (Stokes I)
 * loss: one FP64 value in device memory; grad_phasors = dL/dconj(w) CF32
 * [B,A,F,1,P] (PyTorch's gradient is twice this).  The beams are recomputed
 * tile by tile and never stored.  beam_weights: host array of B floats,
 * B <= 64; P = 2, A <= 128.  precision: BLADE_PREC_F32 (FP32 matrix cores)
 * or BLADE_PREC_F16X2 (FP16 matrix cores with hi + lo phasors and
 * gradients, FP32 accumulation; error at the FP32 level).
 */
int blade_beamformer_stokes_i_grad(int precision, const void* input, const void* phasors,
                                   const float* beam_weights, double* loss, void* grad_phasors,
                                   uint64_t A, uint64_t F, uint64_t T, uint64_t P, uint64_t B,
                                   blade_stream_t stream);

/*
 * Gain and delay steering for calibration:
 *   output[b,a,f,0,p] = phasors[b,a,f,0,p] * u[a,f,p] * exp(-j 2 pi freqs_hz[f] delays[a]),
 * u = gains / |gains| (|gains| floored at 1e-30).  CF32 phasors and output
 * [B,A,F,1,P], CF32 gains [A,F,P] or null (no gain factor), FP64 delays [A] in
 * seconds or null (no delay factor; freqs_hz FP64 [F] needed with delays; the
 * phase is formed in FP64 turns).  P = 2.
 * No reference counterpart: the reference's Phasor applies fixed calibration
 * coefficients (src/modules/phasor/ata.cc:188-204); this makes them unit-modulus
 * gains and delays that calibration can learn.  For gain calibration callers.
 */
int blade_steer_gains(const void* phasors, const void* gains, const double* delays,
                      const double* freqs_hz, void* output, uint64_t B, uint64_t A, uint64_t F,
                      uint64_t P, blade_stream_t stream);

/*
 * Gradient through the gain steering: given grad_phasors = dL/dconj(w_eff) of a
 * real loss for w_eff = phasors * gains / |gains| (as blade_beamformer_stokes_i_grad
 * returns it), grad_gains[a,f,p] = dL/dconj(g) = (G - u^2 conj(G)) / (2|g|) with
 * G = sum_b conj(phasors) grad_phasors and u = g / |g|.  Shapes as blade_steer_gains.
 * No reference counterpart (see blade_steer_gains).  For gain calibration callers.
 */
int blade_gain_gradient(const void* phasors, const void* grad_phasors, const void* gains,
                        void* grad_gains, uint64_t B, uint64_t A, uint64_t F, uint64_t P,
                        blade_stream_t stream);

/*
 * One Adam step on the gains, from grad_phasors as blade_gain_gradient takes it:
 * grad = 2 * grad_scale * dL/dconj(g) (PyTorch's convention for a complex
 * parameter, times a loss scale), then, on the real and imaginary parts of g
 * (CF32 [A,F,P], updated in place) with CF32 exp_avg / exp_avg_sq states of
 * the same shape and *step a device float incremented first:
 *   m = m + (1 - beta1)(grad - m), v = beta2 v + (1 - beta2) grad^2,
 *   g = g - lr / (1 - beta1^t) * m / (sqrt(v) / sqrt(1 - beta2^t) + eps)
 * (torch.optim.Adam without weight decay or amsgrad).  Allocates nothing;
 * capturable in a hipGraph (the step count lives on the device).
 * No reference counterpart (see blade_steer_gains).  For gain calibration callers.
 */
int blade_gain_adam(const void* phasors, const void* grad_phasors, void* gains, void* exp_avg,
                    void* exp_avg_sq, float* step, float lr, float beta1, float beta2, float eps,
                    float grad_scale, uint64_t B, uint64_t A, uint64_t F, uint64_t P,
                    blade_stream_t stream);

/*
 * One Adam step on per-antenna delays (FP64 [A], seconds, updated in place)
 * from steered = blade_steer_gains' output and grad_phasors = dL/dconj(steered):
 *   grad_a = grad_scale * dL/dtau_a = -4 pi grad_scale sum_{b,f,p} freqs_hz[f] Im(conj(steered) grad_phasors)
 * then the FP64 form of blade_gain_adam's update with FP64 exp_avg /
 * exp_avg_sq [A] and its own *step (device float, incremented first).
 * Allocates nothing; capturable.  No reference counterpart (see
 * blade_steer_gains).  For delay calibration callers.
 */
int blade_delay_adam(const void* steered, const void* grad_phasors, const double* freqs_hz,
                     double* delays, double* exp_avg, double* exp_avg_sq, float* step, double lr,
                     double beta1, double beta2, double eps, double grad_scale, uint64_t B,
                     uint64_t A, uint64_t F, uint64_t P, blade_stream_t stream);

/*
 * Channelizer — fftshift(FFT_C) of every C-sample chunk along T.
 * Replaces Modules::Channelizer<CF32,CF32>::process (src/modules/channelizer/base.cc:122-146,
 * pre-kernel channelizer.cu:5-19 and cuFFT C2C forward) with the general-rate
 * semantics of the oracle tests/modules/channelizer/advanced.py:49-65:
 *   out[a, f*C+k, s, p] = fftshift(fft(in[a, f, s*C:(s+1)*C, p]))[k]
 * CF32 [A,F,T,P] -> CF32 [A,F*C,T/C,P]; rate 1 copies; C up to 2^20 and,
 * when C * P > 16384, factoring into 2, 3, 5 and 7 (two passes through a
 * library-owned scratch buffer of the data's size).  in == out (in-place) is
 * supported when C == T, or when C * P > 16384.
 */
int blade_channelizer(const void* input, void* output, uint64_t A, uint64_t F, uint64_t T,
                      uint64_t P, uint64_t rate, blade_stream_t stream);

/*
 * Build (once per device and rate) the twiddle table blade_channelizer uses,
 * so that later calls allocate nothing (e.g. inside hipGraph capture).
 * Replaces the cuFFT plan creation of the Channelizer constructor
 * (src/modules/channelizer/base.cc:58-93).
 */
int blade_channelizer_plan(uint64_t rate, blade_stream_t stream);

/*
 * Channelizer, optional Polarizer and Detector in one pass (P = 2): the ModeH
 * chain (include/blade/bundles/generic/mode_h.hh:55-106: Channelizer ->
 * Polarizer -> Detector) without the CF32 spectra ever reaching memory.  Per
 * fine channel
 *   out[a, f*C+k', g, q] = sum over N consecutive spectra of |X|^2+|Y|^2
 *                          (out_pols 1) or [XX*, YY*, Re XY*, Im XY*] (4),
 * as Detector (src/modules/detector/detector.cu:7-51) applied to the
 * Channelizer output; with polarize != 0, of (L, R) = (X + jY, X - jY)
 * instead (Polarizer, polarizer.cu:5-25, between the two).  N divides
 * S = T/rate (output [A, F*C, S/N, out_pols]), or is a multiple of S (output
 * [A, F*C, 1, out_pols], accumulate != 0 adds this block to it: stepped
 * integration across blocks).
 */
int blade_channelizer_detector(const void* input, void* output, uint64_t A, uint64_t F,
                               uint64_t T, uint64_t rate, uint64_t integration,
                               uint64_t out_pols, int polarize, int accumulate,
                               blade_stream_t stream);

/*
 * Polarizer XY -> LR: out[...,0] = X + jY, out[...,1] = X - jY.
 * Replaces Modules::Polarizer<CF32|CF16>::process (src/modules/polarizer/base.cc:61-68,
 * kernel polarizer.cu:5-25).  type: BLADE_CF32 or BLADE_CF16; n_elements = A*F*T*P
 * (P = 2).  in == out allowed.
 */
int blade_polarizer(int type, const void* input, void* output, uint64_t n_elements,
                    blade_stream_t stream);

/*
 * Detector — integrate `integration` consecutive samples of [X,F,T,2] CF32.
 * Replaces Modules::Detector<CF32,F32>::process (src/modules/detector/base.cc:82-85,
 * kernels detector.cu:7-51): out_pols 1 -> Stokes-I |X|^2+|Y|^2,
 * out_pols 4 -> [XX*, YY*, Re XY*, Im XY*].  Output F32 [X, F, T/integration, out_pols].
 * accumulate != 0 adds into `output` instead of overwriting it (stepped
 * integration across blocks; the reference has none, detector.hh:11).
 */
int blade_detector(const void* input, void* output, uint64_t X, uint64_t F, uint64_t T,
                   uint64_t integration, uint64_t out_pols, int accumulate,
                   blade_stream_t stream);

/*
 * Gather — place one input block at position `step` of `multiplier` along an axis.
 * Replaces Modules::Gather<T,T>::process (src/modules/gather/base.cc:184-225,
 * kernel gather.cu:6-18 and Copy2D): with height = prod(shape[:axis]) and
 * width_bytes = prod(shape[axis:]) * elem_bytes,
 *   out[h, step*width_bytes + i] = in[h, i]   (out row pitch = width_bytes*multiplier).
 * Bit-exact.
 */
int blade_gather(const void* input, void* output, uint64_t height, uint64_t width_bytes,
                 uint64_t multiplier, uint64_t step, blade_stream_t stream);

/*
 * Permutation — out = transpose(in, axes=indexes) for a rank-4 tensor.
 * Replaces Modules::Permutation<T,T>::process (src/modules/permutation/base.cc:319-322,
 * kernel permutation.cu:6-20).  elem_bytes in {1,2,4,8,16}.  Bit-exact.
 */
int blade_permutation(const void* input, void* output, uint64_t elem_bytes,
                      const uint64_t* in_shape4, const uint64_t* indexes4,
                      blade_stream_t stream);

/*
 * ATA phasors from delays on the GPU, FP64 phase argument.
 * Replaces the host loop of Modules::Phasor::ATA<OT>::process
 * (src/modules/phasor/ata.cc:175-208):
 *   w[b,a,f,p] = exp(-j2pi tau[b,a] ((f+fStart) dF + f_obs - BW/2)) * cal[a,f,0,p]
 * delays: F64 [B,A] device; calibration: CF64 [A,F,1,P] device;
 * out_type BLADE_CF32 or BLADE_CF64, output [B,A,F,1,P].
 */
int blade_phasor_ata(const void* delays, const void* calibration, void* phasors, int out_type,
                     uint64_t B, uint64_t A, uint64_t F, uint64_t P,
                     double observation_frequency_hz, double channel_bandwidth_hz,
                     double total_bandwidth_hz, uint64_t frequency_start_index,
                     blade_stream_t stream);

/*
 * Geometric delays of the ATA phasor module, on the host (no device memory,
 * no stream).  Replaces the delay half of Modules::Phasor::ATA<OT>::process
 * (src/modules/phasor/ata.cc:105-173, which calls radiointerferometryc99 and
 * ERFA):
 *   delays[b,a] = (w_a - w_ref)/c for beam b  -  (w_a - w_ref)/c for the boresight,
 * w = cos(dec)cos(H) X - cos(dec)sin(H) Y + sin(dec) Z over the antenna's
 * baseline (X, Y, Z) from the array reference position in the local frame
 * (X to the meridian at the equator, Y east, Z north), dec and RA of date
 * (J2000 precessed with the IAU 1976 angles), H = GMST(UT1, IAU 1982) +
 * longitude - RA.  Nutation and aberration are not modelled.
 * antenna_ecef: [A][3] metres; reference_lla: {lon, lat (rad), alt (m), WGS84};
 * boresight_radec: {ra, dec} rad (J2000); beam_radec: [B][2] rad;
 * julian_date: UTC; dut1: UT1 - UTC in seconds; delays: [B][A] seconds.
 */
int blade_ata_delays(const double* antenna_ecef, uint64_t A, uint64_t reference_antenna,
                     const double* reference_lla, const double* boresight_radec,
                     const double* beam_radec, uint64_t B, double julian_date, double dut1,
                     double* delays);

#ifdef __cplusplus
}
#endif

#endif /* BLADE_AMD_H */
