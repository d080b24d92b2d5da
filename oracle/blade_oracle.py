"""Numpy restatement of the reference's ATA beamforming path (test oracle).

TEST INFRASTRUCTURE ONLY: imported by ``tests/``, ``__graft_entry__.smoke()``
and ``bench.py``'s ``cpu_baseline`` leg, never by the product path.

Layouts are the reference's row-major tensors
(``include/blade/memory/custom.hh:10-163``):

* voltages / beams ``[A|B, F, T, P]`` (ArrayShape)
* phasors ``[B, A, F, 1, P]`` (PhasorShape)
* delays ``[B, A]`` (DelayShape)

All arithmetic is complex128 / float64 unless a function says otherwise,
so the oracle doubles as the FP64 reference for the error budget.
"""

from __future__ import annotations

import numpy as np

__all__ = [
    "ci8_gaussian", "ci8_to_complex", "complex_to_ci8",
    "phasor_ata", "phasor_ata_reference_test_form",
    "beamformer_ata", "channelizer", "polarizer", "detector",
    "detector_stepped", "cast", "gather", "permutation",
    "mode_b", "mode_h", "beamform_chain_c64", "norm_rel_err", "max_rms_rel_err",
    "beamformed_stokes_i_loss", "stokes_i_loss_grad_gains",
    "beamformer_adjoint", "stokes_i_loss_and_grad_phasors",
    "NUMPY_TYPES",
]

# ---------------------------------------------------------------------------
# Data types.  CI8 is stored as int8 pairs (re, im) -> trailing axis of 2.
# ---------------------------------------------------------------------------

NUMPY_TYPES = {
    "i8": np.int8, "i32": np.int32, "f16": np.float16, "f32": np.float32,
    "f64": np.float64, "cf16": np.float16, "cf32": np.complex64,
    "cf64": np.complex128, "ci8": np.int8,
}


def ci8_gaussian(shape, sigma=16.0, rng=None):
    """CI8 voltages: each component N(0, sigma^2), rounded, clipped to [-127, 127].

    Returns int8 array of shape ``shape + (2,)`` (re, im interleaved, the
    reference's CI8 layout, ``include/blade/memory/types.hh``).
    """
    rng = np.random.default_rng(rng)
    v = rng.normal(0.0, sigma, size=tuple(shape) + (2,))
    return np.clip(np.rint(v), -127, 127).astype(np.int8)


def ci8_to_complex(x):
    """int8 ``[..., 2]`` -> complex128 (exact)."""
    x = np.asarray(x)
    return x[..., 0].astype(np.float64) + 1j * x[..., 1].astype(np.float64)


def complex_to_ci8(z):
    """Quantise complex values to CI8 (round half to even, clip to +-127)."""
    z = np.asarray(z)
    out = np.empty(z.shape + (2,), dtype=np.int8)
    out[..., 0] = np.clip(np.rint(z.real), -127, 127)
    out[..., 1] = np.clip(np.rint(z.imag), -127, 127)
    return out


# ---------------------------------------------------------------------------
# Phasor (a1)
# ---------------------------------------------------------------------------

def phasor_ata(delays, calibration, observation_frequency_hz, channel_bandwidth_hz,
               total_bandwidth_hz, frequency_start_index):
    """Phasors from per-(beam, antenna) delays.

    Follows ``src/modules/phasor/ata.cc:175-208``::

        fringe = f_obs - BW/2
        w[b,a,f,p] = exp(-j2pi*tau[b,a]*(f+fStart)*dF - j2pi*tau[b,a]*fringe) * cal[a,f,0,p]

    ``delays``: float64 ``[B, A]`` (seconds); ``calibration``: complex
    ``[A, F, 1, P]``.  Returns complex128 ``[B, A, F, 1, P]``.

    Pinned by: tau = 0 gives cal; unit modulus when cal = 1; phase linear in
    f; agreement with the reference test's product-of-exponentials form
    (``tests/modules/phasor/advanced.py:284-295,339-357``).
    """
    tau = np.asarray(delays, dtype=np.float64)
    cal = np.asarray(calibration, dtype=np.complex128)
    A, F, _, P = cal.shape
    fringe = observation_frequency_hz - total_bandwidth_hz / 2.0
    freq = (np.arange(F, dtype=np.float64) + frequency_start_index) * channel_bandwidth_hz
    phase = -2.0 * np.pi * tau[:, :, None] * (freq[None, None, :] + fringe)   # [B, A, F]
    ph = np.exp(1j * phase)
    return ph[:, :, :, None, None] * cal[None, :, :, :, :]


def phasor_ata_reference_test_form(delays, calibration, observation_frequency_hz,
                                   channel_bandwidth_hz, total_bandwidth_hz,
                                   frequency_start_index):
    """The reference test's independent formulation (product of two exponentials).

    ``tests/modules/phasor/advanced.py:284-295`` (``create_delay_phasors`` and
    ``get_fringe_rate``) and ``:339-357`` (per-pol calibration).
    """
    tau = np.asarray(delays, dtype=np.float64)
    cal = np.asarray(calibration, dtype=np.complex128)
    B, A = tau.shape
    _, F, _, P = cal.shape
    out = np.zeros((B, A, F, 1, P), dtype=np.complex128)
    freqs = np.arange(frequency_start_index, frequency_start_index + F) * channel_bandwidth_hz
    for b in range(B):
        for a in range(A):
            d = np.exp(-1j * 2 * np.pi * tau[b, a] * freqs)
            d = d * np.exp(-1j * 2 * np.pi * tau[b, a] *
                           (observation_frequency_hz - total_bandwidth_hz / 2.0))
            for p in range(P):
                out[b, a, :, 0, p] = d * cal[a, :, 0, p]
    return out


# ---------------------------------------------------------------------------
# Beamformer (a4)
# ---------------------------------------------------------------------------

def beamformer_ata(x, phasors, enable_incoherent_beam=False, enable_incoherent_beam_sqrt=False):
    """ATA beamformer.

    Follows ``src/modules/beamformer/beamformer.cu:13-79`` and the test oracle
    ``tests/modules/beamformer/ata.py:56-63``::

        y[b,f,t,p] = sum_a x[a,f,t,p] * w[b,a,f,0,p]         (no conjugate)
        y[B,f,t,p] = (sum_a |x[a,f,t,p] * w[0,a,f,0,p]|^2, 0)  (optionally sqrt'ed)

    The incoherent beam uses beam 0's phasors like the kernel
    (``beamformer.cu:68-69``), not the last beam like the test (quirk Q1).

    ``x``: complex ``[A, F, T, P]`` (or int8 ``[A, F, T, P, 2]`` CI8).
    Returns complex128 ``[B(+1), F, T, P]``.

    Pinned by: brute-force loops at tiny shapes; linearity in x and w;
    A = 1 gives x*w; w = 1 gives sum_a x; incoherent = sum |x|^2 |w0|^2.
    """
    x = _as_complex(x)
    w = np.asarray(phasors, dtype=np.complex128)[:, :, :, 0, :]   # [B, A, F, P]
    y = np.einsum("bafp,aftp->bftp", w, x, optimize=True)
    if not enable_incoherent_beam:
        return y
    inc = np.einsum("afp,aftp->ftp", np.abs(w[0]) ** 2, np.abs(x) ** 2, optimize=True)
    if enable_incoherent_beam_sqrt:
        inc = np.sqrt(inc)
    return np.concatenate([y, inc[None].astype(np.complex128)], axis=0)


# ---------------------------------------------------------------------------
# Channelizer (a5)
# ---------------------------------------------------------------------------

def channelizer(x, rate):
    """FFT channelizer with fftshift, general rate C.

    Follows the oracle ``tests/modules/channelizer/advanced.py:49-65``::

        out[a, f*C + k, s, p] = fftshift(fft(x[a, f, s*C:(s+1)*C, p]))[k]

    (unnormalised forward DFT, ``e^{-j2pi kn/C}``, the sign of cuFFT's
    forward transform used at ``src/modules/channelizer/base.cc:134-138``).
    ``rate == 1`` is the identity (``base.cc:124-126``).
    Returns complex128 ``[A, F*C, T/C, P]``.

    Pinned by: explicit DFT-matrix brute force; Parseval (x C); impulse gives
    a flat spectrum; a tone at bin k lands at (k + C/2) mod C; C = 1 identity.
    """
    x = _as_complex(x)
    A, F, T, P = x.shape
    C = int(rate)
    if T % C:
        raise ValueError("T must be divisible by the channelizer rate")
    if C == 1:
        return x.copy()
    S = T // C
    v = x.reshape(A, F, S, C, P)
    spec = np.fft.fftshift(np.fft.fft(v, axis=3), axes=3)       # [A, F, S, C, P]
    return np.ascontiguousarray(spec.transpose(0, 1, 3, 2, 4)).reshape(A, F * C, S, P)


# ---------------------------------------------------------------------------
# Polarizer (a6)
# ---------------------------------------------------------------------------

def polarizer(x, input_polarization="xy", output_polarization="lr"):
    """XY -> LR (unnormalised).

    Follows ``src/modules/polarizer/polarizer.cu:5-25`` and the oracle
    ``tests/modules/polarizer/polarizer.py:47-49``::

        out[..., 0] = X + jY ;  out[..., 1] = X - jY

    Bypassed (identity) when input == output polarisation
    (``src/modules/polarizer/base.cc:62-68``).
    """
    x = _as_complex(x)
    if input_polarization == output_polarization:
        return x.copy()
    out = np.empty_like(x)
    out[..., 0] = x[..., 0] + 1j * x[..., 1]
    out[..., 1] = x[..., 0] - 1j * x[..., 1]
    return out


# ---------------------------------------------------------------------------
# Detector (a7)
# ---------------------------------------------------------------------------

def detector(x, integration_size, number_of_output_polarizations):
    """Power detector integrating ``integration_size`` consecutive samples.

    Follows ``src/modules/detector/detector.cu:7-51`` and the oracles
    ``tests/modules/detector/single_pol.py:56-67`` (Stokes-I,
    ``|X|^2 + |Y|^2``) and ``cross_pol.py:56-71`` (``[XX*, YY*, Re XY*, Im XY*]``).

    ``x``: complex ``[X, F, T, 2]``.  Returns float64 ``[X, F, T/INTG, 1|4]``.

    Pinned by: brute-force loops; 4-pol[0] + 4-pol[1] equals the 1-pol output;
    auto powers are non-negative; |XY*|^2 <= XX*YY* (Cauchy-Schwarz) for INTG=1.
    """
    x = _as_complex(x)
    Xn, F, T, P = x.shape
    if P != 2:
        raise ValueError("detector needs 2 input polarisations")
    n = int(integration_size)
    if n <= 0 or T % n:
        raise ValueError("T must be divisible by the integration size")
    X = x[..., 0].reshape(Xn, F, T // n, n)
    Y = x[..., 1].reshape(Xn, F, T // n, n)
    xx = (X.real ** 2 + X.imag ** 2).sum(-1)
    yy = (Y.real ** 2 + Y.imag ** 2).sum(-1)
    if number_of_output_polarizations == 1:
        return (xx + yy)[..., None]
    if number_of_output_polarizations == 4:
        xy = (X * np.conj(Y)).sum(-1)
        return np.stack([xx, yy, xy.real, xy.imag], axis=-1)
    raise ValueError("number_of_output_polarizations must be 1 or 4")


def detector_stepped(blocks, integration_size, number_of_output_polarizations):
    """Detector whose integration spans several consecutive input blocks.

    The reference integrates only within one block
    (``include/blade/modules/detector.hh:11``); stepped integration is the
    same sum taken over the concatenation of the blocks along T, which is
    exactly what this restates.
    """
    return detector(np.concatenate([_as_complex(b) for b in blocks], axis=2),
                    integration_size, number_of_output_polarizations)


# ---------------------------------------------------------------------------
# Cast / Gather / Permutation (a3, a8, a9) - bit-exact data movement
# ---------------------------------------------------------------------------

def cast(x, in_type, out_type):
    """Component-wise ``static_cast`` (``src/modules/cast/cast.cu:5-12``).

    Complex types are cast per scalar component (``cast/base.cc:20-37``
    launches over ``size * cudaSize`` scalars); casting between complex and
    non-complex raises like ``cast/base.cc:39-43``.  ``x`` is given as the
    scalar-component numpy array (CI8 / CF16 with a trailing axis of 2,
    CF32/CF64 as numpy complex).
    """
    in_c = in_type.startswith("c")
    out_c = out_type.startswith("c")
    if in_c != out_c:
        raise ValueError("Cannot cast between complex and non-complex types.")
    comp = _components(x, in_type)
    scalar = {"ci8": np.int8, "cf16": np.float16, "cf32": np.float32, "cf64": np.float64,
              "i8": np.int8, "i32": np.int32, "f16": np.float16, "f32": np.float32,
              "f64": np.float64}[out_type]
    if np.issubdtype(scalar, np.integer) and np.issubdtype(comp.dtype, np.floating):
        # C++ static_cast float->int truncates toward zero.
        res = np.trunc(comp).astype(scalar)
    else:
        res = comp.astype(scalar)
    if out_type in ("cf32", "cf64"):
        return res[..., 0] + 1j * res[..., 1] if out_type == "cf64" else \
            (res[..., 0] + 1j * res[..., 1]).astype(np.complex64)
    return res


def gather(blocks, axis):
    """Concatenate ``multiplier`` successive blocks along ``axis``.

    Follows ``src/modules/gather/base.cc:184-225`` (block ``step`` lands at
    offset ``step * shape[axis]``) and the oracle
    ``tests/modules/gather/gather.py:54-55`` (``np.concatenate``).
    """
    return np.concatenate(list(blocks), axis=axis)


def permutation(x, indexes):
    """``np.transpose(x, axes=indexes)`` (``tests/modules/permutation/permutation.py:49``;
    ``src/modules/permutation/permutation.cu:6-20``: out[perm(coords)] = in[coords]
    with ``permuted[d] = coords[indexes[d]]``)."""
    return np.ascontiguousarray(np.transpose(x, axes=indexes))


# ---------------------------------------------------------------------------
# Bundles (a10)
# ---------------------------------------------------------------------------

def mode_b(x, phasors, channelizer_rate=1, convert_to_circular=False,
           incoherent_beam=False, detector_enable=False, detector_integration_size=1,
           detector_number_of_output_polarizations=1):
    """ATA Mode-B composition (``include/blade/bundles/ata/mode_b.hh:99-199``):

    Cast -> Channelizer -> Polarizer -> (Phasor) -> Beamformer -> [Detector].
    The beamformer's incoherent-beam sqrt follows ``detectorEnable``
    (``mode_b.hh:156``).
    """
    v = channelizer(_as_complex(x), channelizer_rate)
    v = polarizer(v, "xy", "lr" if convert_to_circular else "xy")
    y = beamformer_ata(v, phasors, incoherent_beam, detector_enable)
    if detector_enable:
        return detector(y, detector_integration_size, detector_number_of_output_polarizations)
    return y


def mode_h(x, convert_to_circular=False, detector_integration_size=1,
           detector_number_of_output_polarizations=1):
    """Generic Mode-H (``include/blade/bundles/generic/mode_h.hh:55-106``):
    Cast -> Channelizer(rate = T) -> Polarizer -> Detector."""
    x = _as_complex(x)
    v = channelizer(x, x.shape[2])
    v = polarizer(v, "xy", "lr" if convert_to_circular else "xy")
    return detector(v, detector_integration_size, detector_number_of_output_polarizations)


def beamform_chain_c64(x, phasors, rate=1, polarize=False):
    """Vectorised complex64 CPU port of beamform -> channelize -> [polarize].

    The same composition as ``beamformer_ata`` (``beamformer.cu:13-79``) ->
    ``channelizer`` (``advanced.py:49-65``) -> ``polarizer``
    (``polarizer.py:47-49``) without the incoherent beam, written for CPU
    speed (batched BLAS cgemm over (f, p), numpy FFT) in the precision of the
    GPU's FP32 path.  It is the timed CPU baseline; it is pinned against the
    complex128 functions above by ``tests/test_oracle.py``.
    """
    x = np.asarray(x)
    if x.dtype == np.int8:
        x = (x[..., 0].astype(np.float32) + 1j * x[..., 1].astype(np.float32)).astype(np.complex64)
    x = x.astype(np.complex64, copy=False)
    A, F, T, P = x.shape
    w = np.asarray(phasors).astype(np.complex64)[:, :, :, 0, :]                # [B, A, F, P]
    B = w.shape[0]
    xt = np.ascontiguousarray(x.transpose(1, 3, 0, 2))                          # [F, P, A, T]
    wt = np.ascontiguousarray(w.transpose(2, 3, 0, 1))                          # [F, P, B, A]
    y = np.ascontiguousarray(np.matmul(wt, xt).transpose(2, 0, 3, 1))           # [B, F, T, P]
    C = int(rate)
    if C > 1:
        S = T // C
        spec = np.fft.fftshift(np.fft.fft(y.reshape(B, F, S, C, P), axis=3), axes=3)
        y = np.ascontiguousarray(spec.transpose(0, 1, 3, 2, 4)).reshape(B, F * C, S, P)
        y = y.astype(np.complex64, copy=False)
    if polarize:
        o = np.empty_like(y)
        o[..., 0] = y[..., 0] + 1j * y[..., 1]
        o[..., 1] = y[..., 0] - 1j * y[..., 1]
        y = o
    return y


# ---------------------------------------------------------------------------
# Calibration loss and gradient (closed form, complex128)
# ---------------------------------------------------------------------------

def beamformed_stokes_i_loss(x, gains, phasors, beam=0):
    """Negative Stokes-I power of one beam with per-antenna complex gains.

    ``y[f,t,p] = sum_a g[a,f,p] * w[beam,a,f,p] * x[a,f,t,p]`` and
    ``L = -sum_{f,t,p} |y|^2``.
    """
    x = _as_complex(x)
    w = np.asarray(phasors, dtype=np.complex128)[beam, :, :, 0, :]   # [A, F, P]
    g = np.asarray(gains, dtype=np.complex128)                       # [A, F, P]
    y = np.einsum("afp,aftp->ftp", g * w, x, optimize=True)
    return -float(np.sum(y.real ** 2 + y.imag ** 2))


def beamformer_adjoint(x, grad_output):
    """Adjoint of the beamformer with respect to the phasors.

    For ``y[b,f,t,p] = sum_a w[b,a,f,p] x[a,f,t,p]`` (``beamformer_ata``) and
    ``grad_output = dL/dconj(y)`` ``[B, F, T, P]``, returns
    ``dL/dconj(w)[b,a,f,0,p] = sum_t grad_output[b,f,t,p] conj(x[a,f,t,p])``.

    Pinned by the defining identity of the adjoint,
    ``Re <gy, J dw> = Re <adj(gy), dw>`` for random ``dw`` (``J`` the
    beamformer's linear map in ``w``), and by finite differences of the
    Stokes-I loss, whose gradient is ``2 * adj(-y)``.
    """
    x = _as_complex(x)
    gy = np.asarray(grad_output, dtype=np.complex128)
    g = np.einsum("bftp,aftp->bafp", gy, np.conj(x), optimize=True)
    return g[:, :, :, None, :]


def stokes_i_loss_and_grad_phasors(x, phasors, beam_weights):
    """``L = -sum_b lambda_b sum_{f,t,p} |y_b|^2`` and ``dL/dconj(w)``.

    ``dL/dconj(y_b) = -lambda_b y_b``, so the phasor gradient is
    ``beamformer_adjoint(x, -lambda * y)``.
    """
    lam = np.asarray(beam_weights, dtype=np.float64)
    y = beamformer_ata(x, phasors)
    loss = -float(np.sum(lam[:, None, None, None] * (y.real ** 2 + y.imag ** 2)))
    return loss, beamformer_adjoint(x, -lam[:, None, None, None] * y)


def stokes_i_loss_grad_gains(x, gains, phasors, beam=0):
    """Closed-form gradient of ``beamformed_stokes_i_loss`` w.r.t. the gains.

    Returned in PyTorch's convention for a real loss of complex inputs,
    ``dL/dRe g + j dL/dIm g``:  ``-2 * sum_t conj(w x) y``.
    """
    x = _as_complex(x)
    w = np.asarray(phasors, dtype=np.complex128)[beam, :, :, 0, :]
    g = np.asarray(gains, dtype=np.complex128)
    y = np.einsum("afp,aftp->ftp", g * w, x, optimize=True)
    return -2.0 * np.einsum("afp,aftp,ftp->afp", np.conj(w), np.conj(x), y, optimize=True)


# ---------------------------------------------------------------------------
# Error metrics
# ---------------------------------------------------------------------------

def norm_rel_err(y, ref):
    """Normwise relative error ``||y - ref||_2 / ||ref||_2``."""
    y = np.asarray(y, dtype=np.complex128)
    ref = np.asarray(ref, dtype=np.complex128)
    den = np.linalg.norm(ref.ravel())
    return float(np.linalg.norm((y - ref).ravel()) / (den if den > 0 else 1.0))


def max_rms_rel_err(y, ref):
    """``max|y - ref| / rms(ref)``."""
    y = np.asarray(y, dtype=np.complex128)
    ref = np.asarray(ref, dtype=np.complex128)
    rms = np.sqrt(np.mean(np.abs(ref) ** 2))
    return float(np.max(np.abs(y - ref)) / (rms if rms > 0 else 1.0))


# ---------------------------------------------------------------------------
# helpers
# ---------------------------------------------------------------------------

def _as_complex(x):
    x = np.asarray(x)
    if x.dtype == np.int8:
        return ci8_to_complex(x)
    return x.astype(np.complex128, copy=False)


def _components(x, t):
    x = np.asarray(x)
    if t in ("cf32", "cf64"):
        return np.stack([x.real, x.imag], axis=-1)
    return x
