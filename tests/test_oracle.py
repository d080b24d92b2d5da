"""Pin the numpy oracle (CPU only).

Every oracle function is checked against formulations that do not share its
code: explicit loops, DFT matrices, closed forms and invariants, chosen so
that a dropped term, a wrong sign/index or a transposed operand fails.
"""

import numpy as np
import pytest

import oracle as O

RNG = np.random.default_rng(1234)


def crand(*shape):
    return RNG.normal(size=shape) + 1j * RNG.normal(size=shape)


# --------------------------------------------------------------------- phasor

PH = dict(observation_frequency_hz=6500.125e6, channel_bandwidth_hz=0.5e6,
          total_bandwidth_hz=1.024e9, frequency_start_index=352)


def test_phasor_matches_reference_test_formulation():
    # tests/modules/phasor/advanced.py:284-295,339-357 multiplies two
    # exponentials; the module (ata.cc:186-204) exponentiates the sum.
    tau = RNG.uniform(-5e-8, 5e-8, size=(3, 5))
    cal = crand(5, 7, 1, 2)
    a = O.phasor_ata(tau, cal, **PH)
    b = O.phasor_ata_reference_test_form(tau, cal, **PH)
    assert np.max(np.abs(a - b)) < 1e-9


def test_phasor_special_cases():
    cal = crand(4, 6, 1, 2)
    # tau = 0 -> the calibration itself
    w = O.phasor_ata(np.zeros((2, 4)), cal, **PH)
    assert np.array_equal(w[0], cal) and np.array_equal(w[1], cal)
    # cal = 1 -> unit modulus, phase linear in f with slope -2 pi tau dF
    tau = np.array([[1e-8, -3e-9, 2.5e-8, 0.0]])
    w = O.phasor_ata(tau, np.ones((4, 6, 1, 2)), **PH)
    assert np.allclose(np.abs(w), 1.0, atol=1e-12)
    ratio = w[0, :, 1:, 0, 0] / w[0, :, :-1, 0, 0]
    expect = np.exp(-2j * np.pi * tau[0][:, None] * PH["channel_bandwidth_hz"])
    assert np.allclose(ratio, expect, atol=1e-9)
    # f = 0 phase: -2 pi tau (fStart dF + f_obs - BW/2)
    f0 = PH["frequency_start_index"] * PH["channel_bandwidth_hz"] + \
        PH["observation_frequency_hz"] - PH["total_bandwidth_hz"] / 2
    assert np.allclose(w[0, :, 0, 0, 1], np.exp(-2j * np.pi * tau[0] * f0), atol=1e-9)


# ----------------------------------------------------------------- beamformer

def beamformer_loops(x, w, inc, sq):
    A, F, T, P = x.shape
    B = w.shape[0]
    y = np.zeros((B + (1 if inc else 0), F, T, P), dtype=np.complex128)
    for b in range(B):
        for f in range(F):
            for t in range(T):
                for p in range(P):
                    s = 0j
                    for a in range(A):
                        s += x[a, f, t, p] * w[b, a, f, 0, p]
                    y[b, f, t, p] = s
    if inc:
        for f in range(F):
            for t in range(T):
                for p in range(P):
                    s = 0.0
                    for a in range(A):
                        z = x[a, f, t, p] * w[0, a, f, 0, p]
                        s += z.real * z.real + z.imag * z.imag   # detect() of beamformer.cu:9-11
                    y[B, f, t, p] = np.sqrt(s) if sq else s
    return y


@pytest.mark.parametrize("inc,sq", [(False, False), (True, False), (True, True)])
def test_beamformer_brute_force(inc, sq):
    x = crand(3, 2, 5, 2)
    w = crand(4, 3, 2, 1, 2)
    assert np.allclose(O.beamformer_ata(x, w, inc, sq), beamformer_loops(x, w, inc, sq), atol=1e-12)


def test_beamformer_invariants():
    x = crand(5, 3, 8, 2)
    w = crand(2, 5, 3, 1, 2)
    # linearity in x and in w
    x2 = crand(5, 3, 8, 2)
    assert np.allclose(O.beamformer_ata(2 * x + x2, w), 2 * O.beamformer_ata(x, w) + O.beamformer_ata(x2, w))
    w2 = crand(2, 5, 3, 1, 2)
    assert np.allclose(O.beamformer_ata(x, w - 3j * w2), O.beamformer_ata(x, w) - 3j * O.beamformer_ata(x, w2))
    # A = 1 -> x * w (no conjugation)
    y = O.beamformer_ata(x[:1], w[:, :1])
    assert np.allclose(y, x[0][None] * w[:, 0, :, :, :])
    # w = 1 -> sum over antennas
    y = O.beamformer_ata(x, np.ones((1, 5, 3, 1, 2)))
    assert np.allclose(y[0], x.sum(0))
    # incoherent beam uses beam 0 (beamformer.cu:68-69), not the last one
    y = O.beamformer_ata(x, w, True, False)
    expect = (np.abs(x) ** 2 * np.abs(w[0, :, :, 0, :])[:, :, None, :] ** 2).sum(0)
    assert np.allclose(y[-1].real, expect) and np.all(y[-1].imag == 0)


def test_ci8_widening_is_exact():
    v = np.arange(-128, 128, dtype=np.int8)
    z = O.ci8_to_complex(np.stack([v, v[::-1]], -1))
    assert np.array_equal(z.real, v.astype(np.float64))
    assert np.array_equal(z.imag, v[::-1].astype(np.float64))


def test_ci8_generator_distribution():
    x = O.ci8_gaussian((64, 64, 64), sigma=16, rng=0)
    assert x.dtype == np.int8 and x.shape == (64, 64, 64, 2)
    assert x.min() >= -127 and x.max() <= 127
    assert abs(x.astype(float).std() - 16) < 0.2


# ---------------------------------------------------------------- channelizer

def channelizer_dft_matrix(x, C):
    A, F, T, P = x.shape
    S = T // C
    n = np.arange(C)
    M = np.exp(-2j * np.pi * np.outer(n, n) / C)       # X[k] = sum_n x[n] e^{-j2pi kn/C}
    out = np.zeros((A, F * C, S, P), dtype=np.complex128)
    for a in range(A):
        for f in range(F):
            for s in range(S):
                for p in range(P):
                    X = M @ x[a, f, s * C:(s + 1) * C, p]
                    for k in range(C):
                        out[a, f * C + (k + C // 2) % C, s, p] = X[k]
    return out


@pytest.mark.parametrize("C", [2, 4, 6, 8, 10, 16])
def test_channelizer_dft_matrix(C):
    x = crand(2, 3, 4 * C, 2)
    assert np.allclose(O.channelizer(x, C), channelizer_dft_matrix(x, C), atol=1e-10)


def test_channelizer_invariants():
    C = 16
    x = crand(1, 2, C * 3, 1)
    y = O.channelizer(x, C)
    # Parseval: sum |X|^2 = C sum |x|^2 (per spectrum, here over all)
    assert np.isclose((np.abs(y) ** 2).sum(), C * (np.abs(x) ** 2).sum())
    # impulse at n = 0 -> flat spectrum of ones
    imp = np.zeros((1, 1, C, 1), complex)
    imp[0, 0, 0, 0] = 1
    assert np.allclose(O.channelizer(imp, C), 1.0)
    # tone at bin k -> single bin at (k + C/2) mod C, amplitude C
    for k in (0, 1, 5, C - 1):
        tone = np.exp(2j * np.pi * k * np.arange(C) / C).reshape(1, 1, C, 1)
        y = O.channelizer(tone, C)[0, :, 0, 0]
        j = (k + C // 2) % C
        assert np.isclose(y[j], C) and np.allclose(np.delete(y, j), 0, atol=1e-9)
    # rate 1 is the identity
    assert np.array_equal(O.channelizer(x, 1), x)


# ------------------------------------------------------------------ polarizer

def test_polarizer_closed_form():
    x = crand(2, 3, 4, 2)
    y = O.polarizer(x)
    X, Y = x[..., 0], x[..., 1]
    # explicit real arithmetic of polarizer.cu:18-21
    assert np.allclose(y[..., 0].real, X.real - Y.imag) and np.allclose(y[..., 0].imag, X.imag + Y.real)
    assert np.allclose(y[..., 1].real, X.real + Y.imag) and np.allclose(y[..., 1].imag, X.imag - Y.real)
    # invertible: X = (L + R)/2, Y = (L - R)/(2j)
    assert np.allclose((y[..., 0] + y[..., 1]) / 2, X)
    assert np.allclose((y[..., 0] - y[..., 1]) / 2j, Y)
    # bypass when input == output polarisation
    assert np.array_equal(O.polarizer(x, "xy", "xy"), x)


# ------------------------------------------------------------------- detector

def detector_loops(x, n, npol):
    Xn, F, T, _ = x.shape
    out = np.zeros((Xn, F, T // n, npol))
    for b in range(Xn):
        for f in range(F):
            for s in range(T // n):
                X = x[b, f, s * n:(s + 1) * n, 0]
                Y = x[b, f, s * n:(s + 1) * n, 1]
                if npol == 1:
                    out[b, f, s, 0] = np.sum(X.real ** 2 + X.imag ** 2) + np.sum(Y.real ** 2 + Y.imag ** 2)
                else:
                    c = X * np.conj(Y)
                    out[b, f, s] = [np.sum((X * np.conj(X)).real), np.sum((Y * np.conj(Y)).real),
                                    np.sum(c.real), np.sum(c.imag)]
    return out


@pytest.mark.parametrize("n,npol", [(1, 1), (10, 1), (1, 4), (10, 4), (4, 4)])
def test_detector_brute_force(n, npol):
    x = crand(2, 3, 20, 2)
    assert np.allclose(O.detector(x, n, npol), detector_loops(x, n, npol))


def test_detector_invariants():
    x = crand(2, 3, 16, 2)
    d4 = O.detector(x, 4, 4)
    d1 = O.detector(x, 4, 1)
    assert np.allclose(d4[..., 0] + d4[..., 1], d1[..., 0])
    assert np.all(d4[..., 0] >= 0) and np.all(d4[..., 1] >= 0)
    e = O.detector(x, 1, 4)
    assert np.all(e[..., 2] ** 2 + e[..., 3] ** 2 <= e[..., 0] * e[..., 1] * (1 + 1e-12))
    # stepped integration over two blocks == one block of twice the length
    a, b = crand(1, 2, 8, 2), crand(1, 2, 8, 2)
    assert np.allclose(O.detector_stepped([a, b], 16, 1), O.detector(np.concatenate([a, b], 2), 16, 1))
    with pytest.raises(ValueError):
        O.detector(x, 5, 1)


# ----------------------------------------------------- cast/gather/permutation

def test_cast_exhaustive_int8():
    v = np.arange(-128, 128, dtype=np.int8)
    assert np.array_equal(O.cast(v, "i8", "f32"), v.astype(np.float32))
    ci8 = np.stack([v, v[::-1]], -1)
    z = O.cast(ci8, "ci8", "cf32")
    assert z.dtype == np.complex64
    assert np.array_equal(z.real, v.astype(np.float32)) and np.array_equal(z.imag, v[::-1].astype(np.float32))
    h = O.cast(ci8, "ci8", "cf16")
    assert h.dtype == np.float16 and np.array_equal(h.astype(np.int32), ci8.astype(np.int32))


def test_cast_float_rules():
    f = np.array([1.5, -1.5, 2.7, -2.7, 65519.0, 1e-8, 3.14159], dtype=np.float32)
    assert np.array_equal(O.cast(f, "f32", "i32"), np.array([1, -1, 2, -2, 65519, 0, 3], np.int32))
    h = O.cast(f, "f32", "f16")
    # round to nearest even, same as IEEE conversion
    assert np.array_equal(h, f.astype(np.float16))
    with pytest.raises(ValueError):
        O.cast(f, "f32", "cf32")


def test_gather_and_permutation():
    blocks = [crand(2, 3, 4, 2) for _ in range(3)]
    for axis in range(4):
        g = O.gather(blocks, axis)
        # block k occupies [k*n, (k+1)*n) along axis
        n = blocks[0].shape[axis]
        for k, blk in enumerate(blocks):
            sl = [slice(None)] * 4
            sl[axis] = slice(k * n, (k + 1) * n)
            assert np.array_equal(g[tuple(sl)], blk)
    x = crand(2, 3, 4, 5)
    for idx in ([0, 1, 2, 3], [0, 2, 1, 3], [3, 2, 1, 0], [2, 0, 3, 1]):
        y = O.permutation(x, idx)
        # permutation.cu:13-17: out[permuted] = in[orig], permuted[d] = orig[idx[d]]
        for c in np.ndindex(x.shape):
            assert y[tuple(c[i] for i in idx)] == x[c]


# -------------------------------------------------------------------- bundles

def test_mode_b_one_antenna_unit_phasors_is_channelize_then_polarize():
    # mode_b.hh:104-199 runs Cast -> Channelizer -> Polarizer -> Beamformer.
    # With one antenna and w = 1 the beamformer is the identity, so the
    # output is the polarized spectra of x: a tone e^{j2pi k0 n/C} in X only
    # lands in fine channel (k0 + C/2) % C with amplitude C, in both L and R.
    C, k0 = 4, 1
    n = np.arange(16)
    x = np.zeros((1, 2, 16, 2), complex)
    x[0, :, :, 0] = np.exp(2j * np.pi * k0 * n / C)
    w = np.ones((3, 1, 2 * C, 1, 2))
    y = O.mode_b(x, w, channelizer_rate=C, convert_to_circular=True)
    assert y.shape == (3, 2 * C, 4, 2)
    exp = np.zeros((3, 2 * C, 4, 2), complex)
    for f in range(2):
        exp[:, f * C + (k0 + C // 2) % C, :, :] = C       # L = X + jY, R = X - jY, Y = 0
    assert np.allclose(y, exp, atol=1e-12)
    # a tone in Y: L = +jC, R = -jC
    x2 = x[..., ::-1].copy()
    y2 = O.mode_b(x2, w, channelizer_rate=C, convert_to_circular=True)
    assert np.allclose(y2[..., 0], 1j * exp[..., 0], atol=1e-12)
    assert np.allclose(y2[..., 1], -1j * exp[..., 1], atol=1e-12)


def test_mode_b_polarizes_before_beamforming():
    # Polarization-dependent phasors tell the two orders apart: the
    # polarizer mixes X and Y before each pol's own phasor is applied.
    x = crand(1, 1, 4, 2)
    a, b = 0.5 - 2j, 1.5 + 0.25j
    w = np.zeros((1, 1, 1, 1, 2), complex)
    w[..., 0], w[..., 1] = a, b
    y = O.mode_b(x, w, convert_to_circular=True)
    X, Y = x[0, 0, :, 0], x[0, 0, :, 1]
    assert np.allclose(y[0, 0, :, 0], a * (X + 1j * Y))
    assert np.allclose(y[0, 0, :, 1], b * (X - 1j * Y))
    assert not np.allclose(y[0, 0, :, 0], a * X + 1j * b * Y)     # beamform, then polarize


def test_mode_b_detector_switches_the_incoherent_sqrt():
    # mode_b.hh:156 sets enableIncoherentBeamSqrt = detectorEnable: with the
    # detector the incoherent beam is sqrt(P) and Stokes I of it gives back
    # the summed power P_X + P_Y, P_p = sum_a |x_p|^2 |w0_p|^2 (beam 0's
    # phasors, beamformer.cu:68-69); without it the beam holds (P_p, 0).
    x = crand(3, 2, 6, 2)
    w = crand(2, 3, 2, 1, 2)
    P = np.einsum("aftp,afp->ftp", np.abs(x) ** 2, np.abs(w[0, :, :, 0, :]) ** 2)
    det = O.mode_b(x, w, incoherent_beam=True, detector_enable=True)
    assert det.shape == (3, 2, 6, 1)
    assert np.allclose(det[2, :, :, 0], P.sum(axis=-1))
    # coherent beams: |sum_a x w|^2 summed over pols
    coh = np.einsum("aftp,bafp->bftp", x, w[:, :, :, 0, :])
    assert np.allclose(det[:2, :, :, 0], (np.abs(coh) ** 2).sum(axis=-1))
    raw = O.mode_b(x, w, incoherent_beam=True, detector_enable=False)
    assert raw.shape == (3, 2, 6, 2)
    assert np.allclose(raw[2], P) and np.allclose(raw[:2], coh)


def test_mode_h_tone_and_power():
    # mode_h.hh:55-106: Channelizer(rate = T) -> Polarizer -> Detector.  A tone
    # k0 in X of amplitude 1: Stokes I of (L, R) is 2|C|^2 in fine channel
    # (k0 + T/2) % T and 0 elsewhere; without the polarizer it is |C|^2.
    T, k0 = 16, 3
    x = np.zeros((1, 2, T, 2), complex)
    x[0, :, :, 0] = np.exp(2j * np.pi * k0 * np.arange(T) / T)
    for circ, scale in ((True, 2.0), (False, 1.0)):
        y = O.mode_h(x, convert_to_circular=circ)
        assert y.shape == (1, 2 * T, 1, 1)
        exp = np.zeros(2 * T)
        exp[[(k0 + T // 2) % T, T + (k0 + T // 2) % T]] = scale * T * T
        assert np.allclose(y[0, :, 0, 0], exp, atol=1e-9)


def test_calibration_gradient_closed_form():
    x = crand(3, 2, 6, 2)
    w = crand(1, 3, 2, 1, 2)
    g = crand(3, 2, 2)
    grad = O.stokes_i_loss_grad_gains(x, g, w)
    h = 1e-6
    for idx in [(0, 0, 0), (2, 1, 1), (1, 0, 1)]:
        e = np.zeros_like(g)
        e[idx] = h
        dre = (O.beamformed_stokes_i_loss(x, g + e, w) - O.beamformed_stokes_i_loss(x, g - e, w)) / (2 * h)
        dim = (O.beamformed_stokes_i_loss(x, g + 1j * e, w) - O.beamformed_stokes_i_loss(x, g - 1j * e, w)) / (2 * h)
        assert np.isclose(grad[idx].real, dre, rtol=1e-6) and np.isclose(grad[idx].imag, dim, rtol=1e-6)


def test_adjoint_identity():
    # Re <gy, J dw> == Re <adj(gy), dw> for the beamformer's linear map J in w.
    rng = np.random.default_rng(11)
    x = O.ci8_gaussian((5, 3, 16, 2), rng=rng)
    gy = crand(4, 3, 16, 2)
    for _ in range(3):
        dw = crand(4, 5, 3, 1, 2)
        lhs = np.vdot(gy, O.beamformer_ata(x, dw)).real
        rhs = np.vdot(O.beamformer_adjoint(x, gy), dw).real
        assert np.isclose(lhs, rhs, rtol=1e-10)
    # a dropped conjugate on x breaks the identity
    bad = np.einsum("bftp,aftp->bafp", gy, O.ci8_to_complex(x))[:, :, :, None, :]
    dw = crand(4, 5, 3, 1, 2)
    assert not np.isclose(np.vdot(gy, O.beamformer_ata(x, dw)).real, np.vdot(bad, dw).real, rtol=1e-3)


def test_stokes_i_phasor_gradient_finite_differences():
    # PyTorch's gradient (dL/dRe w + j dL/dIm w) is 2 dL/dconj(w).
    rng = np.random.default_rng(12)
    x = O.ci8_gaussian((4, 2, 8, 2), rng=rng)
    w = crand(3, 4, 2, 1, 2).astype(np.complex128)
    lam = np.array([1.0, 0.0, 0.5])
    loss, g = O.stokes_i_loss_and_grad_phasors(x, w, lam)
    ref_loss = sum(-lam[b] * np.sum(np.abs(O.beamformer_ata(x, w)[b]) ** 2) for b in range(3))
    assert np.isclose(loss, ref_loss, rtol=1e-12)
    h = 1e-6
    for idx in [(0, 0, 0, 0, 0), (2, 3, 1, 0, 1), (1, 2, 0, 0, 1)]:
        e = np.zeros_like(w)
        e[idx] = h
        f = lambda v: O.stokes_i_loss_and_grad_phasors(x, v, lam)[0]
        dre = (f(w + e) - f(w - e)) / (2 * h)
        dim = (f(w + 1j * e) - f(w - 1j * e)) / (2 * h)
        assert np.isclose(2 * g[idx].real, dre, rtol=1e-6, atol=1e-3)
        assert np.isclose(2 * g[idx].imag, dim, rtol=1e-6, atol=1e-3)
    # beam 1 has zero weight: zero gradient
    assert np.all(g[1] == 0)


def test_gain_gradient_is_adjoint_chain_rule():
    # d/dconj(g) of L(g * w0) = conj(w0) * dL/dconj(w_eff): ties the gains
    # gradient (pinned by finite differences above) to the phasor adjoint.
    rng = np.random.default_rng(13)
    x = O.ci8_gaussian((4, 2, 8, 2), rng=rng)
    w = crand(1, 4, 2, 1, 2).astype(np.complex128)
    gains = crand(4, 2, 2).astype(np.complex128)
    weff = w * gains[None, :, :, None, :]
    _, gw = O.stokes_i_loss_and_grad_phasors(x, weff, [1.0])
    via_adjoint = 2 * np.conj(w[0, :, :, 0, :]) * gw[0, :, :, 0, :]
    assert np.allclose(via_adjoint, O.stokes_i_loss_grad_gains(x, gains, w), rtol=1e-10)


def test_cpu_port_matches_oracle():
    x8 = O.ci8_gaussian((5, 3, 32, 2), rng=3)
    w = crand(4, 5, 3, 1, 2)
    for rate, pol in ((1, False), (4, True), (2, False)):
        ref = O.beamformer_ata(x8, w)
        ref = O.channelizer(ref, rate)
        if pol:
            ref = O.polarizer(ref)
        got = O.beamform_chain_c64(x8, w, rate, pol)
        assert got.dtype == np.complex64
        assert O.norm_rel_err(got, ref) < 1e-6


def test_mode_h_closed_forms():
    """Mode H (channelizer at rate T, polarizer, detector), pinned without
    restating its composition: a tone lands in its fftshifted bin with power
    C^2 (|X|^2 + |Y|^2); Parseval holds over the channels; circular
    conversion doubles Stokes I (|X + jY|^2 + |X - jY|^2 = 2 (|X|^2 + |Y|^2))."""
    C, k = 64, 5
    t = np.arange(C)
    ax, ay = 0.75 - 0.5j, -0.25 + 1.0j
    x = np.zeros((1, 1, C, 2), dtype=np.complex128)
    x[0, 0, :, 0] = ax * np.exp(2j * np.pi * k * t / C)
    x[0, 0, :, 1] = ay * np.exp(2j * np.pi * k * t / C)
    y = O.mode_h(x)                                         # [1, C, 1, 1] Stokes I
    assert y.shape == (1, C, 1, 1)
    expect = np.zeros(C)
    expect[k + C // 2] = C * C * (abs(ax) ** 2 + abs(ay) ** 2)
    np.testing.assert_allclose(y[0, :, 0, 0], expect, rtol=1e-12, atol=1e-9)

    rng = np.random.default_rng(4)
    z = rng.normal(size=(2, 3, 32, 2)) + 1j * rng.normal(size=(2, 3, 32, 2))
    s = O.mode_h(z)                                         # [2, 3 * 32, 1, 1]
    per = s.reshape(2, 3, 32).sum(axis=2)
    np.testing.assert_allclose(per, 32 * (np.abs(z) ** 2).sum(axis=(2, 3)), rtol=1e-12)
    np.testing.assert_allclose(O.mode_h(z, convert_to_circular=True), 2 * s, rtol=1e-12)
