"""The benchmark configurations other than the timed one, as parity tests.

* small Mode-B (A20 F192 T512 B8, incoherent beam) against the complex128 oracle;
* mixed-precision Mode A (A28 F192 T16384 B8): FP64-phase phasors at up to
  10 GHz and +-200 ns, split-FP16 beamformer, FFT-8192 channelizer and a
  Stokes-I detector integrating 4 spectra over two blocks, with the error of
  every stage against the FP64 oracle (checked on three coarse channels);
* gain calibration (A42 F384 T2048 B4): the hand-written gradient against a
  complex128 CPU gradient and FP64 finite differences, then 200 Adam steps.

Tolerance: normwise relative error <= 1e-5 per stage.
"""

import math

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

TOL = 1e-5


def host(t):
    import torch
    torch.cuda.synchronize()
    return t.detach().cpu().numpy()


def ci8_on_device(shape, seed, sigma=16.0):
    """CI8 N(0, sigma^2) voltages generated on the GPU (seeded)."""
    import torch
    g = torch.Generator(device="cuda").manual_seed(seed)
    v = torch.randn(tuple(shape) + (2,), device="cuda", generator=g) * sigma
    return torch.clamp(torch.round(v), -127, 127).to(torch.int8)


def test_config_small_mode_b():
    import torch
    from blade_amd import ops
    rng = np.random.default_rng(0)
    A, F, T, B = 20, 192, 512, 8
    x = O.ci8_gaussian((A, F, T, 2), rng=rng)
    tau = rng.uniform(-50e-9, 50e-9, size=(B, A))
    w = O.phasor_ata(tau, np.ones((A, F, 1, 2)), 1.0e9, 0.5e6, 0.0, 0)
    ref = O.beamformer_ata(x, w, enable_incoherent_beam=True)
    xd = torch.from_numpy(x).cuda()
    wd = torch.from_numpy(w.astype(np.complex64)).cuda()
    for prec in ("f32", "f16x2"):
        y = host(ops.beamformer(xd, wd, True, False, prec))
        assert y.shape == (B + 1, F, T, 2)
        err = O.norm_rel_err(y, ref)
        print(f"small Mode-B {prec}: rel err {err:.2e}")
        assert err < TOL


def test_config_mode_a_mixed_precision():
    import torch
    from blade_amd import ops
    A, F, T, B, C, N = 28, 192, 16384, 8, 8192, 4
    rng = np.random.default_rng(2)
    tau = rng.uniform(-200e-9, 200e-9, size=(B, A))
    cal = np.ones((A, F, 1, 2), dtype=np.complex128)
    f_obs, df = 9.904e9, 0.5e6            # channels 9.904 .. 9.9995 GHz
    # stage 1: phasors, FP64 phase argument on the GPU
    w = ops.phasor_ata(torch.from_numpy(tau).cuda(), torch.from_numpy(cal).cuda(), f_obs, df, 0.0, 0,
                       out_type="cf32")
    w_ref = O.phasor_ata(tau, cal, f_obs, df, 0.0, 0)
    err_w = float(np.max(np.abs(host(w) - w_ref)))
    # stages 2-3 over two blocks: split-FP16 beams, then FFT-8192 + Stokes I
    out = torch.empty((B, F * C, 1, 1), dtype=torch.float32, device="cuda")
    fsel = [0, 97, 191]
    ys, xs = [], []
    for blk in range(2):
        x = ci8_on_device((A, F, T, 2), seed=100 + blk)
        y = ops.beamformer(x, w, False, False, "f16x2")
        ops.channelizer_detector(y, C, N, 1, out=out, accumulate=blk > 0)
        xs.append(host(x[:, fsel]))
        ys.append(host(y[:, fsel]))
        del x, y
    y_ref = [O.beamformer_ata(xb, w_ref[:, :, fsel]) for xb in xs]
    err_y = max(O.norm_rel_err(ys[b], y_ref[b]) for b in range(2))
    d_ref = O.detector_stepped([O.channelizer(yr, C) for yr in y_ref], N, 1)     # [B, 3*C, 1, 1]
    d = host(out)
    got = np.concatenate([d[:, f * C:(f + 1) * C] for f in fsel], axis=1)
    err_d = O.norm_rel_err(got, d_ref)
    print(f"Mode A: phasor max abs err {err_w:.2e}; beams (split FP16) {err_y:.2e}; "
          f"Stokes-I filterbank {err_d:.2e}")
    assert err_w < 2e-7
    assert err_y < TOL
    assert err_d < TOL


def calibrator(A, F, T, seed, tau_max=100e-9, snr_db=0.0):
    """Point-source plane wave x per-antenna gains + complex noise (SNR per
    antenna), quantised to CI8; phasors steered at the source."""
    rng = np.random.default_rng(seed)
    freqs = 1.0e9 + np.arange(F) * 0.5e6
    tau = rng.uniform(-tau_max, tau_max, size=A)
    g = (1 + 0.1 * rng.normal(size=(A, F, 2))) * np.exp(1j * rng.uniform(-np.pi, np.pi, size=(A, F, 2)))
    s = (rng.normal(size=(F, T, 2)) + 1j * rng.normal(size=(F, T, 2))) / math.sqrt(2)
    geo = np.exp(-2j * np.pi * freqs[None, :] * tau[:, None])
    sig = g[:, :, None, :] * geo[:, :, None, None] * s[None]
    noise_std = 10 ** (-snr_db / 20)
    noise = (rng.normal(size=sig.shape) + 1j * rng.normal(size=sig.shape)) * (noise_std / math.sqrt(2))
    x = O.complex_to_ci8(12.0 * (sig + noise))
    w = np.conj(geo)[None, :, :, None, None] * np.ones((4, 1, 1, 1, 2))
    return x, w.astype(np.complex64), g


def test_config_calibration_gradient_and_adam():
    import torch
    from blade_amd.calibration import calibrate, stokes_i_power
    A, F, T = 42, 384, 2048
    x, w, g_true = calibrator(A, F, T, seed=4)
    xd, wd = torch.from_numpy(x).cuda(), torch.from_numpy(w).cuda()
    rng = np.random.default_rng(5)
    g0 = ((1 + 0.1 * rng.normal(size=(A, F, 2))) *
          np.exp(1j * rng.uniform(-np.pi, np.pi, size=(A, F, 2)))).astype(np.complex64)
    g = torch.from_numpy(g0).cuda().requires_grad_(True)
    loss = stokes_i_power(xd, wd * g[None, :, :, None, :], [1.0, 0.0, 0.0, 0.0])
    loss.backward()
    grad = host(g.grad)
    # complex128 CPU reference gradient (closed form, pinned by finite differences)
    ref = O.stokes_i_loss_grad_gains(x, g0, w)
    err_g = O.norm_rel_err(grad, ref)
    # FP64 finite differences (step 1e-6) on single coordinates, channel-sliced
    rms = np.sqrt(np.mean(np.abs(ref) ** 2))
    h = 1e-6
    fd_err = 0.0
    for (a, f, p) in [(0, 0, 0), (41, 383, 1), (17, 200, 0)]:
        xs, ws, gs = x[:, f:f + 1], w[:, :, f:f + 1], g0[:, f:f + 1].astype(np.complex128)
        e = np.zeros_like(gs)
        e[a, 0, p] = h
        lf = lambda v: O.beamformed_stokes_i_loss(xs, v, ws)
        dre = (lf(gs + e) - lf(gs - e)) / (2 * h)
        dim = (lf(gs + 1j * e) - lf(gs - 1j * e)) / (2 * h)
        fd_err = max(fd_err, abs(grad[a, f, p] - (dre + 1j * dim)) / rms)
    print(f"calibration: gradient vs complex128 {err_g:.2e}; vs FP64 finite differences {fd_err:.2e}")
    assert err_g < TOL
    assert fd_err < 1e-4
    # 200 Adam steps on unit-modulus gains
    gains, _, losses = calibrate(xd, wd, steps=200, lr=0.05)
    ge = host(gains)
    u = ge * g_true / np.abs(ge * g_true)
    coherence = np.abs(u.sum(axis=0)) / A
    print(f"calibration: loss {losses[0]:.3f} -> {losses[-1]:.3f}; coherence min {coherence.min():.4f} "
          f"median {np.median(coherence):.4f}")
    assert losses[-1] < 2.0 * losses[0]
    assert np.median(coherence) > 0.99
