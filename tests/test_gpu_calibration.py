"""Adjoint beamformer kernel, fused Stokes-I gradient and the autograd functions.

Tolerances: FP32 matrix-core contractions against the complex128 oracle,
normwise relative error <= 1e-5; finite differences in FP64 (step 1e-6)
against the GPU gradient, relative to the gradient's RMS, <= 1e-4.
"""

import math

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

TOL = 1e-5


def dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda")


def host(t):
    import torch
    torch.cuda.synchronize()
    return t.detach().cpu().numpy()


def crand(rng, *shape):
    return (rng.normal(size=shape) + 1j * rng.normal(size=shape)).astype(np.complex64)


@pytest.mark.parametrize("A,F,T,B", [(6, 3, 256, 4), (42, 4, 2048, 4), (28, 2, 300, 9),
                                     (16, 1, 128, 16), (1, 2, 64, 1), (128, 1, 512, 8)])
def test_adjoint_matches_oracle(A, F, T, B):
    from blade_amd import ops
    rng = np.random.default_rng(A * 1000 + T)
    x = O.ci8_gaussian((A, F, T, 2), rng=rng)
    gy = crand(rng, B, F, T, 2)
    got = host(ops.beamformer_adjoint(dev(x), dev(gy)))
    ref = O.beamformer_adjoint(x, gy)
    assert got.shape == ref.shape
    assert O.norm_rel_err(got, ref) < TOL


@pytest.mark.parametrize("A,F,T,B,lam", [(6, 3, 256, 4, [1, 0, 0, 0]), (42, 4, 2048, 4, [1, 0, 0, 0]),
                                         (28, 2, 384, 9, [1, 0.5, 0, 0, 0, 0, 0, 0, 2]),
                                         (20, 2, 100, 1, [1]),
                                         # antenna tiles 1-8, ragged T, several time splits, 3 beam tiles
                                         (4, 1, 64, 1, [1]), (17, 3, 4100, 2, [1, -0.5]),
                                         (64, 2, 1000, 3, [0.2, 1, 3]), (80, 1, 130, 5, [1] * 5),
                                         (100, 1, 320, 20, [(-1) ** b * (b + 1) / 7 for b in range(20)]),
                                         (128, 1, 700, 8, [1, 0, 0, 0, 0, 0, 0, 1]),
                                         # at most 4 beams (FP32: the 4x4-block kernel), all weighted;
                                         # and its fallbacks: A > 64, T not a multiple of 4
                                         (33, 5, 3072, 4, [0.5, -1, 2, 1]), (65, 2, 256, 4, [1, 0.5, 0.25, 2]),
                                         (10, 2, 130, 2, [1, 1])])
@pytest.mark.parametrize("precision", ["f32", "f16x2"])
def test_stokes_i_grad_matches_oracle(A, F, T, B, lam, precision):
    from blade_amd import ops
    rng = np.random.default_rng(A + T)
    x = O.ci8_gaussian((A, F, T, 2), rng=rng)
    w = crand(rng, B, A, F, 1, 2)
    loss, grad = ops.stokes_i_loss_and_grad(dev(x), dev(w), lam, precision=precision)
    ref_loss, ref_grad = O.stokes_i_loss_and_grad_phasors(x, w, lam)
    assert abs(float(host(loss)) - ref_loss) / abs(ref_loss) < TOL
    assert O.norm_rel_err(host(grad), ref_grad) < TOL


@pytest.mark.parametrize("precision", ["f32", "f16x2"])
def test_stokes_i_grad_is_deterministic(precision):
    # Partial sums are combined in a fixed order: repeated calls agree bit for bit.
    from blade_amd import ops
    rng = np.random.default_rng(11)
    x = dev(O.ci8_gaussian((42, 8, 2048, 2), rng=rng))
    w = dev(crand(rng, 4, 42, 8, 1, 2))
    l0, g0 = ops.stokes_i_loss_and_grad(x, w, [1, 0, 0, 0], precision=precision)
    for _ in range(3):
        l1, g1 = ops.stokes_i_loss_and_grad(x, w, [1, 0, 0, 0], precision=precision)
        assert float(host(l1)) == float(host(l0))
        assert np.array_equal(host(g1), host(g0))


def test_adjoint_rejects_bad_input():
    import torch
    from blade_amd import ops
    from blade_amd._lib import BladeError
    x = torch.zeros((129, 1, 64, 2, 2), dtype=torch.int8, device="cuda")
    gy = torch.zeros((1, 1, 64, 2), dtype=torch.complex64, device="cuda")
    with pytest.raises(BladeError, match="128 antennas"):
        ops.beamformer_adjoint(x, gy)
    w = torch.zeros((65, 4, 1, 1, 2), dtype=torch.complex64, device="cuda")
    with pytest.raises(BladeError, match="64 beams"):
        ops.stokes_i_loss_and_grad(x[:4], w, [1.0] * 65)


def test_beamform_autograd_matches_torch_complex128():
    import torch
    from blade_amd.calibration import beamform
    rng = np.random.default_rng(5)
    A, F, T, B = 10, 3, 256, 5
    x = O.ci8_gaussian((A, F, T, 2), rng=rng)
    w0 = crand(rng, B, A, F, 1, 2)
    c = crand(rng, B, F, T, 2)
    # GPU: y through the custom Function, any loss built from torch ops
    w = dev(w0).requires_grad_(True)
    y = beamform(dev(x), w)
    loss = (torch.abs(y) ** 2 * 0.5).sum() + (dev(c).conj() * y).real.sum()
    loss.backward()
    # CPU reference: the same expression in complex128 with torch's own autograd
    xc = torch.from_numpy(O.ci8_to_complex(x))
    wr = torch.from_numpy(w0.astype(np.complex128)).requires_grad_(True)
    yr = torch.einsum("bafp,aftp->bftp", wr[:, :, :, 0, :], xc)
    lr = (torch.abs(yr) ** 2 * 0.5).sum() + (torch.from_numpy(c.astype(np.complex128)).conj() * yr).real.sum()
    lr.backward()
    assert O.norm_rel_err(host(w.grad), wr.grad.numpy()) < TOL


def test_stokes_i_power_finite_differences_fp64():
    import torch
    from blade_amd.calibration import stokes_i_power
    rng = np.random.default_rng(6)
    A, F, T, B = 8, 2, 128, 3
    x = O.ci8_gaussian((A, F, T, 2), rng=rng)
    w0 = crand(rng, B, A, F, 1, 2)
    lam = [1.0, 0.25, 0.0]
    w = dev(w0).requires_grad_(True)
    stokes_i_power(dev(x), w, lam).backward()
    g = host(w.grad)
    wd = w0.astype(np.complex128)
    f = lambda v: O.stokes_i_loss_and_grad_phasors(x, v, lam)[0]
    h = 1e-6
    rms = np.sqrt(np.mean(np.abs(g) ** 2))
    for idx in [(0, 0, 0, 0, 0), (0, 7, 1, 0, 1), (1, 3, 0, 0, 1), (2, 5, 1, 0, 0)]:
        e = np.zeros_like(wd)
        e[idx] = h
        dre = (f(wd + e) - f(wd - e)) / (2 * h)
        dim = (f(wd + 1j * e) - f(wd - 1j * e)) / (2 * h)
        assert abs(g[idx].real - dre) / rms < 1e-4
        assert abs(g[idx].imag - dim) / rms < 1e-4


def test_gain_and_delay_gradients_through_torch():
    import torch
    from blade_amd.calibration import steer_phasors, stokes_i_power
    rng = np.random.default_rng(7)
    A, F, T = 6, 3, 256
    x = O.ci8_gaussian((A, F, T, 2), rng=rng)
    w0 = crand(rng, 1, A, F, 1, 2)
    g0 = crand(rng, A, F, 2)
    freqs = 1.0e9 + np.arange(F) * 0.5e6
    d0 = rng.uniform(-1e-9, 1e-9, size=A)
    # GPU path
    g = dev(g0).requires_grad_(True)
    d = dev(d0).requires_grad_(True)
    stokes_i_power(dev(x), steer_phasors(dev(w0), g, d, dev(freqs)), [1.0]).backward()
    # CPU complex128 reference with torch autograd end to end
    gr = torch.from_numpy(g0.astype(np.complex128)).requires_grad_(True)
    dr = torch.from_numpy(d0).requires_grad_(True)
    weff = steer_phasors(torch.from_numpy(w0.astype(np.complex128)), gr, dr, torch.from_numpy(freqs))
    y = torch.einsum("afp,aftp->ftp", weff[0, :, :, 0, :], torch.from_numpy(O.ci8_to_complex(x)))
    (-(torch.abs(y) ** 2).sum()).backward()
    assert O.norm_rel_err(host(g.grad), gr.grad.numpy()) < 1e-4
    assert abs(float(host(d.grad)[0]) - float(dr.grad[0])) / float(dr.grad.abs().max()) < 1e-4


def test_steer_gains_kernel():
    from blade_amd import ops
    rng = np.random.default_rng(21)
    w0 = crand(rng, 3, 5, 4, 1, 2)
    g0 = crand(rng, 5, 4, 2)
    g0[1, 2, 0] = 0                                      # |g| below the floor: u = g * 1e30 = 0
    got = host(ops.steer_gains(dev(w0), dev(g0)))
    gd = g0.astype(np.complex128)
    u = gd / np.maximum(np.abs(gd), 1e-30)
    ref = w0.astype(np.complex128) * u[None, :, :, None, :]
    assert O.norm_rel_err(got, ref) < 1e-6
    with pytest.raises(ValueError, match="gains"):
        ops.steer_gains(dev(w0), dev(g0[:4]))
    # per-antenna delays, with and without gains: exp(-j 2 pi f tau) from an FP64 phase
    tau = rng.uniform(-100e-9, 100e-9, size=5)
    freqs = 6.5e9 + np.arange(4) * 0.5e6
    rot = np.exp(-2j * np.pi * freqs[None, :] * tau[:, None])                # [A, F]
    got = host(ops.steer_gains(dev(w0), None, dev(tau), dev(freqs)))
    assert O.norm_rel_err(got, w0.astype(np.complex128) * rot[None, :, :, None, None]) < 1e-6
    got = host(ops.steer_gains(dev(w0), dev(g0), dev(tau), dev(freqs)))
    assert O.norm_rel_err(got, ref * rot[None, :, :, None, None]) < 1e-6
    with pytest.raises(ValueError, match="frequencies"):
        ops.steer_gains(dev(w0), None, dev(tau), None)


@pytest.mark.parametrize("B,lam,precision", [(2, [1.0, 0.5], "f32"), (2, [1.0, 0.5], "f16x2"),
                                             (6, [1.0, 0, 0.5, 0, 0, -0.25], "f32")])
def test_gain_gradient_finite_differences_fp64(B, lam, precision):
    # The gain gradient of SteeredStokesIPower against central differences of
    # the complex128 oracle loss in Re g and Im g (FP64, step 1e-6).
    import torch
    from blade_amd.calibration import SteeredStokesIPower
    rng = np.random.default_rng(22 + B)
    A, F, T = 5, 2, 128
    x = O.ci8_gaussian((A, F, T, 2), rng=rng)
    w0 = crand(rng, B, A, F, 1, 2)
    g0 = crand(rng, A, F, 2)
    g = dev(g0).requires_grad_(True)
    SteeredStokesIPower.apply(dev(x), dev(w0), g, lam, precision).backward()
    got = host(g.grad)
    wd, gd = w0.astype(np.complex128), g0.astype(np.complex128)

    def f(gv):
        u = gv / np.abs(gv)
        return O.stokes_i_loss_and_grad_phasors(x, wd * u[None, :, :, None, :], lam)[0]

    h = 1e-6
    rms = np.sqrt(np.mean(np.abs(got) ** 2))
    for idx in [(0, 0, 0), (4, 1, 1), (2, 0, 1), (3, 1, 0)]:
        e = np.zeros_like(gd)
        e[idx] = h
        dre = (f(gd + e) - f(gd - e)) / (2 * h)
        dim = (f(gd + 1j * e) - f(gd - 1j * e)) / (2 * h)
        assert abs(got[idx].real - dre) / rms < 1e-4
        assert abs(got[idx].imag - dim) / rms < 1e-4


@pytest.mark.parametrize("B,precision", [(4, "f32"), (4, "f16x2"), (9, "f32")])
def test_steered_power_matches_torch_chain_rule(B, precision):
    # The fused gain path against torch autograd through steer_phasors and
    # StokesIPower, on the same kernels: same loss, same gain gradient.
    import torch
    from blade_amd.calibration import SteeredStokesIPower, steer_phasors, stokes_i_power
    rng = np.random.default_rng(30 + B)
    A, F, T = 42, 3, 512
    x = dev(O.ci8_gaussian((A, F, T, 2), rng=rng))
    w = dev(crand(rng, B, A, F, 1, 2))
    g0 = crand(rng, A, F, 2)
    lam = [1.0] + [0.3] * (B - 1)
    g1 = dev(g0).requires_grad_(True)
    l1 = SteeredStokesIPower.apply(x, w, g1, lam, precision)
    l1.backward()
    g2 = dev(g0).requires_grad_(True)
    l2 = stokes_i_power(x, steer_phasors(w, g2, None, None), lam, precision)
    l2.backward()
    assert abs(float(host(l1)) - float(host(l2))) <= 1e-6 * abs(float(host(l2)))
    assert O.norm_rel_err(host(g1.grad), host(g2.grad)) < 1e-5


def calibrator(A, F, T, seed, tau_max=100e-9, snr_db=0.0):
    """Point source plane wave x per-antenna gains + complex noise, quantised to CI8."""
    rng = np.random.default_rng(seed)
    freqs = 1.0e9 + np.arange(F) * 0.5e6
    tau = rng.uniform(-tau_max, tau_max, size=A)
    g = (1 + 0.1 * rng.normal(size=(A, F, 2))) * np.exp(1j * rng.uniform(-np.pi, np.pi, size=(A, F, 2)))
    s = (rng.normal(size=(F, T, 2)) + 1j * rng.normal(size=(F, T, 2))) / math.sqrt(2)
    geo = np.exp(-2j * np.pi * freqs[None, :] * tau[:, None])               # [A, F]
    sig = g[:, :, None, :] * geo[:, :, None, None] * s[None]
    noise_std = 10 ** (-snr_db / 20)
    noise = (rng.normal(size=sig.shape) + 1j * rng.normal(size=sig.shape)) * noise_std / math.sqrt(2)
    x = O.complex_to_ci8(12.0 * (sig + noise))
    w = np.conj(geo)[None, :, :, None, None] * np.ones((1, 1, 1, 1, 2))     # steer at the source
    return x, w.astype(np.complex64), g


@pytest.mark.parametrize("precision", ["f32", "f16x2"])
def test_adam_calibration_recovers_gain_phases(precision):
    from blade_amd.calibration import calibrate
    A, F, T = 12, 4, 512
    x, w, g_true = calibrator(A, F, T, seed=8)
    gains, _, losses = calibrate(dev(x), dev(w), steps=200, lr=0.05, precision=precision)
    ge = host(gains)
    u = ge * g_true / np.abs(ge * g_true)
    coherence = np.abs(u.sum(axis=0)) / A                                   # [F, P]
    assert losses[-1] < losses[0] * 2.0                                      # losses are negative
    assert coherence.min() > 0.98


def test_calibration_graph_replays_match_eager_steps():
    # The captured step (steering, fused loss and gradient, chain rule, Adam)
    # replayed 37 times after 3 eager warm-up steps gives every step's loss
    # and the final gains of 40 eager steps; eager torch work runs on the
    # default stream between the replays.
    import torch
    from blade_amd.calibration import calibrate
    A, F, T = 12, 4, 512
    x, w, _ = calibrator(A, F, T, seed=9)
    xd, wd = dev(x), dev(w)
    g_e, _, l_e = calibrate(xd, wd, steps=40, lr=0.05, graph=False)
    g_g, _, l_g = calibrate(xd, wd, steps=40, lr=0.05, graph=True)
    torch.cuda.synchronize()
    np.testing.assert_allclose(l_g, l_e, rtol=1e-5, atol=1e-7)
    assert O.norm_rel_err(host(g_g), host(g_e)) <= 1e-5
    # delays too (their own parameter group)
    freqs = torch.linspace(1.0e9, 1.0015e9, F, dtype=torch.float64, device="cuda")
    _, d_e, l_e = calibrate(xd, wd, steps=12, lr=0.05, fit_delays=True, freqs_hz=freqs, graph=False)
    _, d_g, l_g = calibrate(xd, wd, steps=12, lr=0.05, fit_delays=True, freqs_hz=freqs, graph=True)
    np.testing.assert_allclose(l_g, l_e, rtol=1e-5, atol=1e-7)
    assert O.norm_rel_err(host(d_g), host(d_e)) <= 1e-5


@pytest.mark.parametrize("fit_gains,precision", [(False, "f32"), (True, "f32"), (True, "f16x2")])
def test_fused_delay_adam_matches_torch_adam(fit_gains, precision):
    # Delays (and gains): the library's steering and Adam kernels against
    # torch autograd through steer_phasors and torch.optim.Adam, step by step.
    import torch
    from blade_amd.calibration import calibrate
    A, F, T = 12, 4, 512
    x, w, _ = calibrator(A, F, T, seed=11)
    xd, wd = dev(x), dev(w)
    freqs = torch.linspace(1.0e9, 1.0015e9, F, dtype=torch.float64, device="cuda")
    kw = dict(steps=20, lr=0.05, fit_gains=fit_gains, fit_delays=True, freqs_hz=freqs, delay_lr=2e-10,
              precision=precision)
    g1, d1, l1 = calibrate(xd, wd, fused_adam=True, **kw)
    g2, d2, l2 = calibrate(xd, wd, fused_adam=False, **kw)
    np.testing.assert_allclose(l1, l2, rtol=1e-5, atol=1e-7)
    # The two paths steer in a different order (w r u against (w u) r, the
    # rotation from reduced FP64 turns against torch.polar), so the FP32
    # phasors differ in the last bit and 20 coupled Adam steps carry that
    # into the parameters: 1e-4 normwise, against 1e-5 for the losses.
    assert O.norm_rel_err(host(d1), host(d2)) <= 1e-4
    assert abs(host(d1)).max() > 1e-10                     # the delays moved
    if fit_gains:
        assert O.norm_rel_err(host(g1), host(g2)) <= 1e-4


@pytest.mark.parametrize("precision", ["f32", "f16x2"])
def test_fused_gain_adam_matches_torch_adam(precision):
    # Gains only: the library's gain Adam kernel against torch.optim.Adam
    # (capturable) after autograd through SteeredStokesIPower, step by step.
    from blade_amd.calibration import calibrate
    A, F, T = 12, 4, 512
    x, w, _ = calibrator(A, F, T, seed=10)
    xd, wd = dev(x), dev(w)
    g1, _, l1 = calibrate(xd, wd, steps=30, lr=0.05, precision=precision, fused_adam=True)
    g2, _, l2 = calibrate(xd, wd, steps=30, lr=0.05, precision=precision, fused_adam=False)
    np.testing.assert_allclose(l1, l2, rtol=1e-5, atol=1e-7)
    assert O.norm_rel_err(host(g1), host(g2)) <= 1e-5


def test_gain_adam_rejects_bad_input():
    import torch
    from blade_amd import ops
    from blade_amd._lib import BladeError
    w = torch.zeros((2, 3, 4, 1, 2), dtype=torch.complex64, device="cuda")
    g = torch.ones((3, 4, 2), dtype=torch.complex64, device="cuda")
    m, v = torch.zeros_like(g), torch.zeros_like(g)
    step = torch.zeros((), dtype=torch.float32, device="cuda")
    with pytest.raises(BladeError, match="beta"):
        ops.gain_adam(w, w, g, m, v, step, 0.05, betas=(1.0, 0.999))
    with pytest.raises(ValueError, match="step"):
        ops.gain_adam(w, w, g, m, v, step.double(), 0.05)
    with pytest.raises(ValueError, match="exp_avg"):
        ops.gain_adam(w, w, g, m[:2], v, step, 0.05)


def test_gain_entries_accept_empty_shapes():
    # A zero axis (tensors that may have null data pointers) succeeds without a launch.
    import torch
    from blade_amd import ops
    w = torch.zeros((2, 3, 0, 1, 2), dtype=torch.complex64, device="cuda")
    g = torch.zeros((3, 0, 2), dtype=torch.complex64, device="cuda")
    assert ops.steer_gains(w, g).shape == w.shape
    assert ops.gain_gradient(w, w, g).shape == g.shape
    step = torch.zeros((), dtype=torch.float32, device="cuda")
    ops.gain_adam(w, w, g, torch.zeros_like(g), torch.zeros_like(g), step, 0.05)
    d = torch.zeros((3,), dtype=torch.float64, device="cuda")
    ops.delay_adam(w, w, torch.zeros((0,), dtype=torch.float64, device="cuda"), d, torch.zeros_like(d),
                   torch.zeros_like(d), step, 1e-10)
    torch.cuda.synchronize()
    assert float(step) == 0.0 and float(d.abs().max()) == 0.0
