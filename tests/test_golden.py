"""Golden fixtures of the beamforming path (written by tests/golden/make_golden.py).

CPU tests: the oracle still reproduces every fixture, and the fixtures satisfy
properties that do not depend on the oracle (beam linearity, A-sum identity,
Parseval, unit-modulus phasors, exact bit copies).  GPU tests: the HIP path
against the fixtures, bit-exact for Cast/Gather/Permutation and within the
FP32 tolerance (normwise relative error <= 1e-5) for the rest.
"""

import os

import numpy as np
import pytest

import oracle as O

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
TOL = 1e-5


def load(name):
    with np.load(os.path.join(HERE, f"{name}.npz")) as z:
        return {k: z[k] for k in z.files}


# ---------------------------------------------------------------- CPU: oracle


def test_oracle_reproduces_mode_b_small():
    g = load("mode_b_small")
    y = O.beamformer_ata(g["x"], g["w"], enable_incoherent_beam=True)
    assert O.norm_rel_err(g["y"], y) < 1e-7
    # phasors: unit modulus and the f_k = 1 GHz + k * 0.5 MHz closed form
    f = 1.0e9 + np.arange(4) * 0.5e6
    w = np.exp(-2j * np.pi * g["tau"][:, :, None] * f[None, None, :])
    assert np.max(np.abs(g["w"][:, :, :, 0, 0] - w)) < 1e-6
    # incoherent beam: sum_a |x|^2 |w0|^2 with |w0| = 1
    xc = O.ci8_to_complex(g["x"])
    assert np.allclose(g["y"][-1].real, (np.abs(xc) ** 2).sum(0), rtol=1e-6)


def test_oracle_reproduces_chain_rate4():
    g = load("chain_rate4")
    y = O.polarizer(O.channelizer(O.beamformer_ata(g["x"], g["w"]), 4))
    assert O.norm_rel_err(g["y"], y) < 1e-7
    # The chain is linear in x: doubling the voltages doubles the output.
    y2 = O.polarizer(O.channelizer(O.beamformer_ata(2 * O.ci8_to_complex(g["x"]), g["w"]), 4))
    assert O.norm_rel_err(y2, 2 * y) < 1e-12


def test_oracle_reproduces_channelizer():
    g = load("channelizer")
    assert O.norm_rel_err(g["y4"], O.channelizer(g["x4"], 4)) < 1e-7
    assert O.norm_rel_err(g["yT"], O.channelizer(g["xT"], 256)) < 1e-7
    # Parseval per spectrum: sum |Y|^2 = C sum |x|^2
    for x, y, C in ((g["x4"], g["y4"], 4), (g["xT"], g["yT"], 256)):
        ex = (np.abs(x.astype(np.complex128)) ** 2).sum()
        ey = (np.abs(y.astype(np.complex128)) ** 2).sum()
        assert abs(ey / (C * ex) - 1) < 1e-6


def test_oracle_reproduces_polarizer_detector():
    g = load("polarizer_detector")
    x = g["x"].astype(np.complex128)
    assert np.array_equal(g["lr"], O.polarizer(g["x"]).astype(np.complex64))
    assert np.allclose(g["d1"], O.detector(g["x"], 4, 1), rtol=1e-6)
    assert np.allclose(g["d4"], O.detector(g["x"], 2, 4), rtol=1e-6, atol=1e-6)
    # Stokes I is invariant under the (unnormalised) XY -> LR map up to a factor 2.
    lr = g["lr"].astype(np.complex128)
    assert np.allclose((np.abs(lr) ** 2).sum(-1), 2 * (np.abs(x) ** 2).sum(-1), rtol=1e-6)


def test_oracle_reproduces_phasor():
    g = load("phasor")
    w = O.phasor_ata(g["tau"], g["cal"], float(g["f_obs"]), float(g["chan_bw"]),
                     float(g["total_bw"]), int(g["fstart"]))
    assert np.max(np.abs(w - g["w"])) < 1e-12
    assert np.allclose(np.abs(g["w"]), np.abs(g["cal"])[None, :, :, :, :], rtol=1e-12)


def test_bitexact_fixtures():
    g = load("bitexact")
    assert np.array_equal(g["ci8_cf32"].real, g["ci8"][..., 0].astype(np.float32))
    assert np.array_equal(g["ci8_cf32"].imag, g["ci8"][..., 1].astype(np.float32))
    assert np.array_equal(g["gather_t"], np.concatenate([g["g0"], g["g1"], g["g2"]], axis=2))
    assert np.array_equal(g["perm_0132"], g["perm_in"].transpose(0, 1, 3, 2))
    assert np.array_equal(g["perm_0213"], g["perm_in"].transpose(0, 2, 1, 3))


# ---------------------------------------------------------------- GPU: HIP path


def _dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda")


def _host(t):
    import torch
    torch.cuda.synchronize()
    return t.cpu().numpy()


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["f32", "f16x2"])
def test_gpu_mode_b_small(precision):
    from blade_amd import ops
    g = load("mode_b_small")
    y = _host(ops.beamformer(_dev(g["x"]), _dev(g["w"]), True, False, precision))
    assert O.norm_rel_err(y[:-1], g["y"][:-1]) < TOL
    assert O.norm_rel_err(y[-1], g["y"][-1]) < TOL


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["f32", "f16x2"])
def test_gpu_chain_rate4(precision):
    from blade_amd import ops
    g = load("chain_rate4")
    y = _host(ops.beamform_chain(_dev(g["x"]), _dev(g["w"]), False, False, 4, True, precision))
    assert y.shape == g["y"].shape
    assert O.norm_rel_err(y, g["y"]) < TOL


@pytest.mark.gpu
def test_gpu_channelizer():
    from blade_amd import ops
    g = load("channelizer")
    assert O.norm_rel_err(_host(ops.channelizer(_dev(g["x4"]), 4)), g["y4"]) < TOL
    assert O.norm_rel_err(_host(ops.channelizer(_dev(g["xT"]), 256)), g["yT"]) < TOL


@pytest.mark.gpu
def test_gpu_polarizer_detector():
    from blade_amd import ops
    g = load("polarizer_detector")
    assert np.array_equal(_host(ops.polarizer(_dev(g["x"]))), g["lr"])
    assert O.norm_rel_err(_host(ops.detector(_dev(g["x"]), 4, 1)), g["d1"]) < TOL
    assert O.norm_rel_err(_host(ops.detector(_dev(g["x"]), 2, 4)), g["d4"]) < TOL


@pytest.mark.gpu
def test_gpu_phasor():
    from blade_amd import ops
    g = load("phasor")
    kw = dict(observation_frequency_hz=float(g["f_obs"]), channel_bandwidth_hz=float(g["chan_bw"]),
              total_bandwidth_hz=float(g["total_bw"]), frequency_start_index=int(g["fstart"]))
    w64 = _host(ops.phasor_ata(_dev(g["tau"]), _dev(g["cal"]), out_type="cf64", **kw))
    w32 = _host(ops.phasor_ata(_dev(g["tau"]), _dev(g["cal"]), out_type="cf32", **kw))
    scale = np.abs(g["cal"]).max()
    assert np.max(np.abs(w64 - g["w"])) / scale < 1e-10
    assert np.max(np.abs(w32 - g["w"])) / scale < 2e-7


@pytest.mark.gpu
def test_gpu_bitexact():
    from blade_amd import ops
    g = load("bitexact")
    y = _host(ops.cast(_dev(g["ci8"]), "cf32", in_type="ci8"))
    assert np.array_equal(y.view(np.uint8), g["ci8_cf32"].view(np.uint8))
    out = ops.empty(g["gather_t"].shape, "cf32")
    for step, k in enumerate(("g0", "g1", "g2")):
        ops.gather(_dev(g[k]), out, 2, 3, step)
    assert np.array_equal(_host(out).view(np.uint8), g["gather_t"].view(np.uint8))
    for idx, k in (((0, 1, 3, 2), "perm_0132"), ((0, 2, 1, 3), "perm_0213")):
        y = _host(ops.permutation(_dev(g["perm_in"]), idx))
        assert np.array_equal(y.view(np.uint8), g[k].view(np.uint8))
