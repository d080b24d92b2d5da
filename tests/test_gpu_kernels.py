"""GPU parity of every kernel against the numpy oracle (through the C-ABI).

Bars: FP32 paths <= 1e-5 normwise relative error against the complex128
oracle; split-FP16 <= 1e-5 against the same FP64 oracle; Cast, Gather,
Permutation and Polarizer bit-exact.
"""

import numpy as np
import pytest
import torch

import oracle as O
from blade_amd import ops
from blade_amd._lib import BladeError

pytestmark = pytest.mark.gpu

TOL = 1e-5


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda")


def host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


def phasors_for(B, A, F, rng, tau_max=50e-9, f0=1.0e9, df=0.5e6, P=2):
    tau = rng.uniform(-tau_max, tau_max, size=(B, A))
    cal = np.ones((A, F, 1, P), dtype=np.complex128)
    return O.phasor_ata(tau, cal, f0, df, 0.0, 0)


def crand(rng, *shape):
    return (rng.normal(size=shape) + 1j * rng.normal(size=shape)).astype(np.complex64)


# ------------------------------------------------------------------ beamformer

BF_CASES = [
    # A,  F,   T,   B, inc, sqrt
    (20, 192, 512, 8, True, True),     # config 0 shape
    (2, 192, 512, 1, True, True),      # tests/modules/beamformer/ata.py
    (28, 8, 8192, 16, False, False),   # config 1 per-channel shape
    (28, 4, 16384, 8, True, False),
    (42, 6, 2048, 4, True, True),      # config 4 shape
    (1, 3, 64, 1, True, False),
    (5, 3, 100, 9, True, True),        # B > 8 -> 32-row tile with padding
    (27, 2, 1024, 33, True, False),    # three beam tiles
    (14, 2, 250, 3, False, False),     # T % 4 != 0
    (64, 2, 256, 64, True, True),
    (128, 2, 512, 16, True, True),     # the incoherent beam's antenna limit
    (160, 1, 256, 16, False, False),   # large phasor stages in LDS
    (200, 1, 256, 8, False, False),
]


@pytest.mark.parametrize("A,F,T,B,inc,sq", BF_CASES)
@pytest.mark.parametrize("in_type", ["cf32", "ci8"])
def test_beamformer_fp32(A, F, T, B, inc, sq, in_type):
    rng = np.random.default_rng(A * 1000 + B)
    x8 = O.ci8_gaussian((A, F, T, 2), rng=rng)
    w = phasors_for(B, A, F, rng)
    ref = O.beamformer_ata(x8, w, inc, sq)
    xin = dev(x8) if in_type == "ci8" else dev(O.ci8_to_complex(x8).astype(np.complex64))
    y = host(ops.beamformer(xin, dev(w.astype(np.complex64)), inc, sq, precision="f32"))
    err = O.norm_rel_err(y[:B], ref[:B])
    assert err <= TOL, err
    if inc:
        assert O.norm_rel_err(y[B], ref[B]) <= TOL


@pytest.mark.parametrize("A,F,T,B,inc,sq", BF_CASES)
def test_beamformer_split_f16(A, F, T, B, inc, sq):
    rng = np.random.default_rng(7 + A * 1000 + B)
    x8 = O.ci8_gaussian((A, F, T, 2), rng=rng)
    w = phasors_for(B, A, F, rng, tau_max=200e-9, f0=10e9)
    ref = O.beamformer_ata(x8, w.astype(np.complex64), inc, sq)   # FP64 oracle of the CF32 phasors
    y = host(ops.beamformer(dev(x8), dev(w.astype(np.complex64)), inc, sq, precision="f16x2"))
    err = O.norm_rel_err(y[:B], ref[:B])
    assert err <= TOL, err
    if inc:
        assert O.norm_rel_err(y[B], ref[B]) <= TOL


def test_beamformer_random_phasors_reference_test():
    # tests/modules/beamformer/ata.py: uniform random CF32 input and phasors.
    rng = np.random.default_rng(5)
    x = (rng.random((2, 192, 512, 2)) + 1j * rng.random((2, 192, 512, 2))).astype(np.complex64)
    w = (rng.random((1, 2, 192, 1, 2)) + 1j * rng.random((1, 2, 192, 1, 2))).astype(np.complex64)
    y = host(ops.beamformer(dev(x), dev(w), True, True))
    ref = O.beamformer_ata(x, w, True, True)
    assert O.norm_rel_err(y, ref) <= TOL


@pytest.mark.parametrize("rate", [1, 2, 4])
@pytest.mark.parametrize("polarize", [False, True])
@pytest.mark.parametrize("precision", ["f32", "f16x2"])
def test_beamform_chain(rate, polarize, precision):
    rng = np.random.default_rng(rate * 10 + polarize)
    A, F, T, B = 28, 6, 2048, 16
    x8 = O.ci8_gaussian((A, F, T, 2), rng=rng)
    w = phasors_for(B, A, F, rng).astype(np.complex64)
    ref = O.beamformer_ata(x8, w, True, False)
    ref = O.channelizer(ref, rate)
    if polarize:
        ref = O.polarizer(ref)
    y = host(ops.beamform_chain(dev(x8), dev(w), True, False, rate, polarize, precision))
    assert y.shape == ref.shape
    assert O.norm_rel_err(y, ref) <= TOL


@pytest.mark.parametrize("precision", ["f32", "f16x2"])
@pytest.mark.parametrize("rate,polarize", [(4, True), (1, False)])
def test_beamform_grid_shape_does_not_change_results(monkeypatch, precision, rate, polarize):
    # Chunks per workgroup (BLADE_BF_CHUNKS_PER_WG, else chosen from the
    # kernel's occupancy) only regroup the same per-chunk work: every forced
    # value, dividing the 32 chunks per channel or not, gives the automatic
    # choice's output bit for bit, and that output matches the oracle.
    rng = np.random.default_rng(77)
    A, F, T, B = 28, 4, 8192, 16
    x8 = O.ci8_gaussian((A, F, T, 2), rng=rng)
    w = phasors_for(B, A, F, rng).astype(np.complex64)
    xd, wd = dev(x8), dev(w)
    monkeypatch.delenv("BLADE_BF_CHUNKS_PER_WG", raising=False)
    auto = host(ops.beamform_chain(xd, wd, False, False, rate, polarize, precision))
    ref = O.channelizer(O.beamformer_ata(x8, w, False, False), rate)
    if polarize:
        ref = O.polarizer(ref)
    assert O.norm_rel_err(auto, ref) <= TOL
    for cpw in (1, 3, 5, 8, 32, 100):
        monkeypatch.setenv("BLADE_BF_CHUNKS_PER_WG", str(cpw))
        y = host(ops.beamform_chain(xd, wd, False, False, rate, polarize, precision))
        assert np.array_equal(y.view(np.uint32), auto.view(np.uint32)), cpw


def test_beamform_chain_production_shape():
    # The Mode-B variant the benchmark times (no incoherent beam, rate 4,
    # XY -> LR) at its full shape A28 F192 T8192 B16, both precisions: six
    # channels spread over the band (and so over the grid's channel axis)
    # against the complex128 oracle, and split FP16 against FP32 on the
    # whole block.
    rng = np.random.default_rng(2024)
    A, F, T, B = 28, 192, 8192, 16
    x8 = O.ci8_gaussian((A, F, T, 2), rng=rng)
    w = phasors_for(B, A, F, rng).astype(np.complex64)
    xd, wd = dev(x8), dev(w)
    ys = {}
    for precision in ("f32", "f16x2"):
        y = host(ops.beamform_chain(xd, wd, False, False, 4, True, precision))
        assert y.shape == (B, 4 * F, T // 4, 2)
        for f in (0, 1, 63, 96, 130, 191):
            ref = O.polarizer(O.channelizer(O.beamformer_ata(x8[:, f:f + 1], w[:, :, f:f + 1]), 4))
            err = O.norm_rel_err(y[:, 4 * f:4 * f + 4], ref)
            assert err <= TOL, (precision, f, err)
        ys[precision] = y
    assert O.norm_rel_err(ys["f16x2"], ys["f32"]) <= TOL


@pytest.mark.parametrize("precision", ["f32", "f16x2"])
def test_beamform_chain_caller_workspace(precision):
    # blade_beamform_chain_ws with a caller-owned workspace gives the same bits
    # as the library-owned one, on two calls in a row (the second smaller).
    import ctypes as ct
    from blade_amd import _lib
    rng = np.random.default_rng(31)
    for (A, F, T, B) in ((28, 6, 2048, 16), (20, 3, 1024, 8)):
        x8 = dev(O.ci8_gaussian((A, F, T, 2), rng=rng))
        w = dev(phasors_for(B, A, F, rng).astype(np.complex64))
        ref = host(ops.beamform_chain(x8, w, True, False, 4, True, precision))
        prec = {"f32": 0, "f16x2": 1}[precision]
        lib = _lib.load()
        nbytes = lib.blade_beamformer_workspace_size(prec, A, F, B)
        ws = torch.empty(nbytes, dtype=torch.uint8, device="cuda")
        out = torch.empty_like(torch.from_numpy(ref)).to("cuda")
        _lib.call("blade_beamform_chain_ws", _lib.CI8, prec, ct.c_void_p(x8.data_ptr()),
                  ct.c_void_p(w.data_ptr()), ct.c_void_p(out.data_ptr()), A, F, T, 2, B, 1, 0, 4, 1,
                  ct.c_void_p(ws.data_ptr()), nbytes, ops.stream_handle())
        assert np.array_equal(host(out), ref)
        with pytest.raises(BladeError, match="workspace"):
            _lib.call("blade_beamform_chain_ws", _lib.CI8, prec, ct.c_void_p(x8.data_ptr()),
                      ct.c_void_p(w.data_ptr()), ct.c_void_p(out.data_ptr()), A, F, T, 2, B, 1, 0, 4, 1,
                      ct.c_void_p(ws.data_ptr()), nbytes - 1, ops.stream_handle())


def test_zero_size_calls_are_no_ops():
    # Empty inputs return success without touching the (null) buffers.
    import ctypes as ct
    from blade_amd import _lib
    null = ct.c_void_p(None)
    s = ops.stream_handle()
    cf32 = _lib.TYPE_CODES["cf32"]
    _lib.call("blade_cast", _lib.TYPE_CODES["ci8"], cf32, null, null, 0, s)
    _lib.call("blade_cast", cf32, cf32, null, null, 0, s)
    _lib.call("blade_polarizer", cf32, null, null, 0, s)
    _lib.call("blade_gather", null, null, 0, 16, 4, 1, s)
    _lib.call("blade_permutation", null, null, 8, (ct.c_uint64 * 4)(2, 0, 3, 2),
              (ct.c_uint64 * 4)(0, 2, 3, 1), s)
    _lib.call("blade_detector", null, null, 2, 3, 0, 4, 1, 0, s)
    _lib.call("blade_channelizer", null, null, 2, 3, 0, 2, 4, s)
    # T = 0 leaves no spectra per block: the integration check must not
    # divide by that count.
    _lib.call("blade_channelizer_detector", null, null, 2, 3, 0, 8, 4, 1, 0, 0, s)
    torch.cuda.synchronize()


def test_beamformer_rejects_too_many_antennas_for_lds():
    # 16 beams x 400 antennas of packed phasors exceed the LDS stage.
    x = torch.zeros((400, 1, 256, 2, 2), dtype=torch.int8, device="cuda")
    w = torch.zeros((16, 400, 1, 1, 2), dtype=torch.complex64, device="cuda")
    with pytest.raises(BladeError, match="too many antennas"):
        ops.beamformer(x, w, False, False, "f32")


def test_beamformer_rejects_bad_shapes():
    x = torch.zeros((2, 2, 8, 2), dtype=torch.complex64, device="cuda")
    w = torch.zeros((1, 2, 2, 1, 2), dtype=torch.complex64, device="cuda")
    with pytest.raises(BladeError, match="CI8"):
        ops.beamformer(x, w, precision="f16x2")
    with pytest.raises(BladeError, match="rate"):
        ops.beamform_chain(x, w, rate=8)


# ----------------------------------------------------------------- channelizer

CH_CASES = [
    (5, 100, 200, 1, 200),     # tests/modules/channelizer/meson.build "weird"
    (8, 128, 1024, 2, 1024),   # "hires"
    (8, 128, 1024, 1, 1024),   # "hires-single-pol"
    (4, 16, 8192, 2, 4),       # config 1 rate
    (2, 4, 16384, 2, 8192),    # config 2 rate
    (3, 5, 96, 2, 2),
    (3, 5, 96, 2, 6),
    (3, 5, 96, 2, 12),
    (3, 5, 96, 2, 16),
    (2, 3, 192, 2, 32),
    (2, 3, 384, 2, 96),
    (2, 3, 840, 1, 420),
    (1, 2, 44, 2, 22),         # prime factor 11 -> direct DFT
    (1, 2, 16384, 1, 16384),
]


@pytest.mark.parametrize("A,F,T,P,C", CH_CASES)
def test_channelizer(A, F, T, P, C):
    rng = np.random.default_rng(C)
    x = crand(rng, A, F, T, P)
    y = host(ops.channelizer(dev(x), C))
    ref = O.channelizer(x, C)
    assert y.shape == ref.shape
    assert O.norm_rel_err(y, ref) <= TOL


POW2_RATES = [256, 512, 1024, 2048, 4096, 8192]


@pytest.mark.parametrize("C", POW2_RATES)
def test_channelizer_power_of_two_rates(C):
    # Every radix-16 plan (16^k times 1, 2, 4 or 8); 42 spectra, two per
    # (antenna, channel) so the spectra interleave in the output, and a
    # count that does not fill the last workgroup.
    rng = np.random.default_rng(C + 1)
    x = crand(rng, 3, 7, 2 * C, 2)
    y = host(ops.channelizer(dev(x), C))
    assert O.norm_rel_err(y, O.channelizer(x, C)) <= TOL


@pytest.mark.parametrize("C", [256, 1024, 2048, 4096])
def test_channelizer_many_spectra_per_channel(C):
    # T = 16384: 16384 / C spectra per (antenna, channel), a multiple of the
    # spectra a workgroup holds, which stages them in LDS before the store.
    rng = np.random.default_rng(C + 2)
    x = crand(rng, 2, 3, 16384, 2)
    y = host(ops.channelizer(dev(x), C))
    assert O.norm_rel_err(y, O.channelizer(x, C)) <= TOL


@pytest.mark.parametrize("C", POW2_RATES)
@pytest.mark.parametrize("npol", [1, 4])
def test_channelizer_detector_power_of_two_rates(C, npol):
    rng = np.random.default_rng(C + npol)
    x = crand(rng, 2, 3, 4 * C, 2)
    y = host(ops.channelizer_detector(dev(x), C, 2, npol))
    ref = O.detector(O.channelizer(x, C), 2, npol)
    assert y.shape == ref.shape
    assert O.norm_rel_err(y, ref) <= TOL


def test_channelizer_in_place_rate_equals_t_large():
    rng = np.random.default_rng(4)
    x = crand(rng, 2, 5, 8192, 2)
    t = dev(x)
    ops.channelizer(t, 8192, out=t)
    assert O.norm_rel_err(host(t).reshape(2, 5 * 8192, 1, 2), O.channelizer(x, 8192)) <= TOL


@pytest.mark.parametrize("A,F,T,P,C", [(1, 2, 32768, 2, 32768), (2, 1, 65536, 2, 65536),
                                       (1, 2, 65536, 1, 65536), (1, 1, 98304, 2, 49152),
                                       (1, 2, 2 * 20000, 2, 20000)])
def test_channelizer_large_rates(A, F, T, P, C):
    # Rates past one workgroup's LDS run as two passes (four-step FFT): the
    # reference's cuFFT takes any C == T (src/modules/channelizer/base.cc:58-93),
    # e.g. 65536 in its Mode-BH benchmark (benchmarks/bundles/ata/mode-bh/mode_bh.hh).
    rng = np.random.default_rng(C + P)
    x = crand(rng, A, F, T, P)
    y = host(ops.channelizer(dev(x), C))
    assert O.norm_rel_err(y, O.channelizer(x, C)) <= TOL
    if T == C:   # in place
        t = dev(x)
        ops.channelizer(t, C, out=t)
        assert O.norm_rel_err(host(t).reshape(A, F * C, 1, P), O.channelizer(x, C)) <= TOL


def test_channelizer_in_place_rate_equals_t():
    rng = np.random.default_rng(3)
    x = crand(rng, 2, 3, 1024, 2)
    t = dev(x)
    ops.channelizer(t, 1024, out=t)
    assert O.norm_rel_err(host(t).reshape(2, 3 * 1024, 1, 2), O.channelizer(x, 1024)) <= TOL


@pytest.mark.parametrize("A,F,T,C,N,npol", [(2, 3, 16384, 8192, 2, 1),     # config 2 shape per block
                                            (2, 3, 16384, 8192, 1, 4),
                                            (1, 4, 4096, 1024, 2, 1),
                                            (2, 2, 2400, 200, 3, 4)])
@pytest.mark.parametrize("polarize", [False, True])
def test_channelizer_detector(A, F, T, C, N, npol, polarize):
    rng = np.random.default_rng(C + N)
    x = crand(rng, A, F, T, 2)
    y = host(ops.channelizer_detector(dev(x), C, N, npol, polarize=polarize))
    v = O.channelizer(x, C)
    ref = O.detector(O.polarizer(v) if polarize else v, N, npol)
    assert y.shape == ref.shape
    assert O.norm_rel_err(y, ref) <= TOL


@pytest.mark.parametrize("C", [64, 512, 4096])
def test_channelizer_detector_polarize_stepped(C):
    # XY -> LR before full-Stokes detection, integrated over two blocks
    # (the three kernel families: radix-8 passes, radix 16, large C).
    rng = np.random.default_rng(C)
    blocks = [crand(rng, 2, 3, 2 * C, 2) for _ in range(2)]
    out = torch.empty((2, 3 * C, 1, 4), dtype=torch.float32, device="cuda")
    for k, b in enumerate(blocks):
        ops.channelizer_detector(dev(b), C, 4, 4, out=out, accumulate=k > 0, polarize=True)
    ref = O.detector_stepped([O.polarizer(O.channelizer(b, C)) for b in blocks], 4, 4)
    assert O.norm_rel_err(host(out), ref) <= TOL


def test_channelizer_detector_stepped():
    # 4 spectra integrated over two blocks of 2 (config 2's Stokes-I).
    rng = np.random.default_rng(77)
    blocks = [crand(rng, 2, 3, 4096, 2) for _ in range(2)]
    out = torch.empty((2, 3 * 2048, 1, 1), dtype=torch.float32, device="cuda")
    ops.channelizer_detector(dev(blocks[0]), 2048, 4, 1, out=out, accumulate=False)
    ops.channelizer_detector(dev(blocks[1]), 2048, 4, 1, out=out, accumulate=True)
    ref = O.detector_stepped([O.channelizer(b, 2048) for b in blocks], 4, 1)
    assert O.norm_rel_err(host(out), ref) <= TOL


def test_channelizer_rate_one_is_copy():
    rng = np.random.default_rng(4)
    x = crand(rng, 2, 3, 16, 2)
    assert np.array_equal(host(ops.channelizer(dev(x), 1)), x)


# ------------------------------------------------------------------- polarizer

def test_polarizer_cf32_bit_exact():
    rng = np.random.default_rng(11)
    x = crand(rng, 2, 192, 8750, 2)          # tests/modules/polarizer/polarizer.py shape
    y = host(ops.polarizer(dev(x)))
    X, Y = x[..., 0], x[..., 1]
    exp = np.empty_like(x)
    exp[..., 0] = (X.real - Y.imag) + 1j * (X.imag + Y.real)
    exp[..., 1] = (X.real + Y.imag) + 1j * (X.imag - Y.real)
    assert np.array_equal(y, exp)


def test_polarizer_cf16_bit_exact():
    rng = np.random.default_rng(12)
    h = rng.normal(size=(2, 3, 101, 2, 2)).astype(np.float16)   # odd pair count -> tail path
    y = host(ops.polarizer(dev(h)))
    d = h.astype(np.float64)
    xr, xi, yr, yi = d[..., 0, 0], d[..., 0, 1], d[..., 1, 0], d[..., 1, 1]
    exp = np.stack([np.stack([xr - yi, xi + yr], -1), np.stack([xr + yi, xi - yr], -1)], -2)
    assert np.array_equal(y, exp.astype(np.float16))


# -------------------------------------------------------------------- detector

@pytest.mark.parametrize("intg", [1, 4, 10, 16, 64, 128])
@pytest.mark.parametrize("npol", [1, 4])
def test_detector(intg, npol):
    rng = np.random.default_rng(intg + npol)
    T = 8750 if 8750 % intg == 0 else 8192
    x = crand(rng, 2, 24, T, 2)
    y = host(ops.detector(dev(x), intg, npol))
    ref = O.detector(x, intg, npol)
    assert O.norm_rel_err(y, ref) <= TOL


@pytest.mark.parametrize("npol", [1, 4])
def test_detector_many_outputs_long_integration(npol):
    # 65536 outputs at INTG 64: 16 lanes per output, each summing 4 samples.
    rng = np.random.default_rng(64 + npol)
    x = crand(rng, 4, 128, 8192, 2)
    y = host(ops.detector(dev(x), 64, npol))
    assert O.norm_rel_err(y, O.detector(x, 64, npol)) <= TOL


def test_detector_stepped_accumulate():
    rng = np.random.default_rng(21)
    a, b = crand(rng, 2, 8, 64, 2), crand(rng, 2, 8, 64, 2)
    out = torch.zeros((2, 8, 1, 1), dtype=torch.float32, device="cuda")
    ops.detector(dev(a), 64, 1, out=out, accumulate=False)
    ops.detector(dev(b), 64, 1, out=out, accumulate=True)
    ref = O.detector_stepped([a, b], 128, 1)
    assert O.norm_rel_err(host(out), ref) <= TOL


# ------------------------------------------------------------------------ cast

CAST_PAIRS = [("i8", "i8"), ("i8", "f32"), ("f16", "f16"), ("f16", "f32"), ("f32", "i32"),
              ("f32", "f32"), ("f32", "f16"), ("ci8", "ci8"), ("ci8", "cf32"), ("ci8", "cf16"),
              ("cf32", "cf16"), ("cf32", "cf32"), ("cf16", "cf32"), ("cf16", "cf16"),
              ("f32", "f64"), ("cf32", "cf64"), ("f64", "f32"), ("i32", "f32")]


def make_input(t, rng, n):
    if t in ("i8", "ci8"):
        v = rng.integers(-128, 128, size=(n, 2) if t == "ci8" else n).astype(np.int8)
    elif t in ("f16", "cf16"):
        v = (rng.normal(size=(n, 2) if t == "cf16" else n) * 300).astype(np.float16)
    elif t == "f32":
        v = (rng.normal(size=n) * 1e3).astype(np.float32)
    elif t == "f64":
        v = rng.normal(size=n) * 1e3
    elif t == "i32":
        v = rng.integers(-2**24, 2**24, size=n).astype(np.int32)
    elif t == "cf32":
        v = crand(rng, n) * 1e3
    return v


@pytest.mark.parametrize("it,ot", CAST_PAIRS)
@pytest.mark.parametrize("n", [1, 7, 8, 1000, 100003])
def test_cast_bit_exact(it, ot, n):
    rng = np.random.default_rng(n)
    v = make_input(it, rng, n)
    y = host(ops.cast(dev(v), ot, in_type=it))
    ref = O.cast(v, it, ot)
    if ot in ("ci8", "cf16") and ref.ndim == 1:
        ref = ref.reshape(-1, 2)
    assert y.dtype == ref.dtype
    assert np.array_equal(y.view(np.uint8), np.ascontiguousarray(ref).view(np.uint8))


def test_cast_ci8_all_values():
    v = np.arange(-128, 128, dtype=np.int8)
    ci8 = np.stack(np.meshgrid(v, v, indexing="ij"), -1).reshape(-1, 2).copy()
    y = host(ops.cast(dev(ci8), "cf32", in_type="ci8"))
    assert np.array_equal(y.real, ci8[:, 0].astype(np.float32))
    assert np.array_equal(y.imag, ci8[:, 1].astype(np.float32))


# ---------------------------------------------------------------------- gather

GATHER_CASES = [(2, 192, 8192, 2, 2, 16), (3, 7, 3, 1, 0, 5), (1, 2, 5, 6, 1, 6),
                (11, 1, 13, 9, 2, 7), (9, 2, 1, 3, 3, 8), (9, 2, 5555, 3, 2, 8)]


@pytest.mark.parametrize("A,F,T,P,axis,mult", GATHER_CASES)
@pytest.mark.parametrize("dtype", ["cf32", "ci8"])
def test_gather_bit_exact(A, F, T, P, axis, mult, dtype):
    # tests/modules/gather/meson.build cases; blocks differ per step so a
    # misplaced block is caught.
    rng = np.random.default_rng(A + F + T)
    shape = (A, F, T, P)
    blocks = []
    out_shape = list(shape)
    out_shape[axis] *= mult
    out = ops.empty(out_shape, dtype)
    for step in range(mult):
        if dtype == "cf32":
            blk = crand(rng, *shape)
        else:
            blk = rng.integers(-128, 128, size=shape + (2,)).astype(np.int8)
        blocks.append(blk)
        ops.gather(dev(blk), out, axis, mult, step, type_name=dtype)
    ref = O.gather(blocks, axis)
    assert np.array_equal(host(out), ref)


def test_gather_large_reference_case():
    # (27, 131072, 3, 2) along axis 2 x 16 (tests/modules/gather/meson.build):
    # 2.7 GB of output; compared on the device against torch.cat.
    shape = (27, 131072, 3, 2)
    out = torch.empty((27, 131072, 48, 2), dtype=torch.complex64, device="cuda")
    blocks = []
    g = torch.Generator(device="cuda").manual_seed(0)
    for step in range(16):
        blk = torch.randn(shape, dtype=torch.complex64, device="cuda", generator=g)
        blocks.append(blk)
        ops.gather(blk, out, 2, 16, step)
    torch.cuda.synchronize()
    assert torch.equal(out, torch.cat(blocks, dim=2))


# ----------------------------------------------------------------- permutation

@pytest.mark.parametrize("idx", [[0, 1, 2, 3], [0, 2, 1, 3], [3, 2, 1, 0], [2, 0, 3, 1],
                                 [0, 1, 3, 2], [1, 0, 2, 3]])
@pytest.mark.parametrize("dtype", ["cf32", "ci8", "cf16"])
def test_permutation_bit_exact(idx, dtype):
    rng = np.random.default_rng(sum(idx))
    shape = (2, 192, 8192, 2) if dtype == "cf32" else (3, 17, 129, 2)
    if dtype == "cf32":
        x = crand(rng, *shape)
        ref = O.permutation(x, idx)
    elif dtype == "ci8":
        x = rng.integers(-128, 128, size=shape + (2,)).astype(np.int8)
        ref = np.ascontiguousarray(np.transpose(x, idx + [4]))
    else:
        x = rng.normal(size=shape + (2,)).astype(np.float16)
        ref = np.ascontiguousarray(np.transpose(x, idx + [4]))
    y = host(ops.permutation(dev(x), idx, type_name=dtype))
    assert np.array_equal(y, ref)


@pytest.mark.parametrize("shape,idx,dtype", [
    ((2, 192, 8192, 2), [0, 2, 3, 1], "cf32"),     # AFTP -> ATPF at the beams' shape
    ((2, 100, 130, 2), [0, 2, 3, 1], "cf32"),      # ragged 64-element tiles
    ((2, 100, 130, 2), [3, 2, 1, 0], "cf32"),
    ((3, 70, 97, 1), [2, 0, 3, 1], "cf32"),
    ((1, 70, 300, 2), [0, 2, 3, 1], "ci8"),        # 2-byte elements
    ((2, 65, 66, 2), [1, 3, 2, 0], "cf16")])
def test_permutation_wide_tiles_bit_exact(shape, idx, dtype):
    rng = np.random.default_rng(len(shape) + sum(idx) + shape[1])
    if dtype == "cf32":
        x = crand(rng, *shape)
        ref = O.permutation(x, idx)
    elif dtype == "ci8":
        x = rng.integers(-128, 128, size=shape + (2,)).astype(np.int8)
        ref = np.ascontiguousarray(np.transpose(x, idx + [4]))
    else:
        x = rng.normal(size=shape + (2,)).astype(np.float16)
        ref = np.ascontiguousarray(np.transpose(x, idx + [4]))
    y = host(ops.permutation(dev(x), idx, type_name=dtype))
    assert np.array_equal(y, ref)


# ---------------------------------------------------------------------- phasor

@pytest.mark.parametrize("B,A,F,tau_max,f_obs", [(2, 20, 192, 5e-6, 6500.125e6),
                                                 (8, 28, 192, 200e-9, 10e9),
                                                 (16, 28, 192, 50e-9, 1.0e9)])
def test_phasor(B, A, F, tau_max, f_obs):
    rng = np.random.default_rng(B)
    tau = rng.uniform(-tau_max, tau_max, size=(B, A))
    cal = rng.random((A, F, 1, 2)) + 1j * rng.random((A, F, 1, 2))
    kw = dict(observation_frequency_hz=f_obs, channel_bandwidth_hz=0.5e6,
              total_bandwidth_hz=1.024e9, frequency_start_index=352)
    ref = O.phasor_ata(tau, cal, **kw)
    w32 = host(ops.phasor_ata(dev(tau), dev(cal), out_type="cf32", **kw))
    w64 = host(ops.phasor_ata(dev(tau), dev(cal), out_type="cf64", **kw))
    scale = np.abs(cal).max()
    assert np.max(np.abs(w32 - ref)) / scale < 2e-7
    # the oracle's own FP64 phase carries ~1e-16 * |phase| of rounding
    assert np.max(np.abs(w64 - ref)) / scale < 1e-10
