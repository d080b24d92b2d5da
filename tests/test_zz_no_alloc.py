"""blade_beamform_chain_noalloc (static device scratch, no allocation)
against blade_beamform_chain through the C-ABI: the same kernels, so the
same bits, on two calls in a row, the second with a smaller input, for both
precisions and with and without the fused epilogue."""

import ctypes as ct

import numpy as np
import pytest
import torch

import oracle as O
from blade_amd import _lib

pytestmark = pytest.mark.gpu


def _ptr(t):
    return ct.c_void_p(t.data_ptr())


def _run(name, x, w, out, A, F, T, B, inc, sq, rate, pol, prec):
    _lib.call(name, _lib.CI8, prec, _ptr(x), _ptr(w), _ptr(out), A, F, T, 2, B, inc, sq, rate, pol,
              ct.c_void_p(torch.cuda.current_stream().cuda_stream))


@pytest.mark.parametrize("prec", [_lib.PREC_F32, _lib.PREC_F16X2])
def test_noalloc_matches_chain_two_calls(prec):
    rng = np.random.default_rng(17 + prec)
    cases = [(28, 24, 4096, 16, 0, 0, 4, 1), (20, 6, 1024, 8, 1, 1, 1, 0)]   # second call smaller
    for (A, F, T, B, inc, sq, rate, pol) in cases:
        x8 = O.ci8_gaussian((A, F, T, 2), rng=rng)
        tau = rng.uniform(-50e-9, 50e-9, size=(B, A))
        w = O.phasor_ata(tau, np.ones((A, F, 1, 2)), 1.0e9, 0.5e6, 0.0, 0).astype(np.complex64)
        x = torch.from_numpy(x8).cuda()
        wd = torch.from_numpy(w).cuda()
        shape = (B + inc, F * rate, T // rate, 2)
        ref = torch.full(shape, float("nan"), dtype=torch.complex64, device="cuda")
        got = torch.full(shape, float("nan"), dtype=torch.complex64, device="cuda")
        print(f"noalloc case A{A} F{F} T{T} B{B} inc{inc} rate{rate} pol{pol} prec{prec}", flush=True)
        _run("blade_beamform_chain", x, wd, ref, A, F, T, B, inc, sq, rate, pol, prec)
        _run("blade_beamform_chain_noalloc", x, wd, got, A, F, T, B, inc, sq, rate, pol, prec)
        torch.cuda.synchronize()
        assert torch.equal(got.view(torch.float32), ref.view(torch.float32))
        assert torch.isfinite(got.view(torch.float32)).all()
