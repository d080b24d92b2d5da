"""The C++ API as a drop-in: tests/cpp/mode_b_chain.cc, built against
include/ and libblade_amd.so by tests/cpp/build_cpp.py (run by the build
step), connects Beamformer::ATA -> Channelizer(rate 4) -> Polarizer in a
Runner the way the reference's C++ users do; the Pipeline fuses them into
one launch.  Its beams are checked against the complex128 oracle (normwise
relative error <= 1e-5)."""

import os
import subprocess
from pathlib import Path

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

EXE = Path(__file__).resolve().parent / "cpp" / "_build" / "mode_b_chain"


@pytest.mark.parametrize("A,F,T,B", [(28, 8, 2048, 16), (20, 3, 512, 8)])
def test_cpp_runner_mode_b_chain(tmp_path, A, F, T, B):
    assert EXE.exists(), "tests/cpp/_build/mode_b_chain is missing: run the build step"
    rng = np.random.default_rng(A + B)
    x = O.ci8_gaussian((A, F, T, 2), rng=rng)
    tau = rng.uniform(-50e-9, 50e-9, size=(B, A))
    w = O.phasor_ata(tau, np.ones((A, F, 1, 2)), 1.0e9, 0.5e6, 0.0, 0).astype(np.complex64)
    x.tofile(tmp_path / "x.ci8")
    w.tofile(tmp_path / "w.cf32")
    r = subprocess.run([str(EXE), str(tmp_path / "x.ci8"), str(tmp_path / "w.cf32"), str(tmp_path / "y.cf32"),
                        str(A), str(F), str(T), str(B)], stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stdout[-2000:]
    # the Pipeline runs the three modules as one beamformer launch
    assert "Fused Beamformer -> Channelizer -> Polarizer into one launch." in r.stdout
    y = np.fromfile(tmp_path / "y.cf32", dtype=np.complex64).reshape(B, 4 * F, T // 4, 2)
    ref = O.polarizer(O.channelizer(O.beamformer_ata(x, w), 4), "xy", "lr")
    assert O.norm_rel_err(y, ref) <= 1e-5


JIT_EXE = Path(__file__).resolve().parent / "cpp" / "_build" / "jit_module"


def test_cpp_module_compiled_at_run_time(tmp_path):
    # tests/cpp/jit_module.cc: a module written against the reference's
    # Module base (program in the constructor, createKernel, runKernel) whose
    # kernel hipRTC compiles at run time, in a Runner before the Polarizer.
    assert JIT_EXE.exists(), "tests/cpp/_build/jit_module is missing: run the build step"
    A, F, T = 3, 5, 1000
    rng = np.random.default_rng(9)
    x = (rng.normal(size=(A, F, T, 2)) + 1j * rng.normal(size=(A, F, T, 2))).astype(np.complex64)
    x.tofile(tmp_path / "x.cf32")
    r = subprocess.run([str(JIT_EXE), str(tmp_path / "x.cf32"), str(tmp_path / "y.cf32"),
                        str(tmp_path / "d.u64"), str(A), str(F), str(T)],
                       stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=120)
    assert r.returncode == 0, r.stdout[-2000:]
    y = np.fromfile(tmp_path / "y.cf32", dtype=np.complex64).reshape(A, F, T, 2)
    # scale<CF32, CF32, 2> with factor 0.5 is exact, so is the polarizer
    assert np.array_equal(y, O.polarizer(x.astype(np.complex128)).astype(np.complex64))
    d = np.fromfile(tmp_path / "d.u64", dtype=np.uint64)
    assert list(d) == [A * F * T * 2, A, F, T, 2, F * T * 2, T * 2, 2, 1]


VEC_EXE = Path(__file__).resolve().parent / "cpp" / "_build" / "vector_semantics"


@pytest.mark.gpu
def test_cpp_vector_semantics_with_device_memory():
    # The reference's Vector cases (tests/memory/vector/reference.cc:15-208)
    # with real hipMalloc / hipHostMalloc allocations.
    assert VEC_EXE.exists(), "tests/cpp/_build/vector_semantics is missing: run the build step"
    r = subprocess.run([str(VEC_EXE)], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=120)
    assert r.returncode == 0, r.stdout
    assert "vector semantics: 0 failures" in r.stdout
