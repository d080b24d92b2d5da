"""The C-ABI library loads on the CPU and exports every symbol include/*.h declares.

No compute call is made (there is no GPU here); argument validation that
returns before touching the device is exercised.
"""

import ctypes as ct
import re
from pathlib import Path

import pytest

from blade_amd import _lib

ROOT = Path(__file__).resolve().parents[1]


def declared_symbols():
    names = set()
    for h in (ROOT / "include").glob("*.h"):
        text = re.sub(r"/\*.*?\*/", "", h.read_text(), flags=re.S)
        names |= set(re.findall(r"\b(blade_\w+)\s*\(", text))
    return names


def test_library_exports_every_declared_symbol():
    lib = ct.CDLL(str(_lib.LIB_PATH))
    names = declared_symbols()
    assert "blade_beamformer_ata" in names and len(names) >= 10
    for n in sorted(names):
        assert hasattr(lib, n), f"{n} declared but not exported"
    # every declared symbol has a ctypes signature
    assert names <= set(_lib.SIGNATURES)


def test_version_and_error_string():
    assert "gfx950" in _lib.version()


@pytest.mark.parametrize("args,msg", [
    ((_lib.CF32, _lib.F32, None, None, 4, None), "complex"),
    ((_lib.F64, _lib.F16, None, None, 4, None), "F16"),
])
def test_cast_validation(args, msg):
    with pytest.raises(_lib.BladeError, match=msg):
        _lib.call("blade_cast", *args)


def test_validation_errors_mirror_reference():
    # detector: T % INTG (src/modules/detector/base.cc:22-28)
    with pytest.raises(_lib.BladeError, match="divisible"):
        _lib.call("blade_detector", None, None, 2, 192, 8750, 3, 1, 0, None)
    # detector: number of output polarizations (base.cc:44-51)
    with pytest.raises(_lib.BladeError, match="polarizations"):
        _lib.call("blade_detector", None, None, 2, 192, 8750, 10, 2, 0, None)
    # channelizer: T % rate and odd rate (src/modules/channelizer/base.cc:18-28)
    with pytest.raises(_lib.BladeError, match="divisible"):
        _lib.call("blade_channelizer", None, None, 8, 128, 1025, 1, 1024, None)
    with pytest.raises(_lib.BladeError, match="divisible by 2"):
        _lib.call("blade_channelizer", None, None, 8, 128, 1025, 4, 1025, None)
    # beamformer: P must be 2; split FP16 needs CI8
    with pytest.raises(_lib.BladeError, match="polarizations"):
        _lib.call("blade_beamformer_ata", _lib.CF32, 0, None, None, None, 2, 2, 8, 1, 1, 0, 0, None)
    with pytest.raises(_lib.BladeError, match="CI8"):
        _lib.call("blade_beamformer_ata", _lib.CF32, 1, None, None, None, 2, 2, 8, 2, 1, 0, 0, None)
    # gather: multiplier > 0 (src/modules/gather/base.cc:146-149)
    with pytest.raises(_lib.BladeError, match="multiplier"):
        _lib.call("blade_gather", None, None, 1, 16, 0, 0, None)
    # permutation: repeated index (src/modules/permutation/base.cc:280-287)
    s4 = (ct.c_uint64 * 4)(2, 3, 4, 5)
    i4 = (ct.c_uint64 * 4)(0, 1, 1, 3)
    with pytest.raises(_lib.BladeError, match="repeated"):
        _lib.call("blade_permutation", None, None, 8, s4, i4, None)


def test_run_time_compilation_without_a_device():
    # tests/cpp/jit_compile.cc: Program::compile (hipRTC, the counterpart of
    # the reference's NVRTC/jitify2 programs, include/blade/module.hh:22-83)
    # turns a template instantiation over the device-side tensor types into a
    # gfx950 code object; a source error or an unknown kernel fails.
    import subprocess
    exe = ROOT / "tests" / "cpp" / "_build" / "jit_compile"
    assert exe.exists(), "tests/cpp/_build/jit_compile is missing: run the build step"
    r = subprocess.run([str(exe)], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=120)
    assert r.returncode == 0, r.stdout
    assert "good: 0" in r.stdout and "bad: 1 log=names the file" in r.stdout
