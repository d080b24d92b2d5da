"""ctypes binding of ``libblade_amd.so`` (the C-ABI of ``include/blade_amd.h``).

The library is built in-tree by ``blade_amd.build``; there is no fallback:
if it is missing or fails to load, every entry point raises.
"""

from __future__ import annotations

import ctypes as ct
import os
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libblade_amd.so"

# enum blade_type / blade_precision / blade_result (include/blade_amd.h)
I8, F16, F32, F64, I32, CI8, CF16, CF32, CF64 = range(9)
PREC_F32, PREC_F16X2 = 0, 1
SUCCESS, ERROR, HIP_ERROR, ASSERTION_ERROR = 0, 1, 2, 3

TYPE_CODES = {"i8": I8, "f16": F16, "f32": F32, "f64": F64, "i32": I32,
              "ci8": CI8, "cf16": CF16, "cf32": CF32, "cf64": CF64}

_u64 = ct.c_uint64
_p = ct.c_void_p
_i = ct.c_int
_d = ct.c_double

# name -> argtypes (restype is int unless listed in _RESTYPES)
SIGNATURES = {
    "blade_amd_version": [],
    "blade_amd_last_error": [],
    "blade_cast": [_i, _i, _p, _p, _u64, _p],
    "blade_beamformer_ata": [_i, _i, _p, _p, _p, _u64, _u64, _u64, _u64, _u64, _i, _i, _p],
    "blade_beamform_chain": [_i, _i, _p, _p, _p, _u64, _u64, _u64, _u64, _u64, _i, _i, _u64, _i, _p],
    "blade_beamformer_workspace_size": [_i, _u64, _u64, _u64],
    "blade_beamform_chain_noalloc": [_i, _i, _p, _p, _p, _u64, _u64, _u64, _u64, _u64, _i, _i, _u64, _i,
                                     _p],
    "blade_beamform_chain_ws": [_i, _i, _p, _p, _p, _u64, _u64, _u64, _u64, _u64, _i, _i, _u64, _i,
                                _p, _u64, _p],
    "blade_beamform_batched_workspace_size": [_i, _p, _u64, _p],
    "blade_beamform_batched": [_i, _p, _u64, _p, _p, _u64, _p],
    "blade_beamformer_adjoint": [_p, _p, _p, _u64, _u64, _u64, _u64, _u64, _p],
    "blade_beamformer_stokes_i_grad": [_i, _p, _p, _p, _p, _p, _u64, _u64, _u64, _u64, _u64, _p],
    "blade_steer_gains": [_p, _p, _p, _p, _p, _u64, _u64, _u64, _u64, _p],
    "blade_gain_gradient": [_p, _p, _p, _p, _u64, _u64, _u64, _u64, _p],
    "blade_gain_adam": [_p, _p, _p, _p, _p, _p, ct.c_float, ct.c_float, ct.c_float, ct.c_float, ct.c_float,
                        _u64, _u64, _u64, _u64, _p],
    "blade_delay_adam": [_p, _p, _p, _p, _p, _p, _p, _d, _d, _d, _d, _d, _u64, _u64, _u64, _u64, _p],
    "blade_channelizer": [_p, _p, _u64, _u64, _u64, _u64, _u64, _p],
    "blade_channelizer_plan": [_u64, _p],
    "blade_channelizer_detector": [_p, _p, _u64, _u64, _u64, _u64, _u64, _u64, _i, _i, _p],
    "blade_polarizer": [_i, _p, _p, _u64, _p],
    "blade_detector": [_p, _p, _u64, _u64, _u64, _u64, _u64, _i, _p],
    "blade_gather": [_p, _p, _u64, _u64, _u64, _u64, _p],
    "blade_permutation": [_p, _p, _u64, ct.POINTER(_u64), ct.POINTER(_u64), _p],
    "blade_phasor_ata": [_p, _p, _p, _i, _u64, _u64, _u64, _u64, _d, _d, _d, _u64, _p],
    "blade_ata_delays": [_p, _u64, _u64, _p, _p, _p, _u64, _d, _d, _p],
}
_RESTYPES = {"blade_amd_version": ct.c_char_p, "blade_amd_last_error": ct.c_char_p,
             "blade_beamformer_workspace_size": ct.c_uint64,
             "blade_beamform_batched_workspace_size": ct.c_uint64}


class BeamformBlock(ct.Structure):
    """``blade_beamform_block`` of include/blade_amd.h."""
    _fields_ = [("input", ct.c_void_p), ("phasors", ct.c_void_p), ("output", ct.c_void_p),
                ("A", ct.c_uint64), ("F", ct.c_uint64), ("T", ct.c_uint64), ("B", ct.c_uint64)]


class BatchTuning(ct.Structure):
    """``blade_batch_tuning`` of include/blade_amd.h (0 = automatic)."""
    _fields_ = [("time_samples_per_lane", ct.c_int32), ("beam_tiles", ct.c_int32),
                ("chunks_per_item", ct.c_int32), ("workgroups_per_cu", ct.c_int32)]

_lib = None


class BladeError(RuntimeError):
    """A C-ABI call returned a non-success Result code."""

    def __init__(self, name, code, message):
        super().__init__(f"{name} failed with result {code}: {message}")
        self.code = code


def load(path: str | os.PathLike | None = None) -> ct.CDLL:
    """Load the library (once) and declare every entry point's signature."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = Path(path) if path else LIB_PATH
    if not p.exists():
        raise RuntimeError(
            f"{p} is missing: build it with `python -m blade_amd.build` "
            "(the HIP library is required; there is no CPU fallback)")
    try:
        import torch  # noqa: F401  (one HIP runtime per process; see blade_amd/__init__.py)
    except ImportError:
        pass
    lib = ct.CDLL(str(p))
    for name, argtypes in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = argtypes
        fn.restype = _RESTYPES.get(name, ct.c_int)
    if path is None:
        _lib = lib
    return lib


def call(name: str, *args) -> None:
    """Call a C-ABI entry point and raise BladeError on failure."""
    lib = load()
    rc = getattr(lib, name)(*args)
    if rc != SUCCESS:
        msg = lib.blade_amd_last_error().decode()
        raise BladeError(name, rc, msg)


def version() -> str:
    return load().blade_amd_version().decode()
