"""Tensor-level entry points over the C-ABI, on torch device memory.

PyTorch only provides device buffers and the current HIP stream; every
computation is a call into ``libblade_amd.so``.  Conventions:

* CF32 -> ``torch.complex64``; CF64 -> ``torch.complex128``;
* CI8 -> ``torch.int8`` with a trailing axis of 2 (re, im);
* CF16 -> ``torch.float16`` with a trailing axis of 2.
"""

from __future__ import annotations

import ctypes as ct

import torch

from . import _lib
from ._lib import call

__all__ = ["type_of", "stream_handle", "cast", "beamformer", "beamform_chain", "channelizer",
           "polarizer", "detector", "gather", "permutation", "phasor_ata", "beamformer_adjoint",
           "stokes_i_loss_and_grad", "BatchedBeamformer"]


def stream_handle(stream=None):
    s = stream if stream is not None else torch.cuda.current_stream()
    return ct.c_void_p(s.cuda_stream)


def _ptr(t: torch.Tensor):
    if not t.is_cuda:
        raise ValueError("expected a device tensor")
    if not t.is_contiguous():
        raise ValueError("expected a contiguous tensor")
    return ct.c_void_p(t.data_ptr())


def type_of(t: torch.Tensor) -> str:
    """blade type name of a tensor under the conventions above."""
    if t.dtype == torch.complex64:
        return "cf32"
    if t.dtype == torch.complex128:
        return "cf64"
    if t.dtype == torch.int8:
        return "ci8" if t.dim() and t.shape[-1] == 2 else "i8"
    if t.dtype == torch.float16:
        return "cf16" if t.dim() and t.shape[-1] == 2 else "f16"
    return {torch.float32: "f32", torch.float64: "f64", torch.int32: "i32"}[t.dtype]


def _elements(t: torch.Tensor, type_name: str) -> int:
    if type_name in ("ci8", "cf16"):
        return t.numel() // 2
    return t.numel()


def _shape4(t: torch.Tensor, type_name: str):
    shape = tuple(t.shape[:-1]) if type_name in ("ci8", "cf16") else tuple(t.shape)
    if len(shape) != 4:
        raise ValueError(f"expected a rank-4 [A,F,T,P] tensor, got {shape}")
    return shape


_TORCH_OF = {"i8": torch.int8, "f16": torch.float16, "f32": torch.float32, "f64": torch.float64,
             "i32": torch.int32, "ci8": torch.int8, "cf16": torch.float16,
             "cf32": torch.complex64, "cf64": torch.complex128}


def empty(shape, type_name: str, device="cuda"):
    shape = tuple(shape)
    if type_name in ("ci8", "cf16"):
        shape = shape + (2,)
    return torch.empty(shape, dtype=_TORCH_OF[type_name], device=device)


def cast(x: torch.Tensor, out_type: str, out: torch.Tensor | None = None, in_type: str | None = None,
         stream=None) -> torch.Tensor:
    """Component-wise cast (Modules::Cast)."""
    it = in_type or type_of(x)
    n = _elements(x, it)
    base_shape = tuple(x.shape[:-1]) if it in ("ci8", "cf16") else tuple(x.shape)
    if out is None:
        out = empty(base_shape, out_type, x.device)
    call("blade_cast", _lib.TYPE_CODES[it], _lib.TYPE_CODES[out_type], _ptr(x), _ptr(out), n,
         stream_handle(stream))
    return out


def beamformer(x: torch.Tensor, phasors: torch.Tensor, incoherent=False, incoherent_sqrt=False,
               precision="f32", out: torch.Tensor | None = None, stream=None) -> torch.Tensor:
    """ATA beamformer (Modules::Beamformer::ATA); x CF32 or CI8 [A,F,T,P], phasors CF32 [B,A,F,1,P]."""
    return beamform_chain(x, phasors, incoherent, incoherent_sqrt, 1, False, precision, out, stream)


def beamform_chain(x: torch.Tensor, phasors: torch.Tensor, incoherent=False, incoherent_sqrt=False,
                   rate=1, polarize=False, precision="f32", out: torch.Tensor | None = None,
                   stream=None) -> torch.Tensor:
    """Fused beamform -> channelize(rate 1/2/4) -> [polarize]."""
    it = type_of(x)
    A, F, T, P = _shape4(x, it)
    B = phasors.shape[0]
    if tuple(phasors.shape) != (B, A, F, 1, P) or phasors.dtype != torch.complex64:
        raise ValueError(f"phasors must be CF32 [B,{A},{F},1,{P}], got {tuple(phasors.shape)}")
    nb = B + (1 if incoherent else 0)
    if out is None:
        out = torch.empty((nb, F * rate, T // rate, P), dtype=torch.complex64, device=x.device)
    prec = {"f32": _lib.PREC_F32, "f16x2": _lib.PREC_F16X2}[precision]
    if rate == 1 and not polarize:
        call("blade_beamformer_ata", _lib.TYPE_CODES[it], prec, _ptr(x), _ptr(phasors), _ptr(out),
             A, F, T, P, B, int(incoherent), int(incoherent_sqrt), stream_handle(stream))
    else:
        call("blade_beamform_chain", _lib.TYPE_CODES[it], prec, _ptr(x), _ptr(phasors), _ptr(out),
             A, F, T, P, B, int(incoherent), int(incoherent_sqrt), rate, int(polarize),
             stream_handle(stream))
    return out


class BatchedBeamformer:
    """Beamformer blocks of different shapes in one persistent launch.

    ``blocks``: list of ``(x CI8 [A,F,T,P,2], phasors CF32 [B,A,F,1,P],
    out CF32 [B,F,T,P])`` device tensors (kept referenced).  The block table
    and the workspace are built once; every call launches the batch on
    ``stream`` and allocates nothing (capturable in a hipGraph).
    ``tuning``: dict of the ``blade_batch_tuning`` fields (missing = auto).
    """

    def __init__(self, blocks, precision="f32", tuning=None):
        self.blocks = [tuple(b) for b in blocks]
        self.precision = {"f32": _lib.PREC_F32, "f16x2": _lib.PREC_F16X2}[precision]
        self.table = (_lib.BeamformBlock * max(1, len(self.blocks)))()
        for k, (x, w, y) in enumerate(self.blocks):
            A, F, T, P = _shape4(x, "ci8")
            B = w.shape[0]
            if type_of(x) != "ci8":
                raise ValueError("batched blocks take CI8 voltages")
            if tuple(w.shape) != (B, A, F, 1, P) or w.dtype != torch.complex64:
                raise ValueError(f"block {k}: phasors must be CF32 [B,{A},{F},1,{P}]")
            if tuple(y.shape) != (B, F, T, P) or y.dtype != torch.complex64:
                raise ValueError(f"block {k}: output must be CF32 [{B},{F},{T},{P}]")
            if P != 2:
                raise ValueError("two polarizations")
            self.table[k] = _lib.BeamformBlock(_ptr(x).value, _ptr(w).value, _ptr(y).value, A, F, T, B)
        self.set_tuning(tuning)

    def set_tuning(self, tuning=None):
        t = dict(tuning or {})
        self.tuning = _lib.BatchTuning(int(t.get("time_samples_per_lane", 0)), int(t.get("beam_tiles", 0)),
                                       int(t.get("chunks_per_item", 0)), int(t.get("workgroups_per_cu", 0)))
        lib = _lib.load()
        n = len(self.blocks)
        size = lib.blade_beamform_batched_workspace_size(self.precision, self.table, n, ct.byref(self.tuning))
        if size == 0 and n:
            raise _lib.BladeError("blade_beamform_batched_workspace_size", 1,
                                  lib.blade_amd_last_error().decode())
        dev = self.blocks[0][0].device if n else torch.device("cuda")
        if getattr(self, "workspace", None) is None or self.workspace.numel() < size:
            self.workspace = torch.empty(max(size, 256), dtype=torch.uint8, device=dev)
        return self

    def __call__(self, stream=None):
        call("blade_beamform_batched", self.precision, self.table, len(self.blocks), ct.byref(self.tuning),
             _ptr(self.workspace), self.workspace.numel(), stream_handle(stream))
        return [y for (_, _, y) in self.blocks]


def beamformer_adjoint(x: torch.Tensor, grad_output: torch.Tensor, out: torch.Tensor | None = None,
                       stream=None) -> torch.Tensor:
    """dL/dconj(w) [B,A,F,1,P] = sum_t grad_output[b,f,t,p] conj(x[a,f,t,p]); x CI8 [A,F,T,P]."""
    A, F, T, P = _shape4(x, "ci8")
    if type_of(x) != "ci8" or x.dtype != torch.int8:
        raise ValueError("the adjoint takes CI8 voltages")
    B = grad_output.shape[0]
    if tuple(grad_output.shape) != (B, F, T, P) or grad_output.dtype != torch.complex64:
        raise ValueError(f"grad_output must be CF32 [B,{F},{T},{P}], got {tuple(grad_output.shape)}")
    if out is None:
        out = torch.empty((B, A, F, 1, P), dtype=torch.complex64, device=x.device)
    call("blade_beamformer_adjoint", _ptr(x), _ptr(grad_output), _ptr(out), A, F, T, P, B,
         stream_handle(stream))
    return out


def stokes_i_loss_and_grad(x: torch.Tensor, phasors: torch.Tensor, beam_weights,
                           stream=None, precision: str = "f32") -> tuple[torch.Tensor, torch.Tensor]:
    """L = -sum_b beam_weights[b] sum |y_b|^2 (FP64 device scalar) and dL/dconj(w) CF32.

    ``precision``: "f32" (FP32 matrix cores) or "f16x2" (FP16 matrix cores,
    hi + lo split phasors and gradients, FP32 accumulation)."""
    A, F, T, P = _shape4(x, "ci8")
    B = phasors.shape[0]
    if tuple(phasors.shape) != (B, A, F, 1, P) or phasors.dtype != torch.complex64:
        raise ValueError(f"phasors must be CF32 [B,{A},{F},1,{P}], got {tuple(phasors.shape)}")
    lam = [float(v) for v in beam_weights]
    if len(lam) != B:
        raise ValueError(f"need one beam weight per beam ({B}), got {len(lam)}")
    weights = (ct.c_float * B)(*lam)
    loss = torch.empty((), dtype=torch.float64, device=x.device)
    grad = torch.empty((B, A, F, 1, P), dtype=torch.complex64, device=x.device)
    prec = {"f32": _lib.PREC_F32, "f16x2": _lib.PREC_F16X2}[precision]
    call("blade_beamformer_stokes_i_grad", prec, _ptr(x), _ptr(phasors), ct.cast(weights, ct.c_void_p),
         _ptr(loss), _ptr(grad), A, F, T, P, B, stream_handle(stream))
    return loss, grad


def _gain_shapes(phasors: torch.Tensor, gains: torch.Tensor):
    if phasors.dim() != 5 or phasors.shape[3] != 1 or phasors.dtype != torch.complex64:
        raise ValueError(f"phasors must be CF32 [B,A,F,1,P], got {tuple(phasors.shape)} {phasors.dtype}")
    B, A, F, _, P = phasors.shape
    if tuple(gains.shape) != (A, F, P) or gains.dtype != torch.complex64:
        raise ValueError(f"gains must be CF32 [{A},{F},{P}], got {tuple(gains.shape)} {gains.dtype}")
    if not (phasors.is_contiguous() and gains.is_contiguous()):
        raise ValueError("phasors and gains must be contiguous")
    return B, A, F, P


def _delay_shapes(A: int, F: int, delays: torch.Tensor, freqs_hz: torch.Tensor | None):
    if tuple(delays.shape) != (A,) or delays.dtype != torch.float64 or not delays.is_contiguous():
        raise ValueError(f"delays must be contiguous float64 [{A}], got {tuple(delays.shape)} {delays.dtype}")
    if freqs_hz is None or tuple(freqs_hz.shape) != (F,) or freqs_hz.dtype != torch.float64 \
            or not freqs_hz.is_contiguous():
        raise ValueError(f"delays need contiguous float64 channel frequencies [{F}]")


def steer_gains(phasors: torch.Tensor, gains: torch.Tensor | None, delays: torch.Tensor | None = None,
                freqs_hz: torch.Tensor | None = None, stream=None) -> torch.Tensor:
    """phasors * gains/|gains| * exp(-j 2 pi freqs_hz[f] delays[a]): unit-modulus
    gains [A,F,P] and/or FP64 delays [A] (seconds) on phasors [B,A,F,1,P]."""
    if gains is not None:
        B, A, F, P = _gain_shapes(phasors, gains)
    else:
        if phasors.dim() != 5 or phasors.shape[3] != 1 or phasors.dtype != torch.complex64 \
                or not phasors.is_contiguous():
            raise ValueError("phasors must be contiguous CF32 [B,A,F,1,P]")
        B, A, F, _, P = phasors.shape
    if delays is not None:
        _delay_shapes(A, F, delays, freqs_hz)
    out = torch.empty_like(phasors)
    call("blade_steer_gains", _ptr(phasors), _ptr(gains) if gains is not None else None,
         _ptr(delays) if delays is not None else None, _ptr(freqs_hz) if delays is not None else None,
         _ptr(out), B, A, F, P, stream_handle(stream))
    return out


def gain_gradient(phasors: torch.Tensor, grad_phasors: torch.Tensor, gains: torch.Tensor,
                  stream=None) -> torch.Tensor:
    """dL/dconj(gains) from grad_phasors = dL/dconj(phasors * gains / |gains|)."""
    B, A, F, P = _gain_shapes(phasors, gains)
    if tuple(grad_phasors.shape) != tuple(phasors.shape) or grad_phasors.dtype != torch.complex64 \
            or not grad_phasors.is_contiguous():
        raise ValueError("grad_phasors must be contiguous CF32 of the phasors' shape")
    out = torch.empty_like(gains)
    call("blade_gain_gradient", _ptr(phasors), _ptr(grad_phasors), _ptr(gains), _ptr(out), B, A, F, P,
         stream_handle(stream))
    return out


def gain_adam(phasors: torch.Tensor, grad_phasors: torch.Tensor, gains: torch.Tensor, exp_avg: torch.Tensor,
              exp_avg_sq: torch.Tensor, step: torch.Tensor, lr: float, betas=(0.9, 0.999), eps: float = 1e-8,
              grad_scale: float = 1.0, stream=None) -> None:
    """One Adam step on ``gains`` (in place) from grad_phasors = dL/dconj(phasors * gains/|gains|):
    the gradient ``2 * grad_scale * dL/dconj(gains)`` (PyTorch's convention), Adam
    states ``exp_avg``/``exp_avg_sq`` (CF32, gains' shape) and ``step`` (F32 device
    scalar, incremented first), as torch.optim.Adam without weight decay."""
    B, A, F, P = _gain_shapes(phasors, gains)
    for name, t in (("grad_phasors", grad_phasors), ("exp_avg", exp_avg), ("exp_avg_sq", exp_avg_sq)):
        want = phasors.shape if name == "grad_phasors" else gains.shape
        if tuple(t.shape) != tuple(want) or t.dtype != torch.complex64 or not t.is_contiguous():
            raise ValueError(f"{name} must be contiguous CF32 {tuple(want)}")
    if step.dtype != torch.float32 or step.numel() != 1 or step.device != gains.device:
        raise ValueError("step must be a one-element float32 tensor on the gains' device")
    call("blade_gain_adam", _ptr(phasors), _ptr(grad_phasors), _ptr(gains), _ptr(exp_avg), _ptr(exp_avg_sq),
         _ptr(step), float(lr), float(betas[0]), float(betas[1]), float(eps), float(grad_scale), B, A, F, P,
         stream_handle(stream))


def delay_adam(steered: torch.Tensor, grad_phasors: torch.Tensor, freqs_hz: torch.Tensor, delays: torch.Tensor,
               exp_avg: torch.Tensor, exp_avg_sq: torch.Tensor, step: torch.Tensor, lr: float,
               betas=(0.9, 0.999), eps: float = 1e-8, grad_scale: float = 1.0, stream=None) -> None:
    """One FP64 Adam step on ``delays`` [A] (in place) from steered = steer_gains(...)
    and grad_phasors = dL/dconj(steered); ``exp_avg``/``exp_avg_sq`` float64 [A],
    ``step`` F32 device scalar (incremented first), as torch.optim.Adam."""
    if steered.dim() != 5 or steered.dtype != torch.complex64 or not steered.is_contiguous():
        raise ValueError("steered phasors must be contiguous CF32 [B,A,F,1,P]")
    B, A, F, _, P = steered.shape
    if tuple(grad_phasors.shape) != tuple(steered.shape) or grad_phasors.dtype != torch.complex64 \
            or not grad_phasors.is_contiguous():
        raise ValueError("grad_phasors must be contiguous CF32 of the steered phasors' shape")
    _delay_shapes(A, F, delays, freqs_hz)
    for name, t in (("exp_avg", exp_avg), ("exp_avg_sq", exp_avg_sq)):
        if tuple(t.shape) != (A,) or t.dtype != torch.float64 or not t.is_contiguous():
            raise ValueError(f"{name} must be contiguous float64 [{A}]")
    if step.dtype != torch.float32 or step.numel() != 1 or step.device != delays.device:
        raise ValueError("step must be a one-element float32 tensor on the delays' device")
    call("blade_delay_adam", _ptr(steered), _ptr(grad_phasors), _ptr(freqs_hz), _ptr(delays), _ptr(exp_avg),
         _ptr(exp_avg_sq), _ptr(step), float(lr), float(betas[0]), float(betas[1]), float(eps),
         float(grad_scale), B, A, F, P, stream_handle(stream))


def channelizer(x: torch.Tensor, rate: int, out: torch.Tensor | None = None, stream=None):
    """FFT channelizer (Modules::Channelizer), CF32 [A,F,T,P] -> [A,F*C,T/C,P]."""
    A, F, T, P = _shape4(x, "cf32")
    if out is None:
        out = torch.empty((A, F * rate, T // rate, P), dtype=torch.complex64, device=x.device)
    call("blade_channelizer", _ptr(x), _ptr(out), A, F, T, P, rate, stream_handle(stream))
    return out


def channelizer_detector(x: torch.Tensor, rate: int, integration: int, out_pols: int = 1,
                         out: torch.Tensor | None = None, accumulate=False, stream=None,
                         polarize=False):
    """Fused channelize -> [polarize XY -> LR] -> detect,
    CF32 [A,F,T,2] -> F32 [A,F*C,S/N or 1,out_pols]."""
    A, F, T, P = _shape4(x, "cf32")
    S = T // rate
    groups = S // integration if integration <= S else 1
    if out is None:
        out = torch.empty((A, F * rate, groups, out_pols), dtype=torch.float32, device=x.device)
    call("blade_channelizer_detector", _ptr(x), _ptr(out), A, F, T, rate, integration, out_pols,
         int(polarize), int(accumulate), stream_handle(stream))
    return out


def polarizer(x: torch.Tensor, out: torch.Tensor | None = None, stream=None):
    """XY -> LR (Modules::Polarizer), CF32 or CF16."""
    t = type_of(x)
    if out is None:
        out = torch.empty_like(x)
    call("blade_polarizer", _lib.TYPE_CODES[t], _ptr(x), _ptr(out), _elements(x, t),
         stream_handle(stream))
    return out


def detector(x: torch.Tensor, integration: int, out_pols: int, out: torch.Tensor | None = None,
             accumulate=False, stream=None):
    """Detector (Modules::Detector), CF32 [X,F,T,2] -> F32 [X,F,T/INTG,out_pols]."""
    X, F, T, P = _shape4(x, "cf32")
    if out is None:
        out = torch.empty((X, F, T // integration, out_pols), dtype=torch.float32, device=x.device)
    call("blade_detector", _ptr(x), _ptr(out), X, F, T, integration, out_pols, int(accumulate),
         stream_handle(stream))
    return out


def gather(x: torch.Tensor, out: torch.Tensor, axis: int, multiplier: int, step: int,
           type_name: str | None = None, stream=None):
    """Place block ``step`` of ``multiplier`` along ``axis`` (Modules::Gather)."""
    t = type_name or type_of(x)
    shape = tuple(x.shape[:-1]) if t in ("ci8", "cf16") else tuple(x.shape)
    esize = x.element_size() * (2 if t in ("ci8", "cf16") else 1)
    height = 1
    for d in shape[:axis]:
        height *= d
    width = esize
    for d in shape[axis:]:
        width *= d
    call("blade_gather", _ptr(x), _ptr(out), height, width, multiplier, step, stream_handle(stream))
    return out


def permutation(x: torch.Tensor, indexes, type_name: str | None = None, stream=None):
    """out = transpose(x, indexes) for a rank-4 tensor (Modules::Permutation)."""
    t = type_name or type_of(x)
    shape = tuple(x.shape[:-1]) if t in ("ci8", "cf16") else tuple(x.shape)
    esize = x.element_size() * (2 if t in ("ci8", "cf16") else 1)
    out_shape = tuple(shape[i] for i in indexes)
    out = empty(out_shape, t, x.device)
    s4 = (ct.c_uint64 * 4)(*shape)
    i4 = (ct.c_uint64 * 4)(*indexes)
    call("blade_permutation", _ptr(x), _ptr(out), esize, s4, i4, stream_handle(stream))
    return out


def phasor_ata(delays: torch.Tensor, calibration: torch.Tensor, observation_frequency_hz: float,
               channel_bandwidth_hz: float, total_bandwidth_hz: float, frequency_start_index: int,
               out_type="cf32", out: torch.Tensor | None = None, stream=None):
    """ATA phasors from delays (Modules::Phasor::ATA's phasor loop), FP64 phase."""
    B, A = delays.shape
    A2, F, _, P = calibration.shape
    if A2 != A or delays.dtype != torch.float64 or calibration.dtype != torch.complex128:
        raise ValueError("delays must be F64 [B,A] and calibration CF64 [A,F,1,P]")
    if out is None:
        out = torch.empty((B, A, F, 1, P), dtype=_TORCH_OF[out_type], device=delays.device)
    call("blade_phasor_ata", _ptr(delays), _ptr(calibration), _ptr(out), _lib.TYPE_CODES[out_type],
         B, A, F, P, float(observation_frequency_hz), float(channel_bandwidth_hz),
         float(total_bandwidth_hz), int(frequency_start_index), stream_handle(stream))
    return out
