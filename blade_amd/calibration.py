"""Differentiable beamforming for gain and delay calibration.

Two ``torch.autograd.Function``s back onto the hand-written kernels of the
C-ABI (``blade_beamformer_adjoint``, ``blade_beamformer_stokes_i_grad``):

* :class:`Beamform` -- ``y = beamformer(x, w)`` (FP32 matrix cores); its
  backward is the adjoint contraction ``grad_w = sum_t grad_y conj(x)``.
* :class:`StokesIPower` -- ``L = -sum_b lambda_b sum |y_b|^2``, the negative
  detected (Stokes-I) power of weighted beams.  Forward and gradient come from
  one fused kernel that recomputes the beams tile by tile instead of storing
  them over time; backward only scales the saved gradient.
* :class:`SteeredStokesIPower` -- the same loss of ``w * g/|g|`` as a
  function of the gains ``g``: the steering and the chain rule through it
  are two more kernels (``blade_steer_gains``, ``blade_gain_gradient``).

Complex gradients follow PyTorch's convention (``dL/dRe w + j dL/dIm w``,
i.e. twice the conjugate Wirtinger derivative), so the functions compose with
ordinary torch code for the parameterisation of the phasors.

:func:`calibrate` runs Adam on per-antenna complex gains (unit modulus) and
per-antenna delays that steer ``w_eff = w * g/|g| * exp(-j 2 pi f dtau)``.
The forward model is the reference's beamformer
(``src/modules/beamformer/beamformer.cu:13-79``); the calibration itself has
no reference counterpart.
"""

from __future__ import annotations

import math

import torch

from . import ops

__all__ = ["Beamform", "StokesIPower", "SteeredStokesIPower", "beamform", "stokes_i_power",
           "calibrate", "steer_phasors"]


class Beamform(torch.autograd.Function):
    """``y[b,f,t,p] = sum_a w[b,a,f,p] x[a,f,t,p]``; x CI8 (no gradient)."""

    @staticmethod
    def forward(ctx, x, w):
        ctx.save_for_backward(x)
        return ops.beamformer(x, w.contiguous(), precision="f32")

    @staticmethod
    def backward(ctx, grad_y):
        (x,) = ctx.saved_tensors
        grad_w = None
        if ctx.needs_input_grad[1]:
            grad_w = ops.beamformer_adjoint(x, grad_y.contiguous())
        return None, grad_w


class StokesIPower(torch.autograd.Function):
    """``L = -sum_b beam_weights[b] sum_{f,t,p} |y_b|^2`` (FP64 scalar).

    ``precision``: "f32" or "f16x2" (FP16 matrix cores with hi + lo halves)."""

    @staticmethod
    def forward(ctx, x, w, beam_weights, precision="f32"):
        loss, grad = ops.stokes_i_loss_and_grad(x, w.contiguous(), beam_weights, precision=precision)
        ctx.save_for_backward(grad)
        return loss

    @staticmethod
    def backward(ctx, grad_loss):
        (grad,) = ctx.saved_tensors
        # PyTorch's complex gradient is twice dL/dconj(w).
        return None, grad * (2.0 * grad_loss.to(torch.float32)), None, None


class SteeredStokesIPower(torch.autograd.Function):
    """``StokesIPower`` of ``w * g/|g|`` as a function of the gains ``g`` [A,F,P]
    (``w`` is a constant).  One call runs the steering, the fused loss +
    phasor gradient and the gain gradient; backward only scales the saved
    gain gradient."""

    @staticmethod
    def forward(ctx, x, w, gains, beam_weights, precision="f32"):
        w = w.contiguous()
        gains = gains.detach().contiguous()
        w_eff = ops.steer_gains(w, gains)
        loss, grad_w = ops.stokes_i_loss_and_grad(x, w_eff, beam_weights, precision=precision)
        ctx.save_for_backward(ops.gain_gradient(w, grad_w, gains))
        return loss

    @staticmethod
    def backward(ctx, grad_loss):
        (grad_g,) = ctx.saved_tensors
        # PyTorch's complex gradient is twice dL/dconj(g).
        return None, None, grad_g * (2.0 * grad_loss.to(torch.float32)), None, None


def beamform(x: torch.Tensor, w: torch.Tensor) -> torch.Tensor:
    return Beamform.apply(x, w)


def stokes_i_power(x: torch.Tensor, w: torch.Tensor, beam_weights=None, precision: str = "f32") -> torch.Tensor:
    if beam_weights is None:
        beam_weights = [1.0] + [0.0] * (w.shape[0] - 1)
    return StokesIPower.apply(x, w, list(beam_weights), precision)


def steer_phasors(w: torch.Tensor, gains: torch.Tensor | None, delays: torch.Tensor | None,
                  freqs_hz: torch.Tensor | None) -> torch.Tensor:
    """``w_eff[b,a,f,0,p] = w * g/|g| [a,f,p] * exp(-j 2 pi freq[f] delays[a])``."""
    out = w
    if gains is not None:
        unit = gains / gains.abs().clamp_min(1e-30)
        out = out * unit[None, :, :, None, :]
    if delays is not None:
        phase = -2.0 * math.pi * delays[:, None] * freqs_hz[None, :]          # [A, F]
        rot = torch.polar(torch.ones_like(phase), phase).to(w.dtype)
        out = out * rot[None, :, :, None, None]
    return out


def _replay(step, steps: int, losses: torch.Tensor, graph: bool, device) -> None:
    """Run ``step`` ``steps`` times into ``losses``: eagerly, or a few eager
    steps on a side stream and then one captured step replayed."""
    if not graph:
        for i in range(steps):
            losses[i].copy_(step())
        return
    side = torch.cuda.Stream(device=device)
    side.wait_stream(torch.cuda.current_stream(device))
    warm = min(3, steps)
    with torch.cuda.stream(side):
        for i in range(warm):
            losses[i].copy_(step())
    torch.cuda.current_stream(device).wait_stream(side)
    if steps > warm:
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g, stream=side):
            static_loss = step()
        for i in range(warm, steps):
            g.replay()
            losses[i].copy_(static_loss)


def _calibrate_fused(x, w, steps, lr, bw, graph, precision, fit_gains, fit_delays, freqs_hz, delay_lr):
    """Every step on the library's kernels: steering (``blade_steer_gains``),
    fused loss + phasor gradient, then the gain gradient with its Adam update
    (``blade_gain_adam``) and the delay gradient with its FP64 Adam update
    (``blade_delay_adam``) -- torch.optim.Adam's arithmetic, betas
    (0.9, 0.999), eps 1e-8."""
    A, F, P = w.shape[1], w.shape[2], w.shape[4]
    w = w.contiguous()
    dev = w.device
    gains = torch.ones((A, F, P), dtype=torch.complex64, device=dev) if fit_gains else None
    delays = torch.zeros((A,), dtype=torch.float64, device=dev) if fit_delays else None
    if fit_gains:
        g_m, g_v = torch.zeros_like(gains), torch.zeros_like(gains)
        g_step = torch.zeros((), dtype=torch.float32, device=dev)
    if fit_delays:
        freqs_hz = freqs_hz.to(device=dev, dtype=torch.float64).contiguous()
        d_m, d_v = torch.zeros_like(delays), torch.zeros_like(delays)
        d_step = torch.zeros((), dtype=torch.float32, device=dev)
    losses = torch.zeros(steps, dtype=torch.float64, device=dev)
    if steps == 0:
        return gains, delays, []

    def steered():
        # with both, the gain gradient needs the delay-steered phasors w r
        w_r = ops.steer_gains(w, None, delays, freqs_hz) if fit_delays else w
        return w_r, (ops.steer_gains(w_r, gains) if fit_gains else w_r)

    ref, _ = ops.stokes_i_loss_and_grad(x, steered()[1], bw, precision=precision)
    ref = abs(float(ref))
    inv = 1.0 / ref if ref > 0 else 1.0

    def step():
        w_r, w_eff = steered()
        loss, grad_w = ops.stokes_i_loss_and_grad(x, w_eff, bw, precision=precision)
        if fit_gains:
            ops.gain_adam(w_r, grad_w, gains, g_m, g_v, g_step, lr, grad_scale=inv)
        if fit_delays:
            ops.delay_adam(w_eff, grad_w, freqs_hz, delays, d_m, d_v, d_step, delay_lr, grad_scale=inv)
        return loss * inv

    _replay(step, steps, losses, graph, dev)
    return gains, delays, losses.tolist()


def calibrate(x: torch.Tensor, w: torch.Tensor, steps: int = 200, lr: float = 0.05,
              beam_weights=None, fit_gains: bool = True, fit_delays: bool = False,
              freqs_hz: torch.Tensor | None = None, delay_lr: float = 1e-10, graph: bool = True,
              precision: str = "f32", fused_adam: bool = True):
    """Adam on unit-modulus gains [A,F,P] (and delays [A]) maximising beam power.

    Returns ``(gains, delays, losses)``; the loss is normalised by its value at
    the start so that its history is comparable across data sets.  The losses
    stay on the device until the end (one read).  With ``graph`` the step
    (steering, fused loss and gradient, chain rule, Adam) is captured once in
    a CUDA graph after a few eager steps on the same stream and replayed.
    ``precision`` selects the loss + gradient kernel ("f32" or "f16x2").
    With ``fused_adam`` every step is the library's kernels, the Adam updates
    included (``blade_gain_adam``, ``blade_delay_adam``); otherwise
    torch.optim.Adam (capturable) after autograd.
    """
    if fit_delays and freqs_hz is None:
        raise ValueError("delay calibration needs the channel frequencies")
    if fused_adam and (fit_gains or fit_delays):
        bw = list(beam_weights) if beam_weights is not None else [1.0] + [0.0] * (w.shape[0] - 1)
        return _calibrate_fused(x, w, steps, lr, bw, graph, precision, fit_gains, fit_delays, freqs_hz, delay_lr)
    A, F, P = w.shape[1], w.shape[2], w.shape[4]
    params = []
    gains = delays = None
    if fit_gains:
        gains = torch.ones((A, F, P), dtype=torch.complex64, device=w.device, requires_grad=True)
        params.append({"params": [gains], "lr": lr})
    if fit_delays:
        if freqs_hz is None:
            raise ValueError("delay calibration needs the channel frequencies")
        delays = torch.zeros((A,), dtype=torch.float64, device=w.device, requires_grad=True)
        params.append({"params": [delays], "lr": delay_lr})
    # capturable Adam (step counts on the device) with or without a graph, so
    # both run the same arithmetic
    opt = torch.optim.Adam(params, capturable=True, foreach=True)
    losses = torch.zeros(steps, dtype=torch.float64, device=w.device)
    if steps == 0:
        return gains, delays, []
    with torch.no_grad():
        ref = abs(float(stokes_i_power(x, steer_phasors(w, gains, delays, freqs_hz), beam_weights, precision)))
    inv = 1.0 / ref if ref > 0 else 1.0

    bw = list(beam_weights) if beam_weights is not None else [1.0] + [0.0] * (w.shape[0] - 1)

    def step():
        opt.zero_grad(set_to_none=True)
        if gains is not None and delays is None:
            # gains only: steering and its chain rule on the device kernels
            loss = SteeredStokesIPower.apply(x, w, gains, bw, precision)
        else:
            loss = stokes_i_power(x, steer_phasors(w, gains, delays, freqs_hz), beam_weights, precision)
        (loss * inv).backward()
        opt.step()
        return loss.detach() * inv

    _replay(step, steps, losses, graph, w.device)
    return gains, delays, losses.tolist()
