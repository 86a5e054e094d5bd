"""The ILT objective around the imaging adjoint, as fused device passes.

``resist_objective`` is the relaxed resist of opc/ilt.py:311-337
(``_Objective.__call__`` without its penalties): for intensities I and
targets t over the last two axes,

    f = scale * I,  s = sign * (f - level),  z = 1 / (1 + exp(-clip(a s, +-60))),
    cost = mean (z - t)^2 per image,  a = steepness / max(|level|, 1e-12),

with d cost / d I = sign * (2 / npix) * (z - t) * a * z (1 - z) * scale (the
reference's closed form), the signed field s and the pattern mismatch
count_nonzero((s > 0) != (t > 0.5)) (ilt.py:357), all from one pass over the
images.  ``mask_sigmoid`` is the mask parameterisation of ilt.py:368-389
without its Gaussian: m = 1 / (1 + exp(-clip(beta x, +-60))), d x = d m beta
m (1 - m).  ``RelaxedResist`` and ``MaskSigmoid`` are their autograd forms.

``mask_penalties`` is the TV + discreteness term of ilt.py:339-355 with its
gradient from one pass (``MaskPenalties`` its autograd form), and
``adam_step`` the solver's Adam update (ilt.py:587-591), in place.
"""
from __future__ import annotations

import torch

from . import _lib

__all__ = ["resist_objective", "resist_objective_into", "mask_sigmoid", "mask_sigmoid_grad",
           "mask_penalties", "adam_step", "ilt_schedule", "RelaxedResist", "MaskSigmoid",
           "MaskPenalties"]


def _prec(t: torch.Tensor) -> int:
    if t.dtype == torch.float64:
        return 64
    if t.dtype == torch.float32:
        return 32
    raise TypeError("float32 or float64 tensors expected")


def _on_device(t, name):
    if not isinstance(t, torch.Tensor) or t.device.type != "cuda":
        raise TypeError(f"{name} must be a tensor on the HIP device")


def resist_objective(intensity: torch.Tensor, target: torch.Tensor, scale: float, level: float,
                     steepness: float, sign: float = 1.0, grad: bool = True,
                     signed: bool = False, mismatch: bool = False) -> dict:
    """Costs per image ``cost`` [...], their sum ``total`` (0-d), and on request
    ``d_intensity`` (like ``intensity``), ``signed`` and ``mismatch`` (int32 [...])."""
    _on_device(intensity, "intensity")
    _on_device(target, "target")
    if intensity.dim() < 2:
        raise ValueError("intensity must have (..., ny, nx) image axes")
    prec = _prec(intensity)
    f = intensity.contiguous()
    t = target.to(f.dtype)
    if t.shape != f.shape:
        t = t.expand(f.shape)
    t = t.contiguous()
    lead = tuple(f.shape[:-2])
    npix = f.shape[-1] * f.shape[-2]
    cost = torch.empty(lead, dtype=f.dtype, device=f.device)
    total = torch.empty((), dtype=f.dtype, device=f.device)
    d = torch.empty_like(f) if grad else None
    s = torch.empty_like(f) if signed else None
    mm = torch.empty(lead, dtype=torch.int32, device=f.device) if mismatch else None
    ws = torch.empty(max(workspace_bytes(f.numel() // npix, npix), 256), dtype=torch.uint8,
                     device=f.device)
    resist_objective_into(f, t, scale, level, steepness, sign, cost, total, d, s, mm, ws)
    return {"cost": cost, "total": total, "d_intensity": d, "signed": s, "mismatch": mm}


def workspace_bytes(n_img: int, npix: int) -> int:
    """Scratch bytes of ``resist_objective_into`` for n_img images of npix pixels."""
    return int(_lib.load().lb_resist_objective_workspace_bytes(int(n_img), int(npix)))


def resist_objective_into(f, t, scale, level, steepness, sign, cost, total, d, signed, mismatch,
                          ws) -> None:
    """``resist_objective`` into preallocated, contiguous tensors (any output may be None):
    no allocation, so it can be captured into a graph."""
    npix = f.shape[-1] * f.shape[-2]
    st = torch.cuda.current_stream(f.device).cuda_stream

    def ptr(x):
        return x.data_ptr() if x is not None else None

    _lib.check(_lib.load().lb_resist_objective(
        f.data_ptr(), t.data_ptr(), f.numel() // npix, npix, float(scale), float(level),
        float(steepness), float(sign), ptr(cost), ptr(total), ptr(d), ptr(signed), ptr(mismatch),
        _prec(f), ws.data_ptr(), ws.numel(), st))


def _device_scalar(t: torch.Tensor, n: int, name: str) -> int:
    """Address of n float64 values on the device (a schedule entry)."""
    if t.dtype != torch.float64 or t.device.type != "cuda" or t.numel() < n or not t.is_contiguous():
        raise TypeError(f"{name} must be a contiguous float64 device tensor of {n} value(s)")
    return t.data_ptr()


def mask_sigmoid(x: torch.Tensor, beta, out: torch.Tensor | None = None) -> torch.Tensor:
    """``beta``: a number, or a float64 device tensor read when the pass runs."""
    _on_device(x, "x")
    prec = _prec(x)
    xc = x.contiguous()
    m = torch.empty_like(xc) if out is None else out
    st = torch.cuda.current_stream(xc.device).cuda_stream
    lib = _lib.load()
    if isinstance(beta, torch.Tensor):
        _lib.check(lib.lb_mask_sigmoid_dev(xc.data_ptr(), xc.numel(),
                                           _device_scalar(beta, 1, "beta"), m.data_ptr(), prec, st))
    else:
        _lib.check(lib.lb_mask_sigmoid(xc.data_ptr(), xc.numel(), float(beta), m.data_ptr(), prec,
                                       st))
    return m


def mask_sigmoid_grad(grad_mask: torch.Tensor, mask: torch.Tensor, beta,
                      out: torch.Tensor | None = None) -> torch.Tensor:
    """``beta``: a number, or a float64 device tensor read when the pass runs."""
    g = grad_mask.to(mask.dtype).contiguous()
    res = torch.empty_like(mask) if out is None else out
    st = torch.cuda.current_stream(mask.device).cuda_stream
    lib = _lib.load()
    if isinstance(beta, torch.Tensor):
        _lib.check(lib.lb_mask_sigmoid_grad_dev(g.data_ptr(), mask.data_ptr(), mask.numel(),
                                                _device_scalar(beta, 1, "beta"), res.data_ptr(),
                                                _prec(mask), st))
    else:
        _lib.check(lib.lb_mask_sigmoid_grad(g.data_ptr(), mask.data_ptr(), mask.numel(),
                                            float(beta), res.data_ptr(), _prec(mask), st))
    return res


def mask_sigmoid_grad_theta(grad_mask: torch.Tensor, theta: torch.Tensor, beta: float,
                            g: torch.Tensor | None = None,
                            out: torch.Tensor | None = None) -> torch.Tensor:
    """d theta = g * grad_mask * beta m (1 - m), m = sigmoid(clip(beta theta, +-60))
    recomputed from theta; ``g`` a 0-d device tensor of theta's dtype read when the pass
    runs (None: 1)."""
    _on_device(theta, "theta")
    th = theta.contiguous()
    gm = grad_mask.to(th.dtype).contiguous()
    res = torch.empty_like(th) if out is None else out
    gp = None
    if g is not None:
        gt = g.to(th.dtype).reshape(()).contiguous()
        _on_device(gt, "g")
        gp = gt.data_ptr()
    st = torch.cuda.current_stream(th.device).cuda_stream
    _lib.check(_lib.load().lb_mask_sigmoid_grad_theta(gm.data_ptr(), th.data_ptr(), th.numel(),
                                                      float(beta), gp, res.data_ptr(),
                                                      _prec(th), st))
    return res


def ilt_schedule(schedule: torch.Tensor, growth: float, beta1: float = 0.9,
                 beta2: float = 0.999) -> None:
    """Advance ``schedule`` = [beta, it, 1 - beta1^it, 1 - beta2^it, beta_next] (float64,
    device) by one solver iteration, on the device."""
    _lib.check(_lib.load().lb_ilt_schedule(_device_scalar(schedule, 5, "schedule"), float(growth),
                                           float(beta1), float(beta2),
                                           torch.cuda.current_stream(schedule.device).cuda_stream))


def mask_penalties(mask: torch.Tensor, weight_tv: float, weight_discrete: float,
                   grad: bool = True) -> dict:
    """Per-image penalty ``cost`` [...], their sum ``total`` and, on request, the gradient
    ``grad`` (like ``mask``) over the last two axes, periodic neighbours."""
    _on_device(mask, "mask")
    if mask.dim() < 2:
        raise ValueError("mask must have (..., ny, nx) image axes")
    m = mask.contiguous()
    ny, nx = m.shape[-2], m.shape[-1]
    n_img = m.numel() // (ny * nx)
    cost = torch.empty(tuple(m.shape[:-2]), dtype=m.dtype, device=m.device)
    total = torch.empty((), dtype=m.dtype, device=m.device)
    g = torch.empty_like(m) if grad else None
    ws = torch.empty(max(workspace_bytes(n_img, ny * nx), 256), dtype=torch.uint8,
                     device=m.device)
    st = torch.cuda.current_stream(m.device).cuda_stream
    _lib.check(_lib.load().lb_mask_penalties(
        m.data_ptr(), n_img, ny, nx, float(weight_tv), float(weight_discrete), cost.data_ptr(),
        total.data_ptr(), g.data_ptr() if g is not None else None, _prec(m), ws.data_ptr(),
        ws.numel(), st))
    return {"cost": cost, "total": total, "grad": g}


def adam_step(theta: torch.Tensor, mom1: torch.Tensor, mom2: torch.Tensor, grad: torch.Tensor,
              iteration, step: float, beta1: float = 0.9, beta2: float = 0.999,
              eps: float = 1e-8) -> None:
    """theta, mom1, mom2 updated in place (contiguous tensors of one dtype).  ``iteration``:
    the step number (>= 1), or a float64 device tensor holding the two bias corrections
    [1 - beta1^it, 1 - beta2^it], read when the pass runs."""
    for t in (theta, mom1, mom2, grad):
        _on_device(t, "Adam operand")
        if not t.is_contiguous() or t.dtype != theta.dtype or t.shape != theta.shape:
            raise ValueError("Adam operands must be contiguous tensors of one shape and dtype")
    st = torch.cuda.current_stream(theta.device).cuda_stream
    lib = _lib.load()
    if isinstance(iteration, torch.Tensor):
        _lib.check(lib.lb_adam_step_dev(theta.data_ptr(), mom1.data_ptr(), mom2.data_ptr(),
                                        grad.data_ptr(), theta.numel(),
                                        _device_scalar(iteration, 2, "corrections"), float(step),
                                        float(beta1), float(beta2), float(eps), _prec(theta), st))
    else:
        _lib.check(lib.lb_adam_step(theta.data_ptr(), mom1.data_ptr(), mom2.data_ptr(),
                                    grad.data_ptr(), theta.numel(), int(iteration), float(step),
                                    float(beta1), float(beta2), float(eps), _prec(theta), st))


class MaskPenalties(torch.autograd.Function):
    """Sum over images of the mask penalties; the gradient comes from the same pass."""

    @staticmethod
    def forward(ctx, mask, weight_tv, weight_discrete):
        r = mask_penalties(mask, weight_tv, weight_discrete, grad=mask.requires_grad)
        ctx.save_for_backward(r["grad"] if r["grad"] is not None
                              else torch.empty(0, device=mask.device))
        return r["total"]

    @staticmethod
    def backward(ctx, g):
        (d,) = ctx.saved_tensors
        return d * g, None, None


class RelaxedResist(torch.autograd.Function):
    """Sum over images of the per-image relaxed-resist cost; the derivative is
    computed in the same pass as the cost and scaled by the incoming gradient.

    ``side``: a dict that, when given, receives the same pass's signed field
    (``"signed"``) and per-image pattern mismatch (``"mismatch"``), detached."""

    @staticmethod
    def forward(ctx, intensity, target, scale, level, steepness, sign, side=None):
        r = resist_objective(intensity, target, scale, level, steepness, sign,
                             grad=intensity.requires_grad, signed=side is not None,
                             mismatch=side is not None)
        if side is not None:
            side["signed"] = r["signed"]
            side["mismatch"] = r["mismatch"]
        ctx.save_for_backward(r["d_intensity"] if r["d_intensity"] is not None
                              else torch.empty(0, device=intensity.device))
        return r["total"]

    @staticmethod
    def backward(ctx, g):
        (d,) = ctx.saved_tensors
        return d * g, None, None, None, None, None, None


class MaskSigmoid(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, beta):
        m = mask_sigmoid(x, beta)
        ctx.beta = beta
        ctx.save_for_backward(m)
        return m

    @staticmethod
    def backward(ctx, g):
        (m,) = ctx.saved_tensors
        return mask_sigmoid_grad(g, m, ctx.beta), None
