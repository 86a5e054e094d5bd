"""Inverse-lithography gradient (opc/ilt.py:208-389 of the reference).

``AbbeIntensity`` is a ``torch.autograd.Function`` mask -> raw intensity whose
backward is the device adjoint (``lb_backward``): it recomputes every
per-system coherent field from the mask instead of storing them, and returns
d/dmask of <dI, I> for real masks.  It works for Abbe plans (scalar or
vector, several planes) and SOCS plans alike.

``socs_fields`` / ``socs_gradient`` keep the reference's two-call contract:
the fields are returned, then handed back with the sensitivity, and the
gradient is computed from the fields it is given (``lb_fields_backward``); with
the mask passed as well it recomputes them instead (the autograd path).
"""
from __future__ import annotations

import numpy as np
import torch

from .blur import blur
from .expose import _to_device
from .objective import (MaskPenalties, MaskSigmoid, RelaxedResist, mask_sigmoid_grad_theta,
                        resist_objective)
from .plan import ImagingPlan

__all__ = ["AbbeIntensity", "abbe_intensity", "socs_fields", "socs_gradient",
           "relaxed_resist_loss", "mask_of", "ILTObjective", "ilt_objective"]


class AbbeIntensity(torch.autograd.Function):
    """raw[B, P, N, N] = plan(mask[B, N, N]); backward by the recomputing adjoint.

    The masks' band spectra are computed once and kept for the backward (a few
    hundred KB per mask), so the step analyses each mask once.  A complex mask
    gets torch's complex gradient dJ/dRe m + i dJ/dIm m (its real part is the
    reference's real projection, opc/ilt.py:235-266)."""

    @staticmethod
    def forward(ctx, mask: torch.Tensor, plan: ImagingPlan):
        spec = plan.mask_spectrum(mask)
        raw, _ = plan.forward_spectrum(spec, with_peak=False)
        ctx.plan = plan
        ctx.squeeze = mask.dim() == 2
        ctx.dtype = mask.dtype
        ctx.complex = mask.is_complex()
        ctx.save_for_backward(spec)
        return raw

    @staticmethod
    def backward(ctx, grad_raw: torch.Tensor):
        (spec,) = ctx.saved_tensors
        g = ctx.plan.backward_spectrum(spec, grad_raw.contiguous(),
                                       complex_out=ctx.complex).to(ctx.dtype)
        return (g[0] if ctx.squeeze else g), None


def abbe_intensity(mask: torch.Tensor, plan: ImagingPlan) -> torch.Tensor:
    return AbbeIntensity.apply(mask, plan)


def mask_of(theta: torch.Tensor, beta: float, smooth_px: float = 0.0) -> torch.Tensor:
    """ilt.py:368-378: sigmoid(clip(beta * G(theta), +-60)), G the Gaussian of width
    ``smooth_px`` (the identity below 1e-3 px); the backward is ilt.py:381-389
    (d theta = G(d m * beta m (1 - m)))."""
    return MaskSigmoid.apply(blur(theta, smooth_px), beta)


def relaxed_resist_loss(raw: torch.Tensor, target: torch.Tensor, scale: float, level: float,
                        steepness: float, sign: float = 1.0, sigma_px: float = 0.0,
                        side: dict | None = None) -> torch.Tensor:
    """ilt.py:311-330 without penalties: mean (sigmoid(a s) - t)^2.

    a = steepness / |level| and s = sign * (G(raw * scale) - level), the
    reference's relaxed resist, G the PEB blur of width ``sigma_px`` (identity
    below 1e-3 px).  The mean runs over the last two (image) axes and is
    summed over any leading axes.  ``side`` (a dict) receives the signed field
    s and the per-image mismatch count #{(s > 0) != (t > 0.5)} of the same pass.
    """
    if sigma_px < 1e-3:
        return RelaxedResist.apply(raw, target, scale, level, steepness, sign, side)
    return RelaxedResist.apply(blur(raw * scale, sigma_px), target, 1.0, level, steepness, sign,
                               side)


class ILTObjective(torch.autograd.Function):
    """theta [B, N, N] -> the relaxed-resist cost (opc/ilt.py:311-330, without penalties
    and blurs), summed over images, of the image of the mask sigmoid(clip(beta theta,
    +-60)) (ilt.py:368-378): the value and gradient of
    ``relaxed_resist_loss(abbe_intensity(mask_of(theta, beta), plan), target, ...)``
    in one function.  The sigmoid runs inside the masks' band analysis; the objective is
    evaluated where the images are synthesised, which hands the backward the band spectrum
    of the intensity derivative (plans without that pass write the images and run the
    objective pass on them); the backward runs the recomputing adjoint from the kept band
    spectra, then the sigmoid derivative recomputed from theta and scaled by the incoming
    gradient, read on the device.  No mask, image or intensity-derivative array is
    written and no pass scales the derivative (the adjoint between them is linear)."""

    @staticmethod
    def forward(ctx, theta, plan, target, scale, level, steepness, beta, sign):
        spec = plan.mask_spectrum_sigmoid(theta, beta)
        ctx.plan, ctx.beta, ctx.dtype = plan, float(beta), theta.dtype
        ctx.squeeze = theta.dim() == 2
        t = target if target.dim() < 3 else target.reshape(target.shape[0], 1,
                                                           *target.shape[-2:])
        # the objective where the images are synthesised: the backward then starts from the
        # band spectrum of the intensity derivative (neither array is written)
        fused = plan.forward_objective(spec, t, scale, level, steepness, sign)
        ctx.band = fused is not None
        if ctx.band:
            total, _, sens = fused
            ctx.save_for_backward(spec, sens, theta)
            return total
        raw, _ = plan.forward_spectrum(spec, with_peak=False)
        r = resist_objective(raw, t, scale, level, steepness, sign, grad=True)
        ctx.save_for_backward(spec, r["d_intensity"], theta)
        return r["total"]

    @staticmethod
    def backward(ctx, g):
        spec, d, theta = ctx.saved_tensors
        th = theta[None] if ctx.squeeze else theta
        # the sigmoid derivative inside the mask-gradient synthesis where the band allows
        dth = (ctx.plan.backward_objective_theta(spec, d, th, ctx.beta, g) if ctx.band
               else None)
        if dth is None:
            dm = (ctx.plan.backward_sensitivity_band(spec, d) if ctx.band
                  else ctx.plan.backward_spectrum(spec, d))
            dth = mask_sigmoid_grad_theta(dm, th.to(dm.dtype), ctx.beta, g)
        return ((dth[0] if ctx.squeeze else dth).to(ctx.dtype), None, None, None, None, None,
                None, None)


def ilt_objective(theta: torch.Tensor, plan: ImagingPlan, target: torch.Tensor, scale: float,
                  level: float, steepness: float, beta: float = 8.0,
                  sign: float = 1.0) -> torch.Tensor:
    """The ILT objective of theta in one autograd function (``ILTObjective``)."""
    return ILTObjective.apply(theta, plan, target, scale, level, steepness, beta, sign)


def mask_penalties(mask: torch.Tensor, weight_tv: float = 0.0,
                   weight_discrete: float = 0.0) -> torch.Tensor:
    """ilt.py:339-355: weight_tv * mean |grad m|^2 (periodic differences) +
    weight_discrete * mean m (1 - m), per image, summed over leading axes; cost and
    gradient from one fused pass."""
    if weight_tv <= 0.0 and weight_discrete <= 0.0:
        return mask.new_zeros(())
    return MaskPenalties.apply(mask, max(weight_tv, 0.0), max(weight_discrete, 0.0))


def socs_fields(kernels, mask):
    """ilt.py:208-232: (fields (K, n, n) complex, raw intensity (n, n))."""
    n = kernels.grid.n_pixels
    if tuple(np.shape(mask)) != (n, n):
        raise ValueError(f"mask must be {(n, n)} for these kernels, got {np.shape(mask)}")
    m, was_np, _ = _to_device(mask, kernels.plan.precision)
    fields = kernels.plan.fields(m)[0]
    raw, _ = kernels.plan.forward(m, with_peak=False)
    if was_np:
        return fields.cpu().numpy(), raw[0, 0].double().cpu().numpy()
    return fields, raw[0, 0]


def socs_gradient(kernels, fields, sensitivity, mask=None):
    """ilt.py:235-266: d/dm of sum(sensitivity * I) for the raw SOCS intensity, real.

    Computed from the ``fields`` given (any (K, n, n) complex array, NumPy or a device
    tensor), as the reference does; when ``mask`` is passed the fields are recomputed
    from it instead (``fields`` is then not read)."""
    n = kernels.grid.n_pixels
    plan = kernels.plan
    if mask is not None:
        m, was_np, _ = _to_device(mask, plan.precision)
        g = torch.as_tensor(np.asarray(sensitivity) if was_np else sensitivity,
                            device=m.device, dtype=plan.real_dtype)
        out = plan.backward(m.reshape(1, n, n), g.reshape(1, 1, n, n))[0]
        return out.double().cpu().numpy() if was_np else out
    was_np = not isinstance(fields, torch.Tensor)
    dev = torch.device("cuda", torch.cuda.current_device())
    f = torch.as_tensor(np.asarray(fields), device=dev) if was_np else fields
    if not f.is_complex():
        raise TypeError("fields must be complex (the coherent fields of socs_fields)")
    g = torch.as_tensor(np.asarray(sensitivity, dtype=np.float64) if was_np else sensitivity,
                        device=f.device)
    if tuple(g.shape) != (n, n):
        raise ValueError(f"sensitivity must be {(n, n)} for these kernels, got "
                         f"{tuple(g.shape)}")
    out = plan.fields_backward(f, g)[0]
    return out.double().cpu().numpy() if was_np else out
