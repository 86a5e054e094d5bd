"""Uneven-clip batches: many clips of different sizes and optics in one launch.

``ImagingBatch`` binds one plan per clip (all on a common coarse grid side)
to the library's batch object: each clip's band transforms run at its own
size, and every clip's coherent systems run in one persistent launch of the
coherent-system kernel.  ``BatchIntensity`` is the autograd form, and
``ILTStep`` a fixed-shape inverse-lithography step (forward, relaxed-resist
loss, adjoint) that can be captured into a HIP graph and replayed.
"""
from __future__ import annotations

import ctypes as C

import torch

from . import _lib, objective
from .plan import ImagingPlan


def _ptrs(tensors):
    return (C.c_void_p * len(tensors))(*[t.data_ptr() for t in tensors])


class ImagingBatch:
    """Clips = list of (optics, grid[, planes]); one mask per clip."""

    def __init__(self, clips, precision: int = 32, coarse_m: int = 0):
        lib = _lib.load()
        if coarse_m == 0:
            # the smallest grid every clip can use
            probe = [ImagingPlan(c[0], c[1], c[2] if len(c) > 2 else None, precision=precision)
                     for c in clips]
            coarse_m = max(p.info["m"] for p in probe)
        self.plans = [ImagingPlan(c[0], c[1], c[2] if len(c) > 2 else None, precision=precision,
                                  coarse_m=coarse_m) for c in clips]
        self.precision = precision
        self.sizes = [p.n for p in self.plans]
        handles = (C.c_void_p * len(self.plans))(*[p._h.value for p in self.plans])
        h = C.c_void_p()
        _lib.check(lib.lb_batch_create(handles, len(self.plans), C.byref(h)))
        self._h = h
        self._lib = lib
        self._ws = {}

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                self._lib.lb_batch_destroy(h)
            except Exception:  # noqa: BLE001 - interpreter shutdown
                pass
            self._h = None

    def __len__(self):
        return len(self.plans)

    @property
    def real_dtype(self):
        return self.plans[0].real_dtype

    def workspace(self, backward: bool, device) -> torch.Tensor:
        n = int(self._lib.lb_batch_workspace_bytes(self._h, 1 if backward else 0))
        ws = self._ws.get(backward)
        if ws is None or ws.numel() < n or ws.device != device:
            ws = torch.empty(max(n, 256), dtype=torch.uint8, device=device)
            self._ws[backward] = ws
        return ws

    def _check(self, masks):
        if len(masks) != len(self.plans):
            raise ValueError(f"need {len(self.plans)} masks, got {len(masks)}")
        out = []
        for m, n in zip(masks, self.sizes):
            if m.dim() != 2 or tuple(m.shape) != (n, n):
                raise ValueError(f"mask must be {(n, n)}, got {tuple(m.shape)}")
            if m.device.type != "cuda":
                raise TypeError("masks must be on the HIP device")
            out.append(m.to(self.real_dtype).contiguous())
        return out

    def forward_into(self, masks, images, stream=None) -> None:
        """Raw images into preallocated tensors ([P, N_c, N_c] each); no allocation."""
        st = stream if stream is not None else torch.cuda.current_stream().cuda_stream
        ws = self.workspace(False, masks[0].device)
        _lib.check(self._lib.lb_batch_forward(self._h, _ptrs(masks), _ptrs(images), None,
                                              ws.data_ptr(), ws.numel(), st))

    def backward_into(self, masks, d_images, d_masks, stream=None) -> None:
        st = stream if stream is not None else torch.cuda.current_stream().cuda_stream
        ws = self.workspace(True, masks[0].device)
        _lib.check(self._lib.lb_batch_backward(self._h, _ptrs(masks), _ptrs(d_images),
                                               _ptrs(d_masks), ws.data_ptr(), ws.numel(), st))

    def forward(self, masks):
        masks = self._check(masks)
        images = [torch.empty((len(p.planes), p.n, p.n), dtype=self.real_dtype,
                              device=m.device) for p, m in zip(self.plans, masks)]
        self.forward_into(masks, images)
        return images

    def backward(self, masks, d_images):
        masks = self._check(masks)
        d_images = [g.to(self.real_dtype).contiguous() for g in d_images]
        d_masks = [torch.empty_like(m) for m in masks]
        self.backward_into(masks, d_images, d_masks)
        return d_masks


class BatchIntensity(torch.autograd.Function):
    """(mask_0, ..., mask_k) -> (raw_0, ..., raw_k) through an ImagingBatch."""

    @staticmethod
    def forward(ctx, batch: ImagingBatch, *masks):
        ctx.batch = batch
        ctx.save_for_backward(*masks)
        return tuple(batch.forward(list(masks)))

    @staticmethod
    def backward(ctx, *grads):
        masks = ctx.saved_tensors
        g = ctx.batch.backward(list(masks), [x.contiguous() for x in grads])
        return (None, *[d.to(m.dtype) for d, m in zip(g, masks)])


class ILTStep:
    """One inverse-lithography gradient step for every clip of a batch.

    theta_c -> m_c = sigmoid(beta theta_c) -> raw_c (batch forward) ->
    loss = sum_c mean (sigmoid(a (raw_c scale_c - level)) - t_c)^2, with the
    reference's relaxed resist (a = steepness / |level|, opc/ilt.py:311-337),
    and d loss / d theta_c by the explicit chain rule and the batch adjoint.
    All buffers are preallocated, so ``run`` can be captured into a graph.
    """

    def __init__(self, batch: ImagingBatch, thetas, targets, beta=8.0, level=0.3,
                 steepness=50.0):
        self.batch = batch
        self.beta, self.level, self.steep = beta, level, steepness
        self.theta = [t.to(batch.real_dtype).contiguous() for t in thetas]
        self.target = [t.to(batch.real_dtype).contiguous() for t in targets]
        # the target of every plane of the clip, as the objective reads it
        self.target_planes = [t.expand(len(p.planes), *t.shape[-2:]).contiguous()
                              for p, t in zip(batch.plans, self.target)]
        self.scale = [1.0 / float(p.clear[0]) if p.clear[0] > 0 else 1.0 for p in batch.plans]
        self.mask = [torch.empty_like(t) for t in self.theta]
        self.raw = [torch.empty((len(p.planes), p.n, p.n), dtype=batch.real_dtype,
                                device=t.device) for p, t in zip(batch.plans, self.theta)]
        self.d_raw = [torch.empty_like(r) for r in self.raw]
        self.d_mask = [torch.empty_like(t) for t in self.theta]
        self.grad = [torch.empty_like(t) for t in self.theta]
        self.loss = torch.zeros((), dtype=batch.real_dtype, device=self.theta[0].device)
        self.costs = torch.zeros(len(self.theta), dtype=batch.real_dtype,
                                 device=self.theta[0].device)
        self.obj_ws = [torch.empty(max(objective.workspace_bytes(len(r), r.shape[-1] ** 2), 256),
                                   dtype=torch.uint8, device=r.device) for r in self.raw]
        batch.workspace(False, self.theta[0].device)
        batch.workspace(True, self.theta[0].device)

    def run(self) -> None:
        for th, m in zip(self.theta, self.mask):
            objective.mask_sigmoid(th, self.beta, out=m)
        self.batch.forward_into(self.mask, self.raw)
        for c, (raw, t, s, d, ws) in enumerate(zip(self.raw, self.target_planes, self.scale,
                                                    self.d_raw, self.obj_ws)):
            objective.resist_objective_into(raw, t, s, self.level, self.steep, 1.0, None,
                                            self.costs[c], d, None, None, ws)
        torch.sum(self.costs, 0, out=self.loss)
        self.batch.backward_into(self.mask, self.d_raw, self.d_mask)
        for g, dm, m in zip(self.grad, self.d_mask, self.mask):
            objective.mask_sigmoid_grad(dm, m, self.beta, out=g)
