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

    def set_shape(self, adjoint: bool, threads: int, splits: int) -> None:
        """Coherent-system kernel shape of the batch's forward or adjoint launch."""
        _lib.check(self._lib.lb_batch_set_shape(self._h, 1 if adjoint else 0, int(threads),
                                                int(splits)))

    @property
    def groups(self) -> list[list[int]]:
        """Clips with the same band geometry, whose band passes the library runs as one
        pass when their buffers are consecutive in memory (see ``grouped_empty``)."""
        keys = {}
        for c, p in enumerate(self.plans):
            i = p.info
            keys.setdefault((p.n, len(p.planes), i["band"], i["m"], i["n_systems"]), []).append(c)
        return list(keys.values())

    def grouped_empty(self, per_clip_shape, device) -> list[torch.Tensor]:
        """Per-clip tensors (shape ``per_clip_shape(plan)``) carved consecutively out of one
        allocation per clip group, so batched passes run on them in place."""
        out = [None] * len(self.plans)
        for g in self.groups:
            shape = tuple(per_clip_shape(self.plans[g[0]]))
            block = torch.empty((len(g),) + shape, dtype=self.real_dtype, device=device)
            for k, c in enumerate(g):
                out[c] = block[k]
        return out

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

    def backward_into(self, masks, d_images, d_masks, stream=None,
                      after_forward: bool = False) -> None:
        """``after_forward``: the last ``forward_into`` imaged these same masks (and its
        workspace is untouched since), so their band spectra are reused."""
        st = stream if stream is not None else torch.cuda.current_stream().cuda_stream
        ws = self.workspace(True, d_images[0].device)
        if after_forward:
            fws = self.workspace(False, d_images[0].device)
            _lib.check(self._lib.lb_batch_backward_reuse(self._h, _ptrs(d_images),
                                                         _ptrs(d_masks), ws.data_ptr(),
                                                         ws.numel(), fws.data_ptr(), st))
            return
        _lib.check(self._lib.lb_batch_backward(self._h, _ptrs(masks), _ptrs(d_images),
                                               _ptrs(d_masks), ws.data_ptr(), ws.numel(), st))

    def forward(self, masks):
        masks = self._check(masks)
        images = self.grouped_empty(lambda p: (len(p.planes), p.n, p.n), masks[0].device)
        self.forward_into(masks, images)
        return images

    def backward(self, masks, d_images):
        masks = self._check(masks)
        d_images = [g.to(self.real_dtype).contiguous() for g in d_images]
        d_masks = self.grouped_empty(lambda p: (p.n, p.n), masks[0].device)
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
    Clips of one shape group (``ImagingBatch.groups``) share consecutive buffers:
    their mask transforms, objective and band passes run once per group.
    ``costs`` holds the loss of each group (of each clip where a group's clips
    have different normalisations).
    """

    def __init__(self, batch: ImagingBatch, thetas, targets, beta=8.0, level=0.3,
                 steepness=50.0):
        self.batch = batch
        self.beta, self.level, self.steep = beta, level, steepness
        dev = thetas[0].device
        dt = batch.real_dtype
        plans = batch.plans
        self.theta = batch.grouped_empty(lambda p: (p.n, p.n), dev)
        for th, t in zip(self.theta, thetas):
            th.copy_(t.to(dt))
        self.mask = batch.grouped_empty(lambda p: (p.n, p.n), dev)
        self.raw = batch.grouped_empty(lambda p: (len(p.planes), p.n, p.n), dev)
        self.d_raw = batch.grouped_empty(lambda p: (len(p.planes), p.n, p.n), dev)
        self.d_mask = batch.grouped_empty(lambda p: (p.n, p.n), dev)
        self.grad = batch.grouped_empty(lambda p: (p.n, p.n), dev)
        # the target of every plane of the clip, as the objective reads it
        self.target_planes = batch.grouped_empty(lambda p: (len(p.planes), p.n, p.n), dev)
        for tp, t in zip(self.target_planes, targets):
            tp.copy_(t.to(dt).expand(tp.shape))
        scale = [1.0 / float(p.clear[0]) if p.clear[0] > 0 else 1.0 for p in plans]

        def whole(views, g):  # the group's consecutive block behind its clips' views
            v0 = views[g[0]]
            return v0.as_strided((len(g),) + tuple(v0.shape), (v0.numel(),) + v0.stride())

        # units of work: a whole group where its clips share a normalisation, else clips
        self.units = []
        for g in batch.groups:
            if all(scale[c] == scale[g[0]] for c in g):
                self.units.append((whole(self.theta, g), whole(self.mask, g), whole(self.raw, g),
                                   whole(self.target_planes, g), whole(self.d_raw, g),
                                   whole(self.d_mask, g), whole(self.grad, g), scale[g[0]]))
            else:
                for c in g:
                    self.units.append((self.theta[c], self.mask[c], self.raw[c],
                                       self.target_planes[c], self.d_raw[c], self.d_mask[c],
                                       self.grad[c], scale[c]))
        self.loss = torch.zeros((), dtype=dt, device=dev)
        self.costs = torch.zeros(len(self.units), dtype=dt, device=dev)
        self.obj_ws = [torch.empty(max(objective.workspace_bytes(u[2].numel() // u[2].shape[-1] ** 2,
                                                                 u[2].shape[-1] ** 2), 256),
                                   dtype=torch.uint8, device=dev) for u in self.units]
        batch.workspace(False, dev)
        batch.workspace(True, dev)
        # the units' elementwise passes run concurrently, one stream each (default
        # priority), forked from and joined to the caller's stream
        self.side = [torch.cuda.Stream(device=dev) for _ in self.units] if len(self.units) > 1 \
            else []

    def _each(self, fn) -> None:
        if not self.side:
            for u in range(len(self.units)):
                fn(u)
            return
        cur = torch.cuda.current_stream()
        for u, s in enumerate(self.side):
            s.wait_stream(cur)
            with torch.cuda.stream(s):
                fn(u)
        for s in self.side:
            cur.wait_stream(s)

    def _sigmoid(self, u: int) -> None:
        th, m, *_ = self.units[u]
        objective.mask_sigmoid(th, self.beta, out=m)

    def _objective(self, u: int) -> None:
        _, _, raw, t, d, _, _, s = self.units[u]
        objective.resist_objective_into(raw, t, s, self.level, self.steep, 1.0, None,
                                        self.costs[u], d, None, None, self.obj_ws[u])

    def _sigmoid_grad(self, u: int) -> None:
        _, m, _, _, _, dm, g, _ = self.units[u]
        objective.mask_sigmoid_grad(dm, m, self.beta, out=g)

    def run(self) -> None:
        self._each(self._sigmoid)
        self.batch.forward_into(self.mask, self.raw)
        self._each(self._objective)
        torch.sum(self.costs, 0, out=self.loss)
        self.batch.backward_into(self.mask, self.d_raw, self.d_mask, after_forward=True)
        self._each(self._sigmoid_grad)
