"""Imaging plans: one optical set-up bound to the device library.

An ``ImagingPlan`` owns a C-ABI plan (the coherent systems of one optics /
grid / plane set, their transfer tables on the device, and the kernel shape)
and a growable device workspace.  ``forward`` and ``backward`` take and return
torch tensors on the current device and run on torch's current stream.
"""
from __future__ import annotations

import ctypes as C
import json
import logging
import threading
from collections import OrderedDict
from dataclasses import asdict, is_dataclass

import numpy as np
import torch

from . import _lib
from .source import SourcePoints, source_points

_DTYPES = {32: (torch.float32, torch.complex64), 64: (torch.float64, torch.complex128)}
logger = logging.getLogger(__name__)


def _image_index(optics) -> float:
    n_image = getattr(optics, "n_image", None)
    return optics.n_immersion if n_image is None else n_image


def _dptr(a: np.ndarray, ctype):
    return a.ctypes.data_as(C.POINTER(ctype))


class ImagingPlan:
    """A device plan for the Abbe sum (``kind="abbe"``) or its SOCS form (``kind="socs"``)."""

    def __init__(self, optics, grid, planes=None, precision: int = 32, kind: str = "abbe",
                 tol: float = 1e-8, max_kernels: int | None = None, src: SourcePoints | None = None,
                 coarse_m: int = 0, chunk: int = 0):
        lib = _lib.load()
        if precision not in (32, 64):
            raise ValueError("precision must be 32 or 64")
        if not torch.cuda.is_available():
            raise RuntimeError("the imaging library needs a HIP device (torch.cuda is unavailable)")
        model = optics.imaging_model.lower()
        if model not in ("scalar", "vector"):
            raise ValueError(f"imaging_model must be 'scalar' or 'vector', got "
                             f"'{optics.imaging_model}'")
        mask_model = getattr(optics, "mask_model", "thin")
        if mask_model not in ("thin", "multilayer"):
            raise NotImplementedError(
                f"mask_model={mask_model!r}: the thin and multilayer mask spectra are "
                "implemented")
        if mask_model == "multilayer" and kind != "abbe":
            raise ValueError("SOCS kernels apply to the scalar thin-mask process only")
        if planes is None:
            planes = [(optics.defocus, _image_index(optics))]
        self.planes = [(float(d), float(n)) for d, n in planes]
        self.precision = precision
        self.kind = kind
        self.n = int(grid.n_pixels)
        self.src = src if src is not None else source_points(optics)
        zern = {int(k): float(v) for k, v in (optics.zernike_coeffs or {}).items()}
        for j in zern:
            if not 1 <= j <= 21:  # expose/pupil.py:128-130: the term contributes zero
                logger.warning("Zernike Noll index %d not implemented; returning 0.", j)
        pol = optics.polarisation.lower()
        if pol not in _lib.POLARISATION_CODES:
            raise ValueError(f"polarisation must be one of {list(_lib.POLARISATION_CODES)}, "
                             f"got '{optics.polarisation}'")
        self._keep = []  # arrays the descriptor points into

        def arr(x, dt=np.float64):
            a = np.ascontiguousarray(np.asarray(x, dtype=dt))
            self._keep.append(a)
            return a

        defocus = arr([p[0] for p in self.planes])
        n_image = arr([p[1] for p in self.planes])
        zj = arr(list(zern.keys()) or [0], np.int32)
        zc = arr(list(zern.values()) or [0.0])
        w, fx, fy, phi = (arr(self.src.weight), arr(self.src.fx), arr(self.src.fy),
                          arr(self.src.phi))
        d = _lib.ImagingDesc()
        d.n = self.n
        d.pixel_size = float(grid.pixel_size)
        d.wavelength = float(optics.wavelength)
        d.na = float(optics.NA)
        d.vector = 1 if model == "vector" else 0
        d.polarisation = _lib.POLARISATION_CODES[pol]
        d.obliquity = 1 if optics.obliquity else 0
        d.exact_defocus = 1 if optics.exact_defocus else 0
        d.n_planes = len(self.planes)
        d.defocus = _dptr(defocus, C.c_double)
        d.n_image = _dptr(n_image, C.c_double)
        d.n_zernike = len(zern)
        d.zernike_j = _dptr(zj, C.c_int32)
        d.zernike_c = _dptr(zc, C.c_double)
        d.n_source = len(self.src)
        d.src_weight = _dptr(w, C.c_double)
        d.src_fx = _dptr(fx, C.c_double)
        d.src_fy = _dptr(fy, C.c_double)
        d.src_phi = _dptr(phi, C.c_double)
        d.precision = precision
        d.coarse_m = int(coarse_m)
        d.chunk = int(chunk)
        if mask_model == "multilayer":
            # provider.py:529-545: the stack named by the optics, which must be reflective
            from . import m3d
            stack = m3d.stack_for(optics)
            if not stack.is_reflective:
                raise ValueError(
                    "mask_model='multilayer' describes the Bragg mirror under an EUV mask, "
                    f"but the stack regime is '{stack.regime}'. Use mask_model='thin' or "
                    "'fdtd' for a transmissive mask.")
            theta, mag, phase = m3d.reflection_table(stack, optics.wavelength)
            sines = [m3d.mask_side_sin_theta(fx_, fy_, optics)
                     for fx_, fy_ in zip(self.src.fx, self.src.fy)]
            d.n_refl = len(theta)
            d.refl_theta = _dptr(arr(theta), C.c_double)
            d.refl_mag = _dptr(arr(mag), C.c_double)
            d.refl_phase = _dptr(arr(phase), C.c_double)
            d.refl_doff = float(optics.wavelength / optics.reduction)
            d.src_sin_x = _dptr(arr([v[0] for v in sines]), C.c_double)
            d.src_sin_y = _dptr(arr([v[1] for v in sines]), C.c_double)
        handle = C.c_void_p()
        if kind == "abbe":
            _lib.check(lib.lb_abbe_plan_create(C.byref(d), C.byref(handle)))
            if coarse_m == 0:
                handle = self._tuned_grid(lib, d, handle, precision)
        elif kind == "socs":
            mk = -1 if max_kernels is None else int(max_kernels)
            _lib.check(lib.lb_socs_plan_create(C.byref(d), float(tol), mk, C.byref(handle)))
        else:
            raise ValueError(f"unknown plan kind {kind!r}")
        self._h = handle
        self._lib = lib
        self._ws = None
        self._retired = []  # earlier workspaces, which graphs may still address
        self._lock = threading.Lock()
        # batch size whose tuned kernel shapes are selected; _pinned: shapes set by
        # hand (autotune.set_shape), which the automatic selection leaves alone
        self._shape_for = None
        self._pinned = False
        self._band = None  # mask band side (set on first use; fixed for the plan)
        clear = np.zeros(len(self.planes))
        _lib.check(lib.lb_plan_clear(self._h, _dptr(clear, C.c_double)))
        self.clear = clear
        if kind == "socs":
            K = self.info["n_kernels"]
            wts = np.zeros(max(K, 1))
            e = C.c_double()
            nb = C.c_int32()
            ns = C.c_int32()
            _lib.check(lib.lb_socs_weights(self._h, _dptr(wts, C.c_double), C.byref(e),
                                           C.byref(nb), C.byref(ns)))
            self.weights = wts[:K].copy()
            self.energy_kept = float(e.value)
            self.n_band = int(nb.value)
            self.n_source = int(ns.value)

    def _tuned_grid(self, lib, d, handle, precision: int):
        """The plan rebuilt on the coarse grid ``autotune.tune_grid`` recorded for this
        geometry (precision, image side, pupil box), when that differs from the automatic
        one; the automatic plan where nothing is recorded or the rebuild fails."""
        from . import autotune
        info = _lib.PlanInfo()
        _lib.check(lib.lb_plan_info_get(handle, C.byref(info)))
        m = autotune.recorded_grid(precision, self.n, int(info.box))
        if not m or m == int(info.m):
            return handle
        d.coarse_m = m
        tuned = C.c_void_p()
        if lib.lb_abbe_plan_create(C.byref(d), C.byref(tuned)) != _lib.LB_OK:
            d.coarse_m = 0
            return handle
        lib.lb_plan_destroy(handle)
        return tuned

    # ------------------------------------------------------------------
    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                self._lib.lb_plan_destroy(h)
            except Exception:  # noqa: BLE001 - interpreter shutdown
                pass
            self._h = None

    @property
    def real_dtype(self):
        return _DTYPES[self.precision][0]

    @property
    def complex_dtype(self):
        return _DTYPES[self.precision][1]

    def _workspace(self, nbytes: int, device) -> torch.Tensor:
        """The plan's device workspace, grown when a call needs more.  A replaced
        workspace is kept alive with the plan, not released: a graph captured on it
        still writes there when it replays, after later calls moved to a larger one."""
        if self._ws is None or self._ws.numel() < nbytes or self._ws.device != device:
            if self._ws is not None:
                self._retired.append(self._ws)
            self._ws = torch.empty(max(nbytes, 256), dtype=torch.uint8, device=device)
        return self._ws

    @property
    def info(self) -> dict:
        """Geometry and the current coherent-system kernel shape (lb_plan_info_get)."""
        info = _lib.PlanInfo()
        _lib.check(self._lib.lb_plan_info_get(self._h, C.byref(info)))
        return {k: getattr(info, k) for k, _ in _lib.PlanInfo._fields_}

    def _use_tuning(self, n_masks: int) -> None:
        """Select the autotuned kernel shapes recorded for this batch size (the plan's
        defaults where the table has none), whenever the batch size changes."""
        if self._pinned or self._shape_for == n_masks:
            return
        from . import autotune
        autotune.apply(self, n_masks)
        self._shape_for = n_masks

    def work_items(self, n_masks: int, adjoint: bool = False) -> tuple[int, int]:
        """(systems per work item, work items) of a call of ``n_masks`` masks with the
        shapes currently selected (lb_plan_work_items)."""
        chunk, items = C.c_int32(), C.c_int32()
        _lib.check(self._lib.lb_plan_work_items(self._h, int(n_masks), 1 if adjoint else 0,
                                                C.byref(chunk), C.byref(items)))
        return int(chunk.value), int(items.value)

    def workspace_bytes(self, n_masks: int, backward: bool = False) -> int:
        return int(self._lib.lb_workspace_bytes(self._h, int(n_masks), 1 if backward else 0))

    def _check_masks(self, masks: torch.Tensor) -> tuple[torch.Tensor, int]:
        if not isinstance(masks, torch.Tensor) or masks.device.type != "cuda":
            raise TypeError("masks must be a torch tensor on the HIP device")
        if masks.dim() == 2:
            masks = masks.unsqueeze(0)
        if masks.dim() != 3 or masks.shape[-1] != self.n or masks.shape[-2] != self.n:
            raise ValueError(f"mask must be {(self.n, self.n)} for this plan, got "
                             f"{tuple(masks.shape[-2:])}")
        if masks.is_complex():
            masks = torch.view_as_real(masks.to(self.complex_dtype).contiguous())
            return masks, 1
        return masks.to(self.real_dtype).contiguous(), 0

    def forward(self, masks: torch.Tensor, with_peak: bool = True):
        """Raw images [B, P, N, N] (and peaks [B, P]) of masks [B, N, N] or [N, N]."""
        m, cplx = self._check_masks(masks)
        B = m.shape[0]
        P = len(self.planes)
        out = torch.empty((B, P, self.n, self.n), dtype=self.real_dtype, device=m.device)
        peak = torch.empty((B, P), dtype=self.real_dtype, device=m.device) if with_peak else None
        with self._lock:
            self._use_tuning(B)
            nb = self.workspace_bytes(B, False)
            ws = self._workspace(nb, m.device)
            st = torch.cuda.current_stream(m.device).cuda_stream
            _lib.check(self._lib.lb_forward(self._h, m.data_ptr(), cplx, B, out.data_ptr(),
                                            peak.data_ptr() if peak is not None else None,
                                            ws.data_ptr(), ws.numel(), st))
        return out, peak

    def mask_spectrum(self, masks: torch.Tensor) -> torch.Tensor:
        """Band spectra [B, Wb, Wb] (complex) of masks [B, N, N] or [N, N], the input of
        ``forward_spectrum`` / ``backward_spectrum``."""
        m, cplx = self._check_masks(masks)
        B = m.shape[0]
        wb = self._band_side()
        spec = torch.empty((B, wb, wb), dtype=self.complex_dtype, device=m.device)
        with self._lock:
            ws = self._workspace(self.workspace_bytes(B, False), m.device)
            st = torch.cuda.current_stream(m.device).cuda_stream
            _lib.check(self._lib.lb_mask_spectrum(self._h, m.data_ptr(), cplx, B,
                                                  spec.data_ptr(), ws.data_ptr(), ws.numel(),
                                                  st))
        return spec

    def mask_spectrum_sigmoid(self, theta: torch.Tensor, beta: float) -> torch.Tensor:
        """``mask_spectrum`` of the ILT mask sigmoid(clip(beta theta, +-60)) of real theta
        [B, N, N] (opc/ilt.py:368-378): the sigmoid is applied inside the band analysis, no
        mask array is written.  Grid sides without the fused analysis apply the sigmoid
        first (``lb_mask_sigmoid``) and analyse the mask."""
        from .objective import mask_sigmoid
        th, cplx = self._check_masks(theta)
        if cplx:
            raise TypeError("theta must be real")
        B = th.shape[0]
        wb = self._band_side()
        spec = torch.empty((B, wb, wb), dtype=self.complex_dtype, device=th.device)
        with self._lock:
            ws = self._workspace(self.workspace_bytes(B, False), th.device)
            st = torch.cuda.current_stream(th.device).cuda_stream
            rc = self._lib.lb_mask_spectrum_sigmoid(self._h, th.data_ptr(), B, float(beta),
                                                    spec.data_ptr(), ws.data_ptr(), ws.numel(),
                                                    st)
            if rc == _lib.LB_OK:
                return spec
            if rc != _lib.LB_ERR_UNSUPPORTED:
                _lib.check(rc)
        return self.mask_spectrum(mask_sigmoid(th, beta))

    def _band_side(self) -> int:
        if self._band is None:
            self._band = int(self.info["band"])
            assert self._band ** 2 * (16 if self.precision == 64 else 8) == int(
                self._lib.lb_mask_spectrum_bytes(self._h, 1))
        return self._band

    def _check_spectrum(self, spec: torch.Tensor) -> int:
        wb = self._band_side()
        if (spec.dtype != self.complex_dtype or spec.dim() != 3 or spec.shape[1:] != (wb, wb)
                or not spec.is_contiguous() or spec.device.type != "cuda"):
            raise ValueError(f"spectrum must be a contiguous {self.complex_dtype} device tensor "
                             f"[B, {wb}, {wb}] from mask_spectrum")
        return spec.shape[0]

    def forward_spectrum(self, spec: torch.Tensor, with_peak: bool = True):
        """``forward`` from band spectra (``mask_spectrum``): images [B, P, N, N], peaks."""
        B = self._check_spectrum(spec)
        P = len(self.planes)
        out = torch.empty((B, P, self.n, self.n), dtype=self.real_dtype, device=spec.device)
        peak = torch.empty((B, P), dtype=self.real_dtype, device=spec.device) if with_peak else None
        with self._lock:
            self._use_tuning(B)
            ws = self._workspace(self.workspace_bytes(B, False), spec.device)
            st = torch.cuda.current_stream(spec.device).cuda_stream
            _lib.check(self._lib.lb_forward_spectrum(
                self._h, spec.data_ptr(), B, out.data_ptr(),
                peak.data_ptr() if peak is not None else None, ws.data_ptr(), ws.numel(), st))
        return out, peak

    def backward_spectrum(self, spec: torch.Tensor, d_images: torch.Tensor,
                          complex_out: bool = False) -> torch.Tensor:
        """``backward`` from the band spectra (``mask_spectrum``): d/dmask of
        sum(d_images * I), real [B, N, N]; ``complex_out`` gives dJ/dRe m + i dJ/dIm m
        (complex [B, N, N]), the gradient of a complex mask."""
        B = self._check_spectrum(spec)
        P = len(self.planes)
        g = d_images.to(self.real_dtype).reshape(B, P, self.n, self.n).contiguous()
        dt = self.complex_dtype if complex_out else self.real_dtype
        out = torch.empty((B, self.n, self.n), dtype=dt, device=spec.device)
        fn = (self._lib.lb_backward_spectrum_complex if complex_out
              else self._lib.lb_backward_spectrum)
        with self._lock:
            self._use_tuning(B)
            ws = self._workspace(self.workspace_bytes(B, True), spec.device)
            st = torch.cuda.current_stream(spec.device).cuda_stream
            _lib.check(fn(self._h, spec.data_ptr(), B, g.data_ptr(), out.data_ptr(),
                          ws.data_ptr(), ws.numel(), st))
        return out

    def forward_objective(self, spec: torch.Tensor, target: torch.Tensor, scale: float,
                          level: float, steepness: float, sign: float = 1.0):
        """The relaxed-resist objective (``objective.resist_objective``) of the images of band
        spectra ``spec`` against ``target`` ([N, N], [B, N, N] or [B, P, N, N]), evaluated where the
        images are synthesised (lb_forward_objective): (total (0-d), cost [B, P],
        sens_band), sens_band being the band spectrum of the intensity derivative that
        ``backward_sensitivity_band`` takes.  None when the plan has no such pass (an image
        grid used directly, or an image band wider than N / 8)."""
        B = self._check_spectrum(spec)
        P, n = len(self.planes), self.n
        t = target.to(self.real_dtype)
        t = t.reshape(-1 if t.dim() > 2 else 1, 1 if t.dim() < 4 else t.shape[1], n, n)
        t = t.expand(B, P, n, n).contiguous()
        cost = torch.empty((B, P), dtype=self.real_dtype, device=spec.device)
        total = torch.empty((), dtype=self.real_dtype, device=spec.device)
        with self._lock:
            self._use_tuning(B)
            nb = int(self._lib.lb_forward_objective_workspace_bytes(self._h, B))
            sb = int(self._lib.lb_sensitivity_band_bytes(self._h, B))
            if nb == 0 or sb == 0 or t.data_ptr() % 16:
                return None
            wb = int(round((sb // (B * P * spec.element_size())) ** 0.5))
            sens = torch.empty((B, P, wb, wb), dtype=self.complex_dtype, device=spec.device)
            ws = self._workspace(nb, spec.device)
            st = torch.cuda.current_stream(spec.device).cuda_stream
            rc = self._lib.lb_forward_objective(
                self._h, spec.data_ptr(), B, t.data_ptr(), float(scale), float(level),
                float(steepness), float(sign), cost.data_ptr(), total.data_ptr(), None,
                sens.data_ptr(), None, None, ws.data_ptr(), ws.numel(), st)
            if rc == _lib.LB_ERR_UNSUPPORTED:
                return None
            _lib.check(rc)
        return total, cost, sens

    def backward_sensitivity_band(self, spec: torch.Tensor, sens: torch.Tensor) -> torch.Tensor:
        """``backward_spectrum`` from the band spectrum of the intensity derivative
        (``forward_objective``'s sens_band) instead of the derivative itself: real [B, N, N]."""
        B = self._check_spectrum(spec)
        if sens.dtype != self.complex_dtype or not sens.is_contiguous() or sens.shape[0] != B:
            raise ValueError("sens must be forward_objective's sensitivity band spectrum")
        out = torch.empty((B, self.n, self.n), dtype=self.real_dtype, device=spec.device)
        with self._lock:
            self._use_tuning(B)
            ws = self._workspace(self.workspace_bytes(B, True), spec.device)
            st = torch.cuda.current_stream(spec.device).cuda_stream
            _lib.check(self._lib.lb_backward_sensitivity_band(
                self._h, spec.data_ptr(), B, sens.data_ptr(), out.data_ptr(), ws.data_ptr(),
                ws.numel(), st))
        return out

    def backward_objective_theta(self, spec: torch.Tensor, sens: torch.Tensor,
                                 theta: torch.Tensor, beta: float, g: torch.Tensor | None = None):
        """``backward_sensitivity_band`` followed by ``objective.mask_sigmoid_grad_theta`` in
        one pass (lb_backward_objective_theta): d/dtheta [B, N, N] of the ILT parameter
        whose mask sigmoid(clip(beta theta, +-60)) gave ``spec``, scaled by the device
        scalar ``g``.  None when the plan has no such pass."""
        B = self._check_spectrum(spec)
        th = theta.to(self.real_dtype).reshape(B, self.n, self.n).contiguous()
        gs = None if g is None else g.to(self.real_dtype).reshape(()).contiguous()
        out = torch.empty((B, self.n, self.n), dtype=self.real_dtype, device=spec.device)
        with self._lock:
            self._use_tuning(B)
            ws = self._workspace(self.workspace_bytes(B, True), spec.device)
            st = torch.cuda.current_stream(spec.device).cuda_stream
            rc = self._lib.lb_backward_objective_theta(
                self._h, spec.data_ptr(), B, sens.data_ptr(), th.data_ptr(), float(beta),
                gs.data_ptr() if gs is not None else None, out.data_ptr(), ws.data_ptr(),
                ws.numel(), st)
            if rc == _lib.LB_ERR_UNSUPPORTED:
                return None
            _lib.check(rc)
        return out

    def fields_backward(self, fields: torch.Tensor, sensitivity: torch.Tensor,
                        complex_out: bool = False) -> torch.Tensor:
        """SOCS plans: socs_gradient on the fields given, [K, N, N] or [B, K, N, N] complex,
        with the sensitivity [N, N] or [B, N, N]: real [B, N, N] (or the complex adjoint)."""
        K = self.info["n_systems"]
        f = fields.to(self.complex_dtype)
        if f.dim() == 3:
            f = f.unsqueeze(0)
        if f.dim() != 4 or tuple(f.shape[1:]) != (K, self.n, self.n):
            raise ValueError(f"fields must be ({K}, {self.n}, {self.n}) for these kernels, got "
                             f"{tuple(fields.shape)}")
        B = f.shape[0]
        g = sensitivity.to(device=f.device, dtype=self.real_dtype).reshape(B, self.n, self.n)
        f, g = f.contiguous(), g.contiguous()
        dt = self.complex_dtype if complex_out else self.real_dtype
        out = torch.empty((B, self.n, self.n), dtype=dt, device=f.device)
        with self._lock:
            ws = self._workspace(int(self._lib.lb_fields_backward_workspace_bytes(self._h)),
                                 f.device)
            st = torch.cuda.current_stream(f.device).cuda_stream
            _lib.check(self._lib.lb_fields_backward(self._h, f.data_ptr(), B, g.data_ptr(),
                                                    out.data_ptr(), 1 if complex_out else 0,
                                                    ws.data_ptr(), ws.numel(), st))
        return out

    def backward(self, masks: torch.Tensor, d_images: torch.Tensor) -> torch.Tensor:
        """d/dmask of sum(d_images * I) for real masks [B, N, N]; d_images [B, P, N, N]."""
        m, cplx = self._check_masks(masks)
        if cplx:
            raise NotImplementedError("the adjoint is implemented for real masks")
        B = m.shape[0]
        P = len(self.planes)
        g = d_images.to(self.real_dtype).reshape(B, P, self.n, self.n).contiguous()
        out = torch.empty((B, self.n, self.n), dtype=self.real_dtype, device=m.device)
        with self._lock:
            self._use_tuning(B)
            nb = self.workspace_bytes(B, True)
            ws = self._workspace(nb, m.device)
            st = torch.cuda.current_stream(m.device).cuda_stream
            _lib.check(self._lib.lb_backward(self._h, m.data_ptr(), B, g.data_ptr(),
                                             out.data_ptr(), ws.data_ptr(), ws.numel(), st))
        return out

    def fields(self, masks: torch.Tensor) -> torch.Tensor:
        """SOCS plans: per-kernel coherent fields [B, K, N, N] complex on the full grid."""
        m, cplx = self._check_masks(masks)
        B = m.shape[0]
        K = self.info["n_systems"]
        out = torch.empty((B, K, self.n, self.n), dtype=self.complex_dtype, device=m.device)
        with self._lock:
            nb = int(self._lib.lb_fields_workspace_bytes(self._h, B))
            ws = self._workspace(nb, m.device)
            st = torch.cuda.current_stream(m.device).cuda_stream
            _lib.check(self._lib.lb_fields(self._h, m.data_ptr(), cplx, B, out.data_ptr(),
                                           ws.data_ptr(), ws.numel(), st))
        return out

    def forward_dense(self, mask: torch.Tensor, plane: int = 0, batch: int = 8) -> torch.Tensor:
        """The reference's dataflow on the device (one dense N x N transform per
        coherent system); for comparison -- same image as ``forward``."""
        m, cplx = self._check_masks(mask)
        if cplx or m.shape[0] != 1:
            raise ValueError("forward_dense takes one real mask")
        out = torch.empty((self.n, self.n), dtype=self.real_dtype, device=m.device)
        nb = int(self._lib.lb_dense_workspace_bytes(self._h, batch))
        with self._lock:
            ws = self._workspace(nb, m.device)
            st = torch.cuda.current_stream(m.device).cuda_stream
            _lib.check(self._lib.lb_forward_dense(self._h, m.data_ptr(), plane, batch,
                                                  out.data_ptr(), ws.data_ptr(), ws.numel(), st))
        return out

    def backward_dense(self, mask: torch.Tensor, d_image: torch.Tensor, plane: int = 0,
                       batch: int = 8) -> torch.Tensor:
        """The matching dense adjoint (one dense fft2 / ifft2 pair per coherent system);
        for comparison -- same gradient as ``backward``."""
        m, cplx = self._check_masks(mask)
        if cplx or m.shape[0] != 1:
            raise ValueError("backward_dense takes one real mask")
        g = d_image.to(self.real_dtype).reshape(self.n, self.n).contiguous()
        out = torch.empty((self.n, self.n), dtype=self.real_dtype, device=m.device)
        nb = int(self._lib.lb_dense_workspace_bytes(self._h, batch))
        with self._lock:
            ws = self._workspace(nb, m.device)
            st = torch.cuda.current_stream(m.device).cuda_stream
            _lib.check(self._lib.lb_backward_dense(self._h, m.data_ptr(), plane, batch,
                                                   g.data_ptr(), out.data_ptr(), ws.data_ptr(),
                                                   ws.numel(), st))
        return out

    def profile(self, on: bool) -> None:
        """Record HIP events around every coherent-system kernel launch (resets)."""
        _lib.check(self._lib.lb_profile_enable(self._h, 1 if on else 0))

    def profile_read(self, adjoint: bool = False) -> tuple[int, float]:
        """(launches, summed kernel milliseconds) of the forward (or adjoint)
        coherent-system launches since ``profile(True)``."""
        n = C.c_int32()
        t = C.c_double()
        fn = self._lib.lb_profile_read_adjoint if adjoint else self._lib.lb_profile_read
        _lib.check(fn(self._h, C.byref(n), C.byref(t)))
        return int(n.value), float(t.value)

    def kernels(self, device=None) -> torch.Tensor:
        """SOCS plans: the kernels on the full grid, FFT order, complex128 [K, N, N]."""
        K = self.info["n_kernels"]
        dev = device or torch.device("cuda", torch.cuda.current_device())
        out = torch.empty((K, self.n, self.n), dtype=torch.complex128, device=dev)
        if K:
            st = torch.cuda.current_stream(dev).cuda_stream
            _lib.check(self._lib.lb_socs_kernels(self._h, out.data_ptr(), st))
        return out


def pupil_dense(optics, grid, fs_x: float = 0.0, fs_y: float = 0.0, defocus=None,
                n_image=None) -> np.ndarray:
    """The scalar pupil P(f - f_s) on the whole grid in FFT order (complex128), computed on
    the device (lb_pupil_dense) with the optics' defocus (or ``defocus``), image index,
    Zernike terms and defocus model."""
    lib = _lib.load()
    if not torch.cuda.is_available():
        raise RuntimeError("the imaging library needs a HIP device (torch.cuda is unavailable)")
    keep = []

    def arr(x, dt=np.float64):
        a = np.ascontiguousarray(np.asarray(x, dtype=dt))
        keep.append(a)
        return a

    zern = {int(k): float(v) for k, v in (optics.zernike_coeffs or {}).items()}
    for j in zern:
        if not 1 <= j <= 21:  # expose/pupil.py:128-130: the term contributes zero
            logger.warning("Zernike Noll index %d not implemented; returning 0.", j)
    d = _lib.ImagingDesc()
    d.n = int(grid.n_pixels)
    d.pixel_size = float(grid.pixel_size)
    d.wavelength = float(optics.wavelength)
    d.na = float(optics.NA)
    d.exact_defocus = 1 if optics.exact_defocus else 0
    d.n_planes = 1
    d.defocus = _dptr(arr([optics.defocus if defocus is None else defocus]), C.c_double)
    d.n_image = _dptr(arr([_image_index(optics) if n_image is None else n_image]), C.c_double)
    d.n_zernike = len(zern)
    d.zernike_j = _dptr(arr(list(zern.keys()) or [0], np.int32), C.c_int32)
    d.zernike_c = _dptr(arr(list(zern.values()) or [0.0]), C.c_double)
    d.precision = 64
    n = d.n
    out = torch.empty((n, n), dtype=torch.complex128, device="cuda")
    st = torch.cuda.current_stream().cuda_stream
    _lib.check(lib.lb_pupil_dense(C.byref(d), float(fs_x), float(fs_y), 0, out.data_ptr(), st))
    return out.cpu().numpy()


# ----------------------------------------------------------------------------
# Plan cache keyed on the optical set-up (the role of PrintModel.kernels'
# cache in the reference, opc/model.py:201-221).
# ----------------------------------------------------------------------------

_CACHE: "OrderedDict[str, ImagingPlan]" = OrderedDict()
_CACHE_LOCK = threading.Lock()
_CACHE_SIZE = 16


def _key(optics, grid, planes, precision, kind, extra) -> str:
    od = asdict(optics) if is_dataclass(optics) else dict(vars(optics))
    gd = asdict(grid) if is_dataclass(grid) else dict(vars(grid))
    od["zernike_coeffs"] = {str(k): v for k, v in (od.get("zernike_coeffs") or {}).items()}
    dev = torch.cuda.current_device() if torch.cuda.is_available() else -1
    return json.dumps([od, gd, planes, precision, kind, extra, dev], sort_keys=True, default=str)


def get_plan(optics, grid, planes=None, precision: int = 32, kind: str = "abbe",
             **kw) -> ImagingPlan:
    if planes is None:
        planes = [(optics.defocus, _image_index(optics))]
    planes = [(float(d), float(n)) for d, n in planes]
    key = _key(optics, grid, planes, precision, kind, kw)
    with _CACHE_LOCK:
        plan = _CACHE.get(key)
        if plan is not None:
            _CACHE.move_to_end(key)
            return plan
    plan = ImagingPlan(optics, grid, planes, precision=precision, kind=kind, **kw)
    with _CACHE_LOCK:
        _CACHE[key] = plan
        while len(_CACHE) > _CACHE_SIZE:
            _CACHE.popitem(last=False)
    return plan


def clear_plan_cache() -> None:
    with _CACHE_LOCK:
        _CACHE.clear()
