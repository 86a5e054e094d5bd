"""Sum of coherent systems (expose/hopkins.py of the reference), on the device.

``SOCSKernels.build`` decomposes the optics in FP64 on the device (Gram
matrix of the TCC rows, Householder tridiagonalisation and implicit QL);
``raw`` / ``image`` image masks through the kept kernels with the same coherent-system kernel as
the Abbe sum, in the precision the kernels were built for.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np
import torch

from .expose import _to_device, normalisation_scale
from .plan import ImagingPlan, _image_index


@dataclass(frozen=True)
class SOCSKernels:
    """The eigen-kernels of one optical state (hopkins.py:58-94).

    ``kernels`` (K, n, n) complex128 on the full FFT grid is materialised on
    first access; the device plan holds them on the band only.
    """

    weights: np.ndarray
    clear: float
    energy_kept: float
    n_band: int
    n_source: int
    optics: object
    grid: object
    plan: ImagingPlan = field(repr=False, compare=False)

    @staticmethod
    def supports(optics) -> bool:
        """hopkins.py:98-101."""
        return optics.imaging_model.lower() == "scalar" and getattr(optics, "mask_model",
                                                                    "thin") == "thin"

    @classmethod
    def build(cls, optics, grid, tol: float = 1e-8, max_kernels: int | None = None, *,
              precision: int = 64) -> "SOCSKernels":
        """hopkins.py:103-198; ``precision`` is the imaging precision of the kernels."""
        if not cls.supports(optics):
            raise ValueError(
                "SOCS kernels apply to the scalar thin-mask process only; got "
                f"imaging_model={optics.imaging_model!r}, "
                f"mask_model={getattr(optics, 'mask_model', 'thin')!r}. "
                "Use compute_aerial_image for those.")
        plan = ImagingPlan(optics, grid, [(optics.defocus, _image_index(optics))],
                           precision=precision, kind="socs", tol=tol, max_kernels=max_kernels)
        return cls(weights=plan.weights, clear=float(plan.clear[0]),
                   energy_kept=plan.energy_kept, n_band=plan.n_band, n_source=plan.n_source,
                   optics=optics, grid=grid, plan=plan)

    def __len__(self) -> int:
        return int(self.weights.size)

    @property
    def kernels(self) -> np.ndarray:
        k = self.__dict__.get("_kernels_np")
        if k is None:
            k = self.plan.kernels().cpu().numpy()
            object.__setattr__(self, "_kernels_np", k)
        return k

    def _check(self, mask):
        n = self.grid.n_pixels
        shape = tuple(mask.shape[-2:]) if hasattr(mask, "shape") else np.shape(mask)
        if tuple(shape) != (n, n):
            raise ValueError(f"mask must be {(n, n)} for these kernels, got {tuple(mask.shape)}")

    def raw(self, mask):
        """hopkins.py:205-218: (intensity, peak, clear)."""
        self._check(mask)
        m, was_np, _ = _to_device(mask, self.plan.precision)
        imgs, peak = self.plan.forward(m)
        if was_np or m.dim() == 2:
            raw = imgs[0, 0]
            return (raw.double().cpu().numpy() if was_np else raw), float(peak[0, 0]), self.clear
        return imgs[:, 0], peak[:, 0], self.clear

    def image(self, mask, dose: float = 1.0):
        """hopkins.py:220-223."""
        raw, peak, clear = self.raw(mask)
        if isinstance(peak, torch.Tensor):
            mode = self.optics.normalisation.lower()
            if mode == "peak":
                s = torch.where(peak > 0, dose / peak, torch.full_like(peak, dose))
            else:
                s = torch.full_like(peak, normalisation_scale(mode, 1.0, clear, dose))
            return raw * s[:, None, None]
        return raw * normalisation_scale(self.optics.normalisation, peak, clear, dose)


__all__ = ["SOCSKernels"]
