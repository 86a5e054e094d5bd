"""Depth-resolved aerial image through the resist film (develop/resist3d.py of the reference).

``exposure_volume`` images ``grid.n_z_slices`` optical planes through the film in
one batched launch (the planes share the mask spectrum and every source point's
transfer geometry; only the defocus phase and the image index differ) and
interpolates them linearly onto the voxel grid on the device, each plane's image
normalisation folded into the same pass (lb_volume_interp).
"""
from __future__ import annotations

import numpy as np
import torch

from . import _lib


def focus_reference_depth(resist) -> float:
    """resist3d.py:66-77: depth below the resist top that ``optics.defocus`` refers to [m]."""
    ref = resist.focus_reference
    if ref == "top":
        return 0.0
    if ref == "mid":
        return resist.thickness / 2.0
    if ref == "bottom":
        return resist.thickness
    raise ValueError(f"focus_reference must be 'top', 'mid', or 'bottom', got '{ref}'")


def effective_defocus(depth: float, optics, resist) -> float:
    """resist3d.py:80-110: defocus of the plane ``depth`` below the resist top, in
    immersion-equivalent units."""
    d_ref = focus_reference_depth(resist)
    return optics.defocus - (depth - d_ref) * (optics.n_immersion / resist.n_resist)


def volume_planes(optics, grid, resist):
    """The optical planes (defocus, image index) and their depths below the top
    (resist3d.py:144-171): defocus converted to resist units, image index n_resist."""
    nz = max(int(round(resist.thickness / grid.dz)), 1)
    n_slices = int(np.clip(grid.n_z_slices, 2, nz))
    depths = np.linspace(0.0, resist.thickness, n_slices)
    ratio = resist.n_resist / optics.n_immersion
    planes = [(effective_defocus(float(d), optics, resist) * ratio, float(resist.n_resist))
              for d in depths]
    return planes, depths, nz


def _interp_weights(depths: np.ndarray, nz: int, dz: float):
    """For each voxel (bottom-up), the two planes around its depth and the weight of the
    deeper one, with numpy.interp's clamping outside [depths[0], depths[-1]]."""
    zd = ((np.arange(nz) + 0.5) * dz)[::-1]  # depth below the top, iz = 0 at the bottom
    j = np.clip(np.searchsorted(depths, zd, side="right") - 1, 0, len(depths) - 2)
    t = (zd - depths[j]) / (depths[j + 1] - depths[j])
    t = np.clip(t, 0.0, 1.0)
    return j, t


def exposure_volume(mask, optics, grid, resist, dose: float = 1.0, *,
                    precision: int | None = None):
    """resist3d.py:113-185: (intensity (nz, N, N), z) with iz = 0 at the bottom of the film.

    NumPy masks return float64 NumPy arrays computed in FP64 on the device; a torch
    mask on the device returns device tensors in its precision.
    """
    from .expose import _raw
    planes, depths, nz = volume_planes(optics, grid, resist)
    imgs, peak, plan, was_np, single = _raw(mask, optics, grid, planes, precision)
    B, P, n = imgs.shape[0], imgs.shape[1], imgs.shape[-1]
    dt, dev = imgs.dtype, imgs.device
    # each plane's image normalisation (aerial_image.py:217-241), as compute_aerial_planes
    mode = optics.normalisation.lower()
    if mode not in ("peak", "clear", "none"):
        raise ValueError(f"normalisation must be 'peak', 'clear' or 'none', got "
                         f"'{optics.normalisation}'")
    # voxel weights and fixed plane scales, kept on the plan (no host copies per call)
    cache = plan.__dict__.setdefault("_volume_tables", {})
    key = (nz, float(grid.dz), tuple(float(d) for d in depths), mode, float(dose), str(dev))
    if key not in cache:
        j, t = _interp_weights(depths, nz, grid.dz)
        fixed = None
        if mode != "peak":
            fixed = torch.tensor([dose / c if (mode == "clear" and c > 0) else dose
                                  for c in plan.clear], dtype=dt, device=dev)
        cache[key] = (torch.as_tensor(j, dtype=torch.int32, device=dev),
                      torch.as_tensor(t, dtype=dt, device=dev), fixed)
    jj, tt, fixed = cache[key]
    if mode == "peak":
        scale = torch.where(peak > 0, dose / peak, torch.full_like(peak, dose)).contiguous()
    else:
        scale = fixed.expand(B, P).contiguous()
    vol = torch.empty((B, nz, n, n), dtype=dt, device=dev)
    _lib.check(_lib.load().lb_volume_interp(
        imgs.data_ptr(), B, P, n * n, nz, jj.data_ptr(), tt.data_ptr(), scale.data_ptr(),
        vol.data_ptr(), 64 if dt == torch.float64 else 32,
        torch.cuda.current_stream(dev).cuda_stream))
    z = (np.arange(nz) + 0.5) * grid.dz
    if was_np:
        return vol[0].double().cpu().numpy(), np.asarray(z, dtype=np.float64)
    return (vol[0] if single else vol), torch.as_tensor(z, dtype=dt, device=dev)
