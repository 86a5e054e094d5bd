"""Drop-in aerial-image entry points (expose/aerial_image.py of the reference).

``compute_aerial_image`` / ``compute_aerial_planes`` keep the reference's
signatures and semantics (aerial_image.py:502-587).  NumPy input returns
NumPy float64 computed in FP64 on the device (the reference's precision);
torch input on the device stays on the device, in the precision of its dtype
(float32 / complex64 -> the FP32 path, float64 / complex128 -> FP64), and may
carry a leading batch dimension of masks.
"""
from __future__ import annotations

import numpy as np
import torch

from .plan import ImagingPlan, _image_index, get_plan, pupil_dense
from .source import SourcePoints, source_points  # noqa: F401  (re-exported)

__all__ = ["compute_aerial_image", "compute_aerial_planes", "compute_aerial_batch",
           "normalisation_scale", "source_points", "SourcePoints", "SOURCE_TILT_SIGN",
           "FFT_BATCH_BYTES", "build_fft_pupil"]

#: aerial_image.py:73 -- the pupil is evaluated at f - f_s (a source point at
#: +sigma tilts the mask by exp(-i 2 pi f_s . x)); the device tables follow it.
SOURCE_TILT_SIGN = -1.0
#: aerial_image.py:78 -- kept for API compatibility; the device path batches by
#: work items instead of by bytes.
FFT_BATCH_BYTES = 64 * 2 ** 20


def build_fft_pupil(optics, grid, fs_x: float = 0.0, fs_y: float = 0.0, defocus=None,
                    n_image=None) -> np.ndarray:
    """aerial_image.py:169-209 (_build_fft_pupil): the pupil P(f - f_s) = exp(i W) on the
    in-pupil bins of the N x N grid, zeros elsewhere, in np.fft.fft2 ordering (DC at
    [0, 0]), complex128.  The reference takes the geometry arrays of _pupil_geometry; here
    the source offset (fs_x, fs_y) [1/m] is given and the geometry, defocus OPD (the
    optics' defocus and image index unless passed) and Zernike phase are evaluated on
    the device in FP64."""
    return pupil_dense(optics, grid, fs_x, fs_y, defocus, n_image)


def normalisation_scale(mode: str, peak: float, clear: float, dose: float = 1.0) -> float:
    """aerial_image.py:217-241: the one scalar that turns a raw sum into an image."""
    m = mode.lower()
    if m == "peak":
        return dose / peak if peak > 0 else dose
    if m == "clear":
        return dose / clear if clear > 0 else dose
    if m == "none":
        return dose
    raise ValueError(f"normalisation must be 'peak', 'clear' or 'none', got '{mode}'")


def _check_model(optics) -> None:
    model = optics.imaging_model.lower()
    if model not in ("scalar", "vector"):
        raise ValueError(f"imaging_model must be 'scalar' or 'vector', got "
                         f"'{optics.imaging_model}'")


def _to_device(mask, precision):
    """NumPy / torch -> (device tensor, was_numpy, precision)."""
    if isinstance(mask, torch.Tensor) and mask.device.type == "cuda":
        if precision is None:
            precision = 64 if mask.dtype in (torch.float64, torch.complex128) else 32
        return mask, False, precision
    arr = np.asarray(mask)
    if precision is None:
        precision = 64
    if np.iscomplexobj(arr):
        dt = torch.complex128 if precision == 64 else torch.complex64
    else:
        dt = torch.float64 if precision == 64 else torch.float32
    if not torch.cuda.is_available():
        raise RuntimeError("the imaging library needs a HIP device (torch.cuda is unavailable)")
    return torch.as_tensor(arr).to(device="cuda", dtype=dt), True, precision


def _raw(mask, optics, grid, planes, precision, plan: ImagingPlan | None = None):
    _check_model(optics)
    m, was_np, precision = _to_device(mask, precision)
    if plan is None:
        plan = get_plan(optics, grid, planes, precision=precision)
    imgs, peak = plan.forward(m)
    return imgs, peak, plan, was_np, m.dim() == 2


def compute_aerial_image(mask, optics, grid, dose: float = 1.0, return_raw: bool = False, *,
                         precision: int | None = None):
    """aerial_image.py:502-548: one plane at (optics.defocus, optics.image_index)."""
    planes = [(optics.defocus, _image_index(optics))]
    imgs, peak, plan, was_np, single = _raw(mask, optics, grid, planes, precision)
    clear = float(plan.clear[0])
    if was_np or single:
        raw = imgs[0, 0]
        pk = float(peak[0, 0])
        if return_raw:
            return (raw.double().cpu().numpy() if was_np else raw), pk, clear
        out = raw * normalisation_scale(optics.normalisation, pk, clear, dose)
        return out.double().cpu().numpy() if was_np else out
    # batched device masks [B, N, N]
    raw = imgs[:, 0]
    if return_raw:
        return raw, peak[:, 0], clear
    return _normalise(raw, peak[:, 0], clear, optics.normalisation, dose)


def _normalise(raw: torch.Tensor, peak: torch.Tensor, clear: float, mode: str, dose: float):
    m = mode.lower()
    if m == "peak":
        scale = torch.where(peak > 0, dose / peak, torch.full_like(peak, dose))
    elif m == "clear":
        scale = torch.full_like(peak, dose / clear if clear > 0 else dose)
    elif m == "none":
        scale = torch.full_like(peak, dose)
    else:
        raise ValueError(f"normalisation must be 'peak', 'clear' or 'none', got '{mode}'")
    return raw * scale.reshape(scale.shape + (1, 1))


def compute_aerial_planes(mask, optics, grid, planes, dose: float = 1.0, *,
                          precision: int | None = None):
    """aerial_image.py:551-587: (P, N, N), each plane normalised on its own."""
    planes = [(float(d), float(n)) for d, n in planes]
    imgs, peak, plan, was_np, single = _raw(mask, optics, grid, planes, precision)
    clear = torch.as_tensor(plan.clear, dtype=imgs.dtype, device=imgs.device)
    out = torch.empty_like(imgs)
    for p in range(len(planes)):
        out[:, p] = _normalise(imgs[:, p], peak[:, p], float(clear[p]), optics.normalisation, dose)
    if was_np:
        return out[0].double().cpu().numpy()
    return out[0] if single else out


def compute_aerial_batch(masks: torch.Tensor, optics, grid, planes=None, dose=1.0, *,
                         return_raw: bool = False, plan: ImagingPlan | None = None):
    """Batched device form: masks [B, N, N] x planes -> images [B, P, N, N].

    Dose is a scalar or a tensor broadcastable to [B, P] (dose samples are a
    post-hoc scale on the raw sum, aerial_image.py:217-241).  With
    ``return_raw`` returns (raw, peak, clear) instead.
    """
    if planes is None:
        planes = [(optics.defocus, _image_index(optics))]
    if plan is None:
        prec = 64 if masks.dtype in (torch.float64, torch.complex128) else 32
        plan = get_plan(optics, grid, planes, precision=prec)
    raw, peak = plan.forward(masks)
    clear = torch.as_tensor(plan.clear, dtype=raw.dtype, device=raw.device)
    if return_raw:
        return raw, peak, clear
    mode = optics.normalisation.lower()
    dose_t = torch.as_tensor(dose, dtype=raw.dtype, device=raw.device)
    if mode == "peak":
        denom = torch.where(peak > 0, peak, torch.ones_like(peak))
    elif mode == "clear":
        denom = torch.where(clear > 0, clear, torch.ones_like(clear)).expand_as(peak)
    elif mode == "none":
        denom = torch.ones_like(peak)
    else:
        raise ValueError(f"normalisation must be 'peak', 'clear' or 'none', got '{mode}'")
    return raw * (dose_t / denom)[..., None, None]
