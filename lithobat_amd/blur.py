"""The PEB / ILT Gaussian blur on the device (opc/ilt.py:274-289 of the reference).

``gaussian_blur`` is scipy.ndimage.gaussian_filter(a, sigma) with its defaults
(mode "reflect", truncate 4.0) over the last two axes; below 1e-3 px it returns
its input, as the reference does.  ``GaussianBlur`` is the autograd form: the
blur is its own adjoint (symmetric taps, symmetric extension), which is also
how the reference's chain rule applies it.
"""
from __future__ import annotations

import torch

from . import _lib


def gaussian_blur(a: torch.Tensor, sigma: float) -> torch.Tensor:
    if sigma < 1e-3:
        return a
    if not isinstance(a, torch.Tensor) or a.device.type != "cuda":
        raise TypeError("gaussian_blur takes a tensor on the HIP device")
    if a.dim() < 2 or a.shape[-1] != a.shape[-2]:
        raise ValueError(f"need (..., n, n) planes, got {tuple(a.shape)}")
    if a.dtype not in (torch.float32, torch.float64):
        raise TypeError("gaussian_blur takes float32 or float64 planes")
    lib = _lib.load()
    n = a.shape[-1]
    x = a.contiguous()
    n_img = x.numel() // (n * n)
    prec = 64 if x.dtype == torch.float64 else 32
    out = torch.empty_like(x)
    nb = int(lib.lb_gaussian_blur_workspace_bytes(n_img, n, float(sigma), prec))
    ws = torch.empty(max(nb, 256), dtype=torch.uint8, device=x.device)
    st = torch.cuda.current_stream(x.device).cuda_stream
    _lib.check(lib.lb_gaussian_blur(x.data_ptr(), out.data_ptr(), n_img, n, float(sigma), prec,
                                    ws.data_ptr(), ws.numel(), st))
    return out


class GaussianBlur(torch.autograd.Function):
    @staticmethod
    def forward(ctx, a: torch.Tensor, sigma: float):
        ctx.sigma = sigma
        return gaussian_blur(a, sigma)

    @staticmethod
    def backward(ctx, g: torch.Tensor):
        return gaussian_blur(g.contiguous(), ctx.sigma), None


def blur(a: torch.Tensor, sigma: float) -> torch.Tensor:
    """Differentiable blur; identity below 1e-3 px."""
    if sigma < 1e-3:
        return a
    return GaussianBlur.apply(a, sigma)
