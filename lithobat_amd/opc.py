"""Print and inverse-lithography solve on the device (opc/model.py, opc/ilt.py,
develop/resist.py and bake/peb.py of the reference, threshold develop model).

``PrintModel`` carries the process (optics, resist, grid, dose, tone) and builds
its SOCS kernels once; ``print_mask`` images a mask array and develops it to the
latent field; ``run_ilt`` is the reference's Adam solve on the latent theta with
the mask held in [0, 1] by sigmoid(beta G(theta)), beta annealed, the relaxed
resist objective (PEB blur, TV and discreteness penalties) and the SOCS adjoint
-- every array stays on the device; the host sees one scalar cost per iteration.
"""
from __future__ import annotations

import dataclasses
from dataclasses import dataclass, field
from typing import Callable

import numpy as np
import torch

from .blur import blur, gaussian_blur
from .expose import _to_device, normalisation_scale
from .hopkins import SOCSKernels
from .ilt import abbe_intensity, mask_of, mask_penalties, relaxed_resist_loss
from .objective import adam_step, ilt_schedule, mask_sigmoid_grad


def apply_peb(pac_map, diffusion_sigma: float, pixel_size: float):
    """peb.py:24-70: Gaussian of diffusion_sigma / pixel_size px, clipped to [0, 1]
    (a copy, unblurred, below 1e-3 px).  NumPy in -> FP64 on the device -> NumPy out."""
    was_np = not isinstance(pac_map, torch.Tensor)
    a = (torch.as_tensor(np.asarray(pac_map, dtype=np.float64), device="cuda") if was_np
         else pac_map)
    if a.dim() != 2:
        raise ValueError(f"apply_peb expects a 2-D (ny, nx) map, got ndim={a.dim()}; "
                         "use apply_peb_3d for volumes.")
    if diffusion_sigma < 0:
        raise ValueError(f"diffusion_sigma must be >= 0, got {diffusion_sigma!r}")
    if pixel_size <= 0:
        raise ValueError(f"pixel_size must be > 0, got {pixel_size!r}")
    sigma_px = diffusion_sigma / pixel_size
    out = a.clone() if sigma_px < 1e-3 else torch.clamp(gaussian_blur(a, sigma_px), 0.0, 1.0)
    return out.cpu().numpy() if was_np else out


def develop_field(aerial, cfg, grid, model: str = "threshold"):
    """resist.py:452-503 for the threshold model: (PEB-diffused aerial, threshold)."""
    if model != "threshold":
        raise NotImplementedError(
            f"develop model {model!r} (resist chemistry) is outside this library; the "
            "threshold model is implemented")
    return apply_peb(aerial, cfg.diffusion_sigma, grid.pixel_size), float(cfg.threshold)


@dataclass
class PrintedImage:
    """model.py:60-96: the mask, its aerial image, the latent field and its level."""

    mask: object
    aerial: object
    field: object
    level: float
    tone: str
    grid: object

    @property
    def signed(self):
        s = self.field - self.level
        return s if self.tone == "clear" else -s

    @property
    def printed(self):
        return self.signed > 0.0


@dataclass
class PrintModel:
    """model.py:126-215 (the imaging part): process + SOCS kernels built once per optics.

    ``precision`` is the device precision of the kernels (64 by default, as the
    reference computes in float64)."""

    optics: object
    resist: object
    grid: object
    dose: float = 1.0
    tone: str = "clear"
    model: str = "threshold"
    normalisation: str | None = "clear"
    imaging: str = "auto"
    socs_tol: float = 1e-8
    precision: int = 64
    _cache: dict = field(default_factory=dict, repr=False, compare=False)

    def __post_init__(self) -> None:
        if self.tone not in ("clear", "dark"):
            raise ValueError(f"tone must be 'clear' or 'dark', got '{self.tone}'")
        if self.imaging not in ("auto", "socs", "abbe"):
            raise ValueError(f"imaging must be 'auto', 'socs' or 'abbe', got '{self.imaging}'")
        if self.normalisation is not None and self.normalisation != self.optics.normalisation:
            self.optics = dataclasses.replace(self.optics, normalisation=self.normalisation)
        if self.imaging == "socs" and not SOCSKernels.supports(self.optics):
            raise ValueError(
                "imaging='socs' needs the scalar thin-mask process; got "
                f"imaging_model={self.optics.imaging_model!r}, "
                f"mask_model={getattr(self.optics, 'mask_model', 'thin')!r}")

    def kernels(self) -> SOCSKernels | None:
        """model.py:201-221: the kernels, rebuilt only when optics / grid change."""
        if self.imaging == "abbe":
            return None
        if self.imaging == "auto" and not SOCSKernels.supports(self.optics):
            return None
        key = (repr(self.optics), self.grid.n_pixels, self.grid.pixel_size, self.socs_tol,
               self.precision)
        cached = self._cache.get("kernels")
        if cached is None or cached[0] != key:
            cached = (key, SOCSKernels.build(self.optics, self.grid, tol=self.socs_tol,
                                             precision=self.precision))
            self._cache["kernels"] = cached
        return cached[1]


def print_mask(mask, model: PrintModel) -> PrintedImage:
    """ilt.py:624-648: image a mask array through the model and develop it."""
    from .expose import compute_aerial_image
    kernels = model.kernels()
    if kernels is not None:
        raw, peak, clear = kernels.raw(mask)
    else:
        raw, peak, clear = compute_aerial_image(mask, model.optics, model.grid, return_raw=True)
    aerial = raw * normalisation_scale(model.optics.normalisation, peak, clear, model.dose)
    fld, level = develop_field(aerial, model.resist, model.grid, model=model.model)
    return PrintedImage(mask, aerial, fld, level, model.tone, model.grid)


def dose_for_target(target, model: PrintModel, mask=None) -> float:
    """ilt.py:651-721: the dose at which ``mask`` (default: the target) prints the target's
    area, ``model.dose`` when the target is empty or full.

    The unit-dose field blur(scale * raw) is imaged and blurred on the device once and
    sorted there.  The printed area sign * (dose * unit - level) > 0 is then a suffix
    (clear tone) or prefix (dark tone) of the sorted values, so each of the reference's 60
    geometric bisection steps counts it by a binary search that makes the same
    floating-point comparison: the same counts, hence the same dose.
    """
    n = model.grid.n_pixels
    tgt = np.asarray(target, dtype=np.float64)
    m = tgt if mask is None else np.asarray(mask, dtype=np.float64)
    if m.shape != (n, n):
        raise ValueError(f"mask must be {(n, n)} for this grid, got {m.shape}")
    want = int(np.count_nonzero(tgt > 0.5))
    if want == 0 or want == tgt.size:
        return float(model.dose)
    kernels = model.kernels()
    if kernels is None:
        raise ValueError("dose_for_target needs a model that images through SOCS kernels")
    md, _, _ = _to_device(m, kernels.plan.precision)
    raw, _peak, clear = kernels.raw(md)
    sigma_px = float(model.resist.diffusion_sigma / model.grid.pixel_size)
    unit = blur(raw * normalisation_scale(model.optics.normalisation, 1.0, clear, 1.0), sigma_px)
    u = torch.sort(unit.reshape(-1).double()).values.cpu().numpy()
    level = float(model.resist.threshold)
    clear_tone = model.tone == "clear"
    total = u.size

    def area(d: float) -> int:
        # clear tone: d * u - level > 0 (a suffix); dark tone: level - d * u > 0 (a prefix).
        # d * u is non-decreasing in u, so the first index where the comparison switches
        # is found by bisection over the sorted values.
        lo_i, hi_i = 0, total
        while lo_i < hi_i:
            mid_i = (lo_i + hi_i) // 2
            above = float(d * u[mid_i]) - level > 0.0 if clear_tone else \
                -(float(d * u[mid_i]) - level) <= 0.0
            if above:
                hi_i = mid_i
            else:
                lo_i = mid_i + 1
        return total - lo_i if clear_tone else lo_i

    lo, hi = 1e-3, 1e3
    rising = area(hi) > area(lo)
    for _ in range(60):
        mid = float(np.sqrt(lo * hi))
        if (area(mid) < want) == rising:
            lo = mid
        else:
            hi = mid
    return float(np.sqrt(lo * hi))


@dataclass
class ILTFrame:
    """ilt.py:60-78: one progress report."""

    iteration: int
    mask: np.ndarray
    signed: np.ndarray
    cost: float
    pattern_error: int
    beta: float


@dataclass
class ILTResult:
    """ilt.py:81-95."""

    mask: np.ndarray
    theta: np.ndarray
    target: np.ndarray
    printed: PrintedImage
    history: object
    iterations: int
    converged: bool


def run_ilt(target, model: PrintModel, *, initial=None, max_iter: int = 120, step: float = 0.4,
            tol: float = 1e-7, beta: float = 4.0, beta_final: float | None = 12.0,
            initial_grey: float = 0.9, steepness: float = 12.0, weight_tv: float = 0.0,
            weight_discrete: float = 0.0, smooth: float = 0.0,
            progress: Callable[[ILTFrame], None] | None = None,
            graph: bool = True) -> ILTResult:
    """ilt.py:397-621 on the device (same parameters, checks, schedule and history).

    ``graph``: after the first iteration, one iteration (the schedule update, the step
    back through the mask parameterisation, Adam, the new mask, the forward image,
    objective and adjoint) is captured into a HIP graph and replayed; its per-iteration
    scalars (beta, the Adam bias corrections) live on the device, so replays and eager
    iterations run the same kernels on the same values.  Each iteration still reads
    its cost, pattern error and gradient norm back for the history and the
    convergence test, as the reference does."""
    kernels = model.kernels()
    if kernels is None:
        raise ValueError(
            "inverse lithography images through the SOCS kernels; this model images through "
            f"the Abbe sum (imaging={model.imaging!r}, "
            f"imaging_model={model.optics.imaging_model!r}, "
            f"mask_model={getattr(model.optics, 'mask_model', 'thin')!r}). The adjoint is "
            "written against the kernel decomposition and has no thick-mask form.")
    if model.optics.normalisation.lower() == "peak":
        raise ValueError(
            "inverse lithography needs a mask-independent normalisation; 'peak' rescales the "
            "image by whatever the current mask's brightest point happens to be, which moves "
            "the objective every step. Use normalisation='clear'.")
    n = model.grid.n_pixels
    tgt = np.asarray(target, dtype=np.float64)
    if tgt.shape != (n, n):
        raise ValueError(f"target must be {(n, n)} for this grid, got {tgt.shape}")
    if smooth < 0.0:
        raise ValueError(f"smooth is a length and cannot be negative, got {smooth:g}")
    smooth_px = float(smooth / model.grid.pixel_size)
    scale = normalisation_scale(model.optics.normalisation, 1.0, kernels.clear, model.dose)
    level = float(model.resist.threshold)
    sigma_px = float(model.resist.diffusion_sigma / model.grid.pixel_size)
    sign = 1.0 if model.tone == "clear" else -1.0
    plan = kernels.plan
    dt = plan.real_dtype
    dev = torch.device("cuda", torch.cuda.current_device())
    t_dev = torch.as_tensor(tgt, dtype=dt, device=dev)

    def objective(m):
        """ilt.py:311-360 on the device: cost (device scalar), d cost / d mask, the signed
        field and the pattern error (device int), the last two from the loss's own pass."""
        m = m.detach().requires_grad_(True)
        raw = abbe_intensity(m[None], plan)  # [1, 1, n, n]
        side = {}
        cost = (relaxed_resist_loss(raw, t_dev[None, None], scale, level, steepness, sign,
                                    sigma_px, side=side)
                + mask_penalties(m, weight_tv, weight_discrete))
        cost.backward()
        return cost.detach(), m.grad, side["signed"][0, 0], side["mismatch"][0, 0]

    def grad_theta(grad_m, m, beta_k):
        """ilt.py:381-389: back through the sigmoid's slope, then the Gaussian."""
        return blur(mask_sigmoid_grad(grad_m, m, beta_k), smooth_px)

    p = float(np.clip(initial_grey, 0.5 + 1e-3, 1.0 - 1e-6))
    if initial is None:
        m0 = 0.5 + (np.clip(tgt, 0.0, 1.0) - 0.5) * (2.0 * p - 1.0)
    else:
        m0 = np.clip(np.asarray(initial, dtype=np.float64), 1.0 - p, p)
    theta = torch.as_tensor(np.log(m0 / (1.0 - m0)) / beta, dtype=dt, device=dev).contiguous()
    growth = (1.0 if beta_final is None or max_iter <= 1
              else (beta_final / beta) ** (1.0 / (max_iter - 1)))
    mom1 = torch.zeros_like(theta)
    mom2 = torch.zeros_like(theta)
    b1, b2, eps = 0.9, 0.999, 1e-8
    beta_k = float(beta)
    # [beta, it, 1 - b1^it, 1 - b2^it, beta_next] on the device (ilt_schedule)
    sched = torch.tensor([beta_k, 0.0, 0.0, 0.0, beta_k], dtype=torch.float64, device=dev)
    with torch.no_grad():
        mask = mask_of(theta, beta_k, smooth_px)
    cost, grad_m, signed, perr = objective(mask)
    lit = int((signed > 0.0).sum())
    if lit == 0 or lit == signed.numel():
        peak = float((signed.abs() + level).max())
        raise ValueError(
            f"the initial mask prints {'nothing' if lit == 0 else 'everywhere'} — no contour "
            f"exists, so the objective has no gradient. The resist level is {level:g} and the "
            f"field reaches {peak:.4g} at dose {model.dose:g}. Size the dose first.")
    rows = []

    def record(it, cost, grad_m, signed, perr, mask):
        # the three history values reach the host in one copy
        vals = torch.stack([cost.double(), perr.double(),
                            torch.linalg.vector_norm(grad_m).double()]).cpu().numpy()
        c, pe = float(vals[0]), int(vals[1])
        rows.append({"iteration": it, "cost": c, "pattern_error": pe, "beta": beta_k,
                     "grad_norm": float(vals[2])})
        if progress is not None:
            progress(ILTFrame(it, mask.cpu().numpy(), signed.cpu().numpy(), c, pe, beta_k))
        return c

    record(0, cost, grad_m, signed, perr, mask)
    # the iteration's inputs and outputs at fixed addresses (a graph replays on them)
    state = {"mask": mask.detach().clone(), "grad_m": grad_m.detach().clone()}

    def iteration():
        ilt_schedule(sched, growth, b1, b2)
        with torch.no_grad():
            grad_t = grad_theta(state["grad_m"], state["mask"], sched[0:1]).contiguous()
            adam_step(theta, mom1, mom2, grad_t, sched[2:4], step, b1, b2, eps)
            new_mask = mask_of(theta, sched[4:5], smooth_px)
        out = objective(new_mask)
        with torch.no_grad():
            state["mask"].copy_(new_mask)
            state["grad_m"].copy_(out[1])
        return out

    prev_cost, converged, taken = np.inf, False, 0
    g = None
    for it in range(1, max_iter + 1):
        if g is None:
            outs = iteration()
        else:
            g.replay()
        beta_k *= growth
        cost, _, signed, perr = outs
        taken = it
        c = record(it, cost, state["grad_m"], signed, perr, state["mask"])
        if abs(prev_cost - c) < tol:
            converged = True
            break
        prev_cost = c
        if graph and g is None and it < max_iter:
            # capture records the iteration without running it; replays run it
            torch.cuda.current_stream().synchronize()
            try:
                g = torch.cuda.CUDAGraph()
                with torch.cuda.graph(g):
                    outs = iteration()
            except RuntimeError:  # a stack that cannot capture this step: stay eager
                g, graph = None, False
                torch.cuda.synchronize()
    mask = state["mask"]
    try:
        import pandas as pd
        history = pd.DataFrame(rows)
    except ImportError:  # pragma: no cover - pandas is a reference dependency
        history = rows
    mask_np = mask.double().cpu().numpy()
    printed = print_mask(mask_np if model.precision == 64 else mask, model)
    return ILTResult(mask=mask_np, theta=theta.double().cpu().numpy(), target=tgt > 0.5,
                     printed=printed, history=history, iterations=taken, converged=converged)
