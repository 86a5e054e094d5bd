"""Illumination sources and their point lists (host-side plan set-up).

The source is rasterised once per optical set-up on a small (source_grid^2)
pupil grid in float64, with the reference's exact grid formula, because the
rasterisation decides which points exist: a point sitting exactly on a sigma
boundary is in or out depending on the float64 value of rho, and that must
match the reference bit for bit.  The point list is then uploaded with the
plan; nothing here runs per mask.

Conventions (expose/pupil.py:35-48): rows = eta, columns = xi, both from
``linspace(-1, 1, n)``.  Sources are normalised to unit sum.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Any

import numpy as np


def pupil_grid(n: int):
    """(rho, phi, xi, eta) on the normalised [-1, 1] pupil grid (pupil.py:35-48)."""
    c = np.linspace(-1.0, 1.0, n)
    xi, eta = np.meshgrid(c, c)
    rho = np.sqrt(xi ** 2 + eta ** 2)
    return rho, np.arctan2(eta, xi), xi, eta


def _unit(a: np.ndarray) -> np.ndarray:
    s = a.sum()
    return a / s if s > 0 else a


def conventional(n: int, sigma: float) -> np.ndarray:
    """Filled disc (illumination.py:26-42)."""
    return _unit((pupil_grid(n)[0] <= sigma).astype(np.float64))


def annular(n: int, sigma_outer: float, sigma_inner: float) -> np.ndarray:
    """Ring (illumination.py:45-70)."""
    if sigma_inner >= sigma_outer:
        raise ValueError(
            f"sigma_inner ({sigma_inner:.3f}) must be < sigma_outer ({sigma_outer:.3f})")
    rho = pupil_grid(n)[0]
    return _unit(((rho >= sigma_inner) & (rho <= sigma_outer)).astype(np.float64))


def _discs(n, sigma_outer, sigma_inner, centres):
    r_pole = (sigma_outer - sigma_inner) / 2.0
    _, _, xi, eta = pupil_grid(n)
    out = np.zeros((n, n), dtype=np.float64)
    for cx, cy in centres:
        out += (np.sqrt((xi - cx) ** 2 + (eta - cy) ** 2) <= r_pole).astype(np.float64)
    return _unit(out)


def dipole(n: int, sigma_outer: float, sigma_inner: float, axis: str = "x") -> np.ndarray:
    """Two disc poles on an axis (illumination.py:73-98)."""
    c = (sigma_outer + sigma_inner) / 2.0
    centres = [(c, 0.0), (-c, 0.0)] if axis == "x" else [(0.0, c), (0.0, -c)]
    return _discs(n, sigma_outer, sigma_inner, centres)


def quadrupole(n: int, sigma_outer: float, sigma_inner: float,
               rotation_deg: float = 45.0) -> np.ndarray:
    """Four disc poles at rotation + k*90 deg (illumination.py:101-129)."""
    c = (sigma_outer + sigma_inner) / 2.0
    ang = np.deg2rad(np.array([0.0, 90.0, 180.0, 270.0]) + rotation_deg)
    return _discs(n, sigma_outer, sigma_inner, [(c * np.cos(a), c * np.sin(a)) for a in ang])


@dataclass(frozen=True)
class Pole:
    """Annular sector pole (expose/source.py:84-132)."""

    sigma_inner: float = 0.0
    sigma_outer: float = 0.9
    angle_deg: float = 0.0
    opening_deg: float = 360.0
    weight: float = 1.0

    def __post_init__(self):
        if self.sigma_outer <= 0:
            raise ValueError(f"sigma_outer must be > 0, got {self.sigma_outer}")
        if self.sigma_inner < 0:
            raise ValueError(f"sigma_inner must be >= 0, got {self.sigma_inner}")
        if self.sigma_inner >= self.sigma_outer:
            raise ValueError(f"sigma_inner ({self.sigma_inner:.3f}) must be < "
                             f"sigma_outer ({self.sigma_outer:.3f})")
        if not 0.0 < self.opening_deg <= 360.0:
            raise ValueError(f"opening_deg must be in (0, 360], got {self.opening_deg}")
        if self.weight < 0:
            raise ValueError(f"weight must be >= 0, got {self.weight}")

    def render(self, n: int) -> np.ndarray:
        rho, phi, _, _ = pupil_grid(n)
        radial = (rho >= self.sigma_inner) & (rho <= self.sigma_outer)
        if self.opening_deg >= 360.0:
            angular = np.ones_like(radial, dtype=bool)
        else:
            d = np.angle(np.exp(1j * (phi - np.deg2rad(self.angle_deg))))
            angular = np.abs(d) <= np.deg2rad(self.opening_deg) / 2.0
        return self.weight * (radial & angular).astype(np.float64)


@dataclass
class Source:
    """Composable source: sector poles or a free-form map (expose/source.py:143-219)."""

    poles: list = field(default_factory=list)
    pixels: np.ndarray | None = None
    blur: float = 0.0
    prune: float = 1e-3
    name: str = "source"

    def render(self, n: int, normalise: bool = True) -> np.ndarray:
        if n < 3:
            raise ValueError(f"source grid must be >= 3, got {n}")
        if self.pixels is not None:
            src = self._render_pixels(n)
        else:
            src = np.zeros((n, n), dtype=np.float64)
            for pole in self.poles:
                src += pole.render(n)
        if self.blur > 0.0:
            from scipy.ndimage import gaussian_filter
            src = gaussian_filter(src, sigma=self.blur * (n - 1) / 2.0, mode="constant")
        rho = pupil_grid(n)[0]
        src = np.where(rho <= 1.0, src, 0.0)
        if self.prune > 0.0 and src.max() > 0:
            src = np.where(src >= self.prune * src.max(), src, 0.0)
        if normalise:
            total = src.sum()
            if total > 0:
                src = src / total
        return src

    def _render_pixels(self, n: int) -> np.ndarray:
        px = np.asarray(self.pixels, dtype=np.float64)
        if px.ndim != 2:
            raise ValueError(f"pixels must be 2-D, got shape {px.shape}")
        if px.shape == (n, n):
            return px.copy()
        from scipy.ndimage import zoom
        out = zoom(px, (n / px.shape[0], n / px.shape[1]), order=1, mode="nearest")
        if out.shape != (n, n):
            fixed = np.zeros((n, n), dtype=np.float64)
            k = min(n, out.shape[0]), min(n, out.shape[1])
            fixed[: k[0], : k[1]] = out[: k[0], : k[1]]
            out = fixed
        return np.clip(out, 0.0, None)

    @classmethod
    def quadrupole(cls, sigma_outer=0.9, sigma_inner=0.6, opening_deg=40.0, rotation_deg=45.0,
                   **kw):
        return cls(poles=[Pole(sigma_inner, sigma_outer, a + rotation_deg, opening_deg)
                          for a in (0.0, 90.0, 180.0, 270.0)], name="quadrupole", **kw)

    @classmethod
    def monopole(cls, sigma_outer=0.9, sigma_inner=0.6, angle_deg=0.0, opening_deg=60.0, **kw):
        return cls(poles=[Pole(sigma_inner, sigma_outer, angle_deg, opening_deg)],
                   name="monopole", **kw)

    @classmethod
    def quasar(cls, sigma_outer=0.9, sigma_inner=0.6, opening_deg=40.0, rotation_deg=45.0, **kw):
        s = cls.quadrupole(sigma_outer, sigma_inner, opening_deg, rotation_deg, **kw)
        s.name = "quasar"
        return s

    @classmethod
    def cquad(cls, sigma_outer=0.9, sigma_inner=0.6, opening_deg=40.0, **kw):
        s = cls.quadrupole(sigma_outer, sigma_inner, opening_deg, 0.0, **kw)
        s.name = "cquad"
        return s

    @classmethod
    def from_dict(cls, data: dict[str, Any]) -> "Source":
        px = data.get("pixels")
        return cls(poles=[Pole(**p) for p in data.get("poles", [])],
                   pixels=None if px is None else np.asarray(px, dtype=np.float64),
                   blur=float(data.get("blur", 0.0)), prune=float(data.get("prune", 1e-3)),
                   name=str(data.get("name", "source")))


def build_source(n_pixels: int, source_type: str, sigma_outer: float, sigma_inner: float = 0.0,
                 spec: dict | None = None, **kwargs) -> np.ndarray:
    """illumination.py:132-219: normalised source map, rows = eta."""
    if spec is not None:
        return Source.from_dict(spec).render(n_pixels)
    builders = {
        "conventional": lambda: conventional(n_pixels, sigma_outer),
        "annular": lambda: annular(n_pixels, sigma_outer, sigma_inner),
        "dipole": lambda: dipole(n_pixels, sigma_outer, sigma_inner, **kwargs),
        "quadrupole": lambda: quadrupole(n_pixels, sigma_outer, sigma_inner, **kwargs),
        "monopole": lambda: Source.monopole(sigma_outer, sigma_inner, **kwargs).render(n_pixels),
        "quasar": lambda: Source.quasar(sigma_outer, sigma_inner, **kwargs).render(n_pixels),
        "cquad": lambda: Source.cquad(sigma_outer, sigma_inner, **kwargs).render(n_pixels),
    }
    key = source_type.lower()
    if key not in builders:
        raise ValueError(f"Unknown source type '{source_type}'. Available: {list(builders)}")
    if kwargs and key in ("conventional", "annular"):
        raise ValueError(
            f"Source type '{key}' accepts no extra arguments, got {sorted(kwargs)}. Use "
            "'monopole', 'dipole', 'quasar' or 'cquad' for shapes with an angle or opening.")
    return builders[key]()


@dataclass(frozen=True)
class SourcePoints:
    """The source as the point set the Abbe sum runs over (aerial_image.py:249-272)."""

    weight: np.ndarray
    fx: np.ndarray
    fy: np.ndarray
    phi: np.ndarray
    grid: int

    def __len__(self) -> int:
        return int(self.weight.size)


def source_points(optics) -> SourcePoints:
    """Non-zero source pixels as (weight, fx, fy, phi) (aerial_image.py:275-304)."""
    ns = max(int(optics.source_grid), 3)
    src = build_source(ns, optics.source_type, optics.sigma_outer, optics.sigma_inner,
                       spec=getattr(optics, "source_spec", None),
                       **(getattr(optics, "source_kwargs", None) or {}))
    idx = np.argwhere(src > 0)
    rows = idx[:, 0].astype(np.float64)
    cols = idx[:, 1].astype(np.float64)
    half = (ns - 1) / 2.0
    k_max = optics.NA / optics.wavelength
    return SourcePoints(
        weight=src[idx[:, 0], idx[:, 1]].astype(np.float64),
        fx=((cols - half) / half) * k_max,
        fy=((rows - half) / half) * k_max,
        phi=np.arctan2(rows - half, cols - half),
        grid=ns,
    )
