"""Configuration carriers with the reference's field names and validation.

``OpticsConfig`` and ``GridConfig`` mirror core/config.py:238-487 and
core/config.py:720-759 of the reference (same attributes, defaults and
construction-time checks), so a caller's existing objects or keyword sets work
unchanged.  Only the thin-mask model is imaged here; the thick-mask fields are
kept so configurations round-trip, and are refused by the imaging calls.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Any

IMAGING_MODELS = ("scalar", "vector")
POLARISATIONS = ("unpolarised", "x", "y", "te", "tm")
NORMALISATIONS = ("peak", "clear", "none")
MASK_MODELS = ("thin", "multilayer", "fdtd")
SOURCE_TYPES = ("conventional", "annular", "monopole", "dipole", "quadrupole", "quasar", "cquad")


def _choice(owner, name, value, choices):
    if value not in choices:
        raise ValueError(f"{owner}.{name} must be one of {list(choices)}, got {value!r}")


def _bounded(owner, name, value, lo=None, hi=None, *, lo_open=False, hi_open=False):
    """core/config.py:190-214: raise unless lo <= value <= hi."""
    try:
        v = float(value)
    except (TypeError, ValueError):
        raise TypeError(f"{owner}.{name} must be a number, got {value!r}") from None
    if v != v:
        raise ValueError(f"{owner}.{name} must not be NaN")
    if lo is not None and (v < lo or (lo_open and v == lo)):
        raise ValueError(f"{owner}.{name} must be {'>' if lo_open else '>='} {lo:g}, got {value!r}")
    if hi is not None and (v > hi or (hi_open and v == hi)):
        raise ValueError(f"{owner}.{name} must be {'<' if hi_open else '<='} {hi:g}, got {value!r}")


@dataclass
class OpticsConfig:
    """Projection optics (core/config.py:383-406 field list and defaults)."""

    wavelength: float = 193e-9
    NA: float = 0.93
    n_immersion: float = 1.0
    defocus: float = 0.0
    sigma_outer: float = 0.80
    sigma_inner: float = 0.0
    source_type: str = "conventional"
    source_grid: int = 21
    source_kwargs: dict[str, Any] = field(default_factory=dict)
    source_spec: dict[str, Any] | None = None
    zernike_coeffs: dict[int, float] = field(default_factory=dict)
    exact_defocus: bool = False
    n_image: float | None = None
    imaging_model: str = "scalar"
    polarisation: str = "unpolarised"
    obliquity: bool = True
    normalisation: str = "peak"
    mask_model: str = "thin"
    reduction: float = 4.0
    chief_ray_deg: float = 0.0
    chief_ray_axis: str = "x"
    mask_stack: dict[str, Any] | None = None
    mask_geometry: dict[str, Any] | None = None
    m3d_angles: int = 3

    def __post_init__(self) -> None:
        """core/config.py:408-451."""
        o = "OpticsConfig"
        _bounded(o, "wavelength", self.wavelength, 0.0, lo_open=True)
        _bounded(o, "NA", self.NA, 0.0, lo_open=True)
        _bounded(o, "n_immersion", self.n_immersion, 0.0, lo_open=True)
        if self.n_image is not None:
            _bounded(o, "n_image", self.n_image, 0.0, lo_open=True)
        if self.NA > self.image_index:
            raise ValueError(
                f"{o}: NA {self.NA:g} exceeds the image-space index {self.image_index:g}, "
                "which would make sin θ > 1. Raise n_immersion (or n_image) or lower NA.")
        _bounded(o, "defocus", self.defocus)
        _bounded(o, "sigma_outer", self.sigma_outer, 0.0, 1.0, lo_open=True)
        _bounded(o, "sigma_inner", self.sigma_inner, 0.0, self.sigma_outer)
        _bounded(o, "source_grid", self.source_grid, 3)
        if self.source_spec is None:
            _choice(o, "source_type", self.source_type, SOURCE_TYPES)
        _choice(o, "imaging_model", self.imaging_model, IMAGING_MODELS)
        _choice(o, "polarisation", self.polarisation, POLARISATIONS)
        _choice(o, "normalisation", self.normalisation, NORMALISATIONS)
        _choice(o, "mask_model", self.mask_model, MASK_MODELS)
        _choice(o, "chief_ray_axis", self.chief_ray_axis, ("x", "y"))
        _bounded(o, "reduction", self.reduction, 0.0, lo_open=True)
        _bounded(o, "chief_ray_deg", self.chief_ray_deg, -90.0, 90.0)
        _bounded(o, "m3d_angles", self.m3d_angles, 1)
        for j, c in (self.zernike_coeffs or {}).items():
            if not isinstance(j, int) or isinstance(j, bool) or j < 1:
                raise ValueError(f"{o}.zernike_coeffs keys are Noll indices (int ≥ 1), got {j!r}")
            _bounded(o, f"zernike_coeffs[{j}]", c)

    @property
    def k_max(self) -> float:
        return self.NA / self.wavelength

    @property
    def image_index(self) -> float:
        """core/config.py:469-477: n_image, defaulting to n_immersion."""
        return self.n_immersion if self.n_image is None else self.n_image

    @property
    def sin_theta_max(self) -> float:
        return self.NA / self.image_index


@dataclass
class GridConfig:
    """Simulation grid (core/config.py:741-754)."""

    n_pixels: int = 128
    pixel_size: float = 4e-9
    dz: float = 2e-9
    n_z_slices: int = 21

    def __post_init__(self) -> None:
        g = "GridConfig"
        if isinstance(self.n_pixels, bool) or int(self.n_pixels) != self.n_pixels:
            raise TypeError(f"{g}.n_pixels must be an integer, got {self.n_pixels!r}")
        _bounded(g, "n_pixels", self.n_pixels, 1)
        _bounded(g, "pixel_size", self.pixel_size, 0.0, lo_open=True)
        _bounded(g, "dz", self.dz, 0.0)
        _bounded(g, "n_z_slices", self.n_z_slices, 1)

    @property
    def grid_size(self) -> float:
        return self.n_pixels * self.pixel_size


FOCUS_REFERENCES = ("top", "mid", "bottom")


TONES = ("positive", "negative")


@dataclass
class ResistConfig:
    """The resist fields the imaging path reads (core/config.py:634-672, checks :686-709):
    tone, develop threshold and PEB diffusion length (the threshold develop model and the
    ILT objective), film thickness, film index and the film plane ``OpticsConfig.defocus``
    refers to (the depth-resolved image).  The reference's full ``ResistConfig``
    (chemistry, develop rates) is accepted by duck typing wherever this one is."""

    tone: str = "positive"
    threshold: float = 0.30
    diffusion_sigma: float = 20e-9
    thickness: float = 100e-9
    n_resist: float = 1.70
    focus_reference: str = "mid"

    def __post_init__(self) -> None:
        r = "ResistConfig"
        _choice(r, "tone", self.tone, TONES)
        _bounded(r, "threshold", self.threshold, 0.0, 1.0, lo_open=True, hi_open=True)
        _bounded(r, "diffusion_sigma", self.diffusion_sigma, 0.0)
        _bounded(r, "thickness", self.thickness, 0.0)
        _bounded(r, "n_resist", self.n_resist, 0.0, lo_open=True)
        _choice(r, "focus_reference", self.focus_reference, FOCUS_REFERENCES)
