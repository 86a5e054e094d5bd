"""The reflective (EUV multilayer) mask model: thin-mask spectrum, each order
reflected by the blank at its own angle (expose/m3d/provider.py:190-293).

Every source point s gets the spectrum M(f) r(sin theta_s(f)), where the
mask-side direction sine of order f under illumination s is

    sin_s(f) = | (sin_x(s) + lambda f_x / m, sin_y(s) + lambda f_y / m) |,
    (sin_x, sin_y)(s) = -lambda f_s / m + (chief-ray sine on its axis)

(provider.py:72-122, m = reduction), and r is the blank's amplitude
reflection, the mean of its s and p values, interpolated in angle from a
table of 501 angles over [0, 25 deg] (magnitude and unwrapped phase
separately; zero beyond sin = 1).  On the device r multiplies each system's
transfer table on its box, and the clear field of a source point is scaled by
|r|^2 at its specular order (aerial_image.py:425-438).

The blank is a planar film stack, solved here once per plan on the host by the
characteristic-matrix method (Macleod, Thin-Film Optical Filters, ch. 2):
per layer [[cos d, i sin d / eta], [i eta sin d, cos d]] with
d = 2 pi n cos(theta) t / lambda and tilted admittance eta = n cos(theta) (s) or
n / cos(theta) (p), the stored index n + ik conjugated for propagation
(coat/films.py:88-213).
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Any

import numpy as np

__all__ = ["MASK_MATERIALS", "mask_material", "Multilayer", "MaskStack", "MASK_STACK_PRESETS",
           "reflection_table", "mask_side_sin_theta", "SOURCE_TILT_SIGN"]

# The engine shifts the pupil instead of the spectrum, which reverses the handedness of
# the illumination tilt (aerial_image.py:73).
SOURCE_TILT_SIGN = -1.0

# Angle table of the blank reflection (provider.py:63-64).
THETA_TABLE_MAX_DEG = 25.0
THETA_TABLE_N = 501

# n + ik, positive k absorbing, by wavelength (expose/m3d/materials.py:41-66; the EUV
# values are the CXRO/Henke constants at 13.5 nm as the reference tabulates them).
MASK_MATERIALS: dict[float, dict[str, complex]] = {
    13.5e-9: {
        "vacuum": 1.0 + 0.0j,
        "Si": 0.99931 + 0.00182j,
        "Mo": 0.92108 + 0.00644j,
        "Ru": 0.88651 + 0.01710j,
        "TaBN": 0.94900 + 0.03120j,
        "TaBO": 0.97000 + 0.02500j,
        "Ta": 0.94270 + 0.04170j,
        "SiO2": 0.97800 + 0.01070j,
    },
    193e-9: {
        "vacuum": 1.0 + 0.0j,
        "air": 1.0 + 0.0j,
        "quartz": 1.5603 + 0.0j,
        "Cr": 0.84 + 1.65j,
        "CrON": 1.95 + 1.05j,
        "MoSi": 2.343 + 0.586j,
    },
}


def mask_material(name: str, wavelength: float, tol: float = 1e-11) -> complex:
    """materials.py:69-102: the tabulated index of ``name`` at ``wavelength``."""
    for lam, table in MASK_MATERIALS.items():
        if abs(lam - wavelength) <= tol:
            if name not in table:
                raise KeyError(f"Unknown mask material '{name}' at {lam * 1e9:.1f} nm. "
                               f"Known: {sorted(table)}")
            return table[name]
    known = ", ".join(f"{w * 1e9:.1f} nm" for w in sorted(MASK_MATERIALS))
    raise KeyError(f"No mask-material table at {wavelength * 1e9:.3f} nm. Tabulated: {known}")


@dataclass(frozen=True)
class Multilayer:
    """multilayer.py:45-152: a Mo/Si Bragg mirror with a capping layer."""

    n_bilayers: int = 40
    period: float = 6.94e-9
    gamma: float = 0.4
    cap_material: str = "Ru"
    cap_thickness: float = 2.5e-9
    substrate: str = "Si"

    def __post_init__(self) -> None:
        if self.n_bilayers < 1:
            raise ValueError(f"n_bilayers must be >= 1, got {self.n_bilayers}")
        if not 0.0 < self.gamma < 1.0:
            raise ValueError(f"gamma must be in (0, 1), got {self.gamma}")
        if self.period <= 0.0:
            raise ValueError(f"period must be positive, got {self.period}")

    def layers(self, wavelength: float):
        """Ambient, cap, n_bilayers x (Si, Mo), substrate: (indices, thicknesses)."""
        t_mo = self.period * self.gamma
        t_si = self.period - t_mo
        n = [mask_material("vacuum", wavelength),
             mask_material(self.cap_material, wavelength)]
        t = [0.0, self.cap_thickness]
        n_si, n_mo = mask_material("Si", wavelength), mask_material("Mo", wavelength)
        for _ in range(self.n_bilayers):
            n += [n_si, n_mo]
            t += [t_si, t_mo]
        n.append(mask_material(self.substrate, wavelength))
        t.append(0.0)
        return np.asarray(n, dtype=complex), np.asarray(t, dtype=float)

    @classmethod
    def from_spec(cls, spec: dict[str, Any] | None) -> "Multilayer":
        return cls() if not spec else cls(**spec)


REGIMES = ("euv_reflective", "duv_transmissive")


@dataclass(frozen=True)
class MaskStack:
    """stack.py:44-160: an absorber on a substrate; the open field is the blank."""

    regime: str = "euv_reflective"
    absorber: str = "TaBN"
    absorber_thickness: float = 60e-9
    sidewall_deg: float = 90.0
    multilayer: Multilayer | None = field(default_factory=Multilayer)
    substrate: str = "SiO2"

    def __post_init__(self) -> None:
        if self.regime not in REGIMES:
            raise ValueError(f"regime must be one of {list(REGIMES)}, got '{self.regime}'")
        if self.absorber_thickness < 0.0:
            raise ValueError(f"absorber_thickness must be >= 0, got {self.absorber_thickness}")
        if not 0.0 < self.sidewall_deg <= 90.0:
            raise ValueError(f"sidewall_deg must be in (0, 90], got {self.sidewall_deg}")
        if self.is_reflective and self.multilayer is None:
            raise ValueError("regime='euv_reflective' needs a multilayer — a reflective mask "
                             "with no mirror reflects nothing.")

    @property
    def is_reflective(self) -> bool:
        return self.regime == "euv_reflective"

    def blank_layers(self, wavelength: float):
        """The stack where there is no absorber (stack.py:106-125)."""
        if self.is_reflective:
            return self.multilayer.layers(wavelength)
        n = [mask_material("vacuum", wavelength), mask_material(self.substrate, wavelength),
             mask_material(self.substrate, wavelength)]
        return np.asarray(n, dtype=complex), np.zeros(3)

    def blank_reflection(self, wavelength: float, theta: float = 0.0,
                         polarisation: str = "s") -> complex:
        n, t = self.blank_layers(wavelength)
        return complex(stack_reflection(n, t, wavelength, np.array([theta]), polarisation)[0])

    @classmethod
    def from_spec(cls, spec: dict[str, Any] | None) -> "MaskStack":
        if not spec:
            return cls()
        data = dict(spec)
        ml = data.pop("multilayer", None)
        return cls(multilayer=Multilayer.from_spec(ml) if ml else None, **data)

    @classmethod
    def for_wavelength(cls, wavelength: float) -> "MaskStack":
        if abs(wavelength - 13.5e-9) < 1e-11:
            return cls(**MASK_STACK_PRESETS["EUV TaBN 60nm"])
        return cls(**MASK_STACK_PRESETS["DUV binary Cr"])


# stack.py:176-205
MASK_STACK_PRESETS: dict[str, dict[str, Any]] = {
    "EUV TaBN 60nm": {"regime": "euv_reflective", "absorber": "TaBN",
                      "absorber_thickness": 60e-9, "substrate": "SiO2"},
    "EUV TaBN 45nm": {"regime": "euv_reflective", "absorber": "TaBN",
                      "absorber_thickness": 45e-9, "substrate": "SiO2"},
    "DUV binary Cr": {"regime": "duv_transmissive", "absorber": "Cr",
                      "absorber_thickness": 70e-9, "multilayer": None, "substrate": "quartz"},
    "DUV att-PSM 6%": {"regime": "duv_transmissive", "absorber": "MoSi",
                       "absorber_thickness": 71.9e-9, "multilayer": None,
                       "substrate": "quartz"},
}


def stack_reflection(n_stored, thickness, wavelength: float, theta0, polarisation: str):
    """Amplitude reflection of a planar stack (ambient first, substrate last) seen from
    the top of its first real layer, for every incidence angle in ``theta0`` at once."""
    if not wavelength > 0:
        raise ValueError(f"wavelength must be positive, got {wavelength}")
    if polarisation not in ("s", "p"):
        raise ValueError(f"polarisation must be 's' or 'p', got '{polarisation}'")
    n = np.conj(np.asarray(n_stored, dtype=complex))          # propagation index n - ik
    th = np.asarray(theta0, dtype=float)
    kx = n[0] * np.sin(th)                                     # conserved tangential index
    cos_t = np.sqrt(1.0 - (kx[:, None] / n[None, :]) ** 2)     # (angles, layers)
    eta = n[None, :] * cos_t if polarisation == "s" else n[None, :] / cos_t
    # product of the characteristic matrices of layers 1 .. L-2, applied to [1, eta_sub]
    m00 = np.ones(th.shape, dtype=complex)
    m01 = np.zeros(th.shape, dtype=complex)
    m10 = np.zeros(th.shape, dtype=complex)
    m11 = np.ones(th.shape, dtype=complex)
    for j in range(1, len(n) - 1):
        d = 2.0 * np.pi * n[j] * cos_t[:, j] * thickness[j] / wavelength
        c, s = np.cos(d), np.sin(d)
        a01 = 1j * s / eta[:, j]
        a10 = 1j * eta[:, j] * s
        m00, m01, m10, m11 = (m00 * c + m01 * a10, m00 * a01 + m01 * c,
                              m10 * c + m11 * a10, m10 * a01 + m11 * c)
    b = m00 + m01 * eta[:, -1]
    cc = m10 + m11 * eta[:, -1]
    y = cc / b
    return (eta[:, 0] - y) / (eta[:, 0] + y)


def reflection_table(stack: MaskStack, wavelength: float):
    """(theta [rad], |r|, unwrapped arg r) of the blank over the angle table, r the mean
    of the s and p reflections (provider.py:226-250)."""
    n, t = stack.blank_layers(wavelength)
    theta = np.linspace(0.0, np.deg2rad(THETA_TABLE_MAX_DEG), THETA_TABLE_N)
    r = 0.5 * (stack_reflection(n, t, wavelength, theta, "s")
               + stack_reflection(n, t, wavelength, theta, "p"))
    return theta, np.abs(r), np.unwrap(np.angle(r))


def mask_side_sin_theta(fs_x: float, fs_y: float, optics) -> tuple[float, float]:
    """provider.py:72-122: mask-side direction sines of one source point's illumination."""
    lam, m = optics.wavelength, optics.reduction
    sin_x = SOURCE_TILT_SIGN * lam * fs_x / m
    sin_y = SOURCE_TILT_SIGN * lam * fs_y / m
    cra = float(np.sin(np.radians(optics.chief_ray_deg)))
    axis = optics.chief_ray_axis.lower()
    if axis == "x":
        sin_x += cra
    elif axis == "y":
        sin_y += cra
    else:
        raise ValueError(f"chief_ray_axis must be 'x' or 'y', got '{optics.chief_ray_axis}'")
    return sin_x, sin_y


def stack_for(optics) -> MaskStack:
    """The mask stack an optics configuration names (provider.py:529-543)."""
    spec = optics.mask_stack
    return MaskStack.from_spec(spec) if spec is not None else MaskStack.for_wavelength(
        optics.wavelength)
