"""Float64 NumPy/SciPy restatement of the reference's partially coherent imaging path.

TEST INFRASTRUCTURE ONLY.  This module is the checker: only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import it.  The product (``lithobat_amd``) never imports, calls or falls back to
anything here.

Every function restates the algorithm of one reference function and cites it
(paths relative to the reference checkout, ``file:line``).  The arithmetic is
float64 / complex128 throughout, with the reference's own FFT routines
(``numpy.fft`` and ``scipy.fft`` with ``workers=-1``) and its 64 MiB transform
batching, so that the timed CPU baseline is the reference algorithm.

Pinning (what checks each function independently of this file):

* ``make_freq_grid``, ``pupil_grid``: closed forms (``fftfreq`` definition,
  ``linspace``) -- tests/test_oracle.py.
* source rasterisers / ``source_points``: unit sum, annulus bounds, rows = eta,
  golden source tables written by the imported reference (tests/golden).
* ``zernike_noll``, ``defocus_opd``, ``vector_coefficients``: closed forms
  (Z1..Z4, paraxial OPD, two-beam TE/TM laws, |v_z| = sin theta).
* ``abbe_sum``: brute-force direct DFT on tiny grids (no FFT), tilt-sign
  equivalence, clear field = 1, planes == singles, golden images from the
  imported reference (defocused / aberrated scalar and every vector state).
* ``socs_build`` / ``socs_raw``: full rank == Abbe (1e-12), eigen-reconstruction
  of the TCC, golden eigenvalues from the imported reference.
* ``socs_gradient``, ``abbe_gradient``, ``objective``: central finite
  differences, and the reference's own outputs (tests/golden/ilt32.npz) for
  ``objective``, ``socs_gradient``, ``mask_of`` / ``grad_theta``,
  ``dose_for_target``, ``run_ilt`` (history, mask, theta, print) and
  ``exposure_volume`` (every focus reference).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import Any

import numpy as np
from scipy import fft as sfft

# --------------------------------------------------------------------------
# Configuration carrier (attribute names follow core/config.py:383-406)
# --------------------------------------------------------------------------


@dataclass
class Optics:
    """Duck-typed stand-in for ``OpticsConfig`` (core/config.py:238-487).

    Only the fields the imaging path reads.  Any object with the same
    attribute names (including ``lithobat_amd.config.OpticsConfig``) works
    wherever the oracle takes ``optics``.
    """

    wavelength: float = 193e-9
    NA: float = 0.93
    n_immersion: float = 1.0
    defocus: float = 0.0
    sigma_outer: float = 0.80
    sigma_inner: float = 0.0
    source_type: str = "conventional"
    source_grid: int = 21
    source_kwargs: dict = field(default_factory=dict)
    zernike_coeffs: dict = field(default_factory=dict)
    exact_defocus: bool = False
    n_image: float | None = None
    imaging_model: str = "scalar"
    polarisation: str = "unpolarised"
    obliquity: bool = True
    normalisation: str = "peak"

    @property
    def image_index(self) -> float:  # core/config.py:469-477
        return self.n_immersion if self.n_image is None else self.n_image


@dataclass
class Grid:
    """Stand-in for ``GridConfig`` (core/config.py:720-759)."""

    n_pixels: int = 128
    pixel_size: float = 4e-9


def _image_index(optics: Any) -> float:
    n_image = getattr(optics, "n_image", None)
    return optics.n_immersion if n_image is None else n_image


# --------------------------------------------------------------------------
# Grids (core/utils.py:45-66, expose/pupil.py:35-48)
# --------------------------------------------------------------------------


def make_freq_grid(n: int, pixel_size: float):
    """core/utils.py:45-66 -- fftfreq meshgrid, FX varies along columns."""
    f = np.fft.fftfreq(n, d=pixel_size)
    fx, fy = np.meshgrid(f, f)
    return fx.astype(np.float64), fy.astype(np.float64)


def pupil_grid(n: int):
    """expose/pupil.py:35-48 -- (rho, phi, xi, eta), rows = eta, cols = xi."""
    c = np.linspace(-1.0, 1.0, n)
    xi, eta = np.meshgrid(c, c)
    return np.sqrt(xi ** 2 + eta ** 2), np.arctan2(eta, xi), xi, eta


# --------------------------------------------------------------------------
# Sources (expose/illumination.py:21-219)
# --------------------------------------------------------------------------


def _unit_sum(a):
    """illumination.py:21-23."""
    s = a.sum()
    return a / s if s > 0 else a


def source_conventional(n, sigma):
    """illumination.py:26-42."""
    rho = pupil_grid(n)[0]
    return _unit_sum((rho <= sigma).astype(np.float64))


def source_annular(n, sigma_outer, sigma_inner):
    """illumination.py:45-70."""
    if sigma_inner >= sigma_outer:
        raise ValueError("sigma_inner must be < sigma_outer")
    rho = pupil_grid(n)[0]
    return _unit_sum(((rho >= sigma_inner) & (rho <= sigma_outer)).astype(np.float64))


def _disc_poles(n, sigma_outer, sigma_inner, centres_of):
    radius = (sigma_outer - sigma_inner) / 2.0
    mid = (sigma_outer + sigma_inner) / 2.0
    _, _, xi, eta = pupil_grid(n)
    acc = np.zeros((n, n))
    for cx, cy in centres_of(mid):
        acc += (np.sqrt((xi - cx) ** 2 + (eta - cy) ** 2) <= radius).astype(np.float64)
    return _unit_sum(acc)


def source_dipole(n, sigma_outer, sigma_inner, axis="x"):
    """illumination.py:73-98 (legacy disc poles)."""
    def centres(c):
        return [(c, 0.0), (-c, 0.0)] if axis == "x" else [(0.0, c), (0.0, -c)]
    return _disc_poles(n, sigma_outer, sigma_inner, centres)


def source_quadrupole(n, sigma_outer, sigma_inner, rotation_deg=45.0):
    """illumination.py:101-129 (legacy disc poles, rotated)."""
    ang = np.deg2rad(np.array([0.0, 90.0, 180.0, 270.0]) + rotation_deg)

    def centres(c):
        return [(c * np.cos(a), c * np.sin(a)) for a in ang]
    return _disc_poles(n, sigma_outer, sigma_inner, centres)


def build_source(n, source_type, sigma_outer, sigma_inner=0.0, **kw):
    """illumination.py:132-219 for the four analytic shapes."""
    key = source_type.lower()
    if key == "conventional":
        return source_conventional(n, sigma_outer)
    if key == "annular":
        return source_annular(n, sigma_outer, sigma_inner)
    if key == "dipole":
        return source_dipole(n, sigma_outer, sigma_inner, **kw)
    if key == "quadrupole":
        return source_quadrupole(n, sigma_outer, sigma_inner, **kw)
    raise ValueError(f"oracle has no rasteriser for source type {source_type!r}")


@dataclass(frozen=True)
class SourcePts:
    """aerial_image.py:249-272."""

    weight: np.ndarray
    fx: np.ndarray
    fy: np.ndarray
    phi: np.ndarray
    grid: int

    def __len__(self):
        return int(self.weight.size)


def source_points(optics) -> SourcePts:
    """aerial_image.py:275-304: non-zero source pixels -> (w, fx, fy, phi)."""
    ns = max(int(optics.source_grid), 3)
    src = build_source(ns, optics.source_type, optics.sigma_outer, optics.sigma_inner,
                       **(getattr(optics, "source_kwargs", None) or {}))
    idx = np.argwhere(src > 0)
    rows = idx[:, 0].astype(np.float64)
    cols = idx[:, 1].astype(np.float64)
    half = (ns - 1) / 2.0
    k_max = optics.NA / optics.wavelength
    return SourcePts(
        weight=src[idx[:, 0], idx[:, 1]].astype(np.float64),
        fx=((cols - half) / half) * k_max,
        fy=((rows - half) / half) * k_max,
        phi=np.arctan2(rows - half, cols - half),
        grid=ns,
    )


# --------------------------------------------------------------------------
# Pupil physics (expose/pupil.py:56-369)
# --------------------------------------------------------------------------

NOLL = {  # pupil.py:56-78, Noll j -> (n, m)
    1: (0, 0), 2: (1, 1), 3: (1, -1), 4: (2, 0), 5: (2, -2), 6: (2, 2),
    7: (3, -1), 8: (3, 1), 9: (3, -3), 10: (3, 3), 11: (4, 0), 12: (4, 2),
    13: (4, -2), 14: (4, 4), 15: (4, -4), 16: (5, 1), 17: (5, -1), 18: (5, 3),
    19: (5, -3), 20: (5, 5), 21: (5, -5),
}


def radial_poly(n, m, rho):
    """pupil.py:81-100 -- R_n^|m| by the factorial sum, rho clipped to [0,1]."""
    ma = abs(m)
    r = np.clip(rho, 0.0, 1.0)
    out = np.zeros_like(r, dtype=np.float64)
    for s in range((n - ma) // 2 + 1):
        c = ((-1) ** s * math.factorial(n - s)
             / (math.factorial(s) * math.factorial((n + ma) // 2 - s)
                * math.factorial((n - ma) // 2 - s)))
        out += c * r ** (n - 2 * s)
    return out


def zernike_noll(j, rho, phi):
    """pupil.py:103-141 -- unknown j gives zeros."""
    if j not in NOLL:
        return np.zeros_like(rho, dtype=np.float64)
    n, m = NOLL[j]
    norm = math.sqrt(n + 1) if m == 0 else math.sqrt(2.0 * (n + 1))
    R = radial_poly(n, m, rho)
    if m > 0:
        return norm * R * np.cos(m * phi)
    if m < 0:
        return norm * R * np.sin(abs(m) * phi)
    return norm * R


def defocus_opd(rho, wavelength, NA, defocus, n_image=1.0, exact=False):
    """pupil.py:155-221 -- phase [rad]; exactly zero when defocus == 0."""
    if defocus == 0.0:
        return np.zeros_like(rho, dtype=np.float64)
    rho = np.asarray(rho, dtype=np.float64)
    if not exact:
        return np.pi * NA ** 2 * rho ** 2 * defocus / (wavelength * n_image)
    sin_sq = np.clip((rho * NA / n_image) ** 2, 0.0, 1.0)
    return (2.0 * np.pi / wavelength) * (n_image * defocus * (1.0 - np.sqrt(1.0 - sin_sq)))


def jones_states(polarisation, phi_source):
    """pupil.py:232-289 -- list of (Jx, Jy, weight)."""
    p = polarisation.lower()
    if p == "x":
        return [(1.0, 0.0, 1.0)]
    if p == "y":
        return [(0.0, 1.0, 1.0)]
    if p == "te":
        return [(-math.sin(phi_source), math.cos(phi_source), 1.0)]
    if p == "tm":
        return [(math.cos(phi_source), math.sin(phi_source), 1.0)]
    if p in ("unpolarised", "unpolarized"):
        return [(1.0, 0.0, 0.5), (0.0, 1.0, 0.5)]
    raise ValueError(f"unknown polarisation {polarisation!r}")


def vector_coefficients(rho, phi, jx, jy, sin_theta_max, obliquity=True):
    """pupil.py:292-369 -- s/p projection, p tips by theta, sqrt(cos) obliquity."""
    sin_sq = np.clip((np.asarray(rho, dtype=np.float64) * sin_theta_max) ** 2, 0.0, 1.0)
    cos_t = np.sqrt(1.0 - sin_sq)
    sin_t = np.sqrt(sin_sq)
    cp, sp = np.cos(phi), np.sin(phi)
    e_s = -jx * sp + jy * cp
    e_p = jx * cp + jy * sp
    vx = -e_s * sp + e_p * cos_t * cp
    vy = e_s * cp + e_p * cos_t * sp
    vz = -e_p * sin_t
    if obliquity:
        a = np.sqrt(cos_t)
        vx, vy, vz = vx * a, vy * a, vz * a
    return vx.astype(np.complex128), vy.astype(np.complex128), vz.astype(np.complex128)


def pupil_geometry(FX, FY, wavelength, NA, fs_x=0.0, fs_y=0.0, need_phi=False):
    """aerial_image.py:86-139 -- rho, phi (optional) and inside on the FFT grid."""
    k_max = NA / wavelength
    fx = FX - fs_x
    fy = FY - fs_y
    rho = np.sqrt(fx ** 2 + fy ** 2) / k_max
    phi = np.arctan2(fy, fx) if need_phi else None
    return rho, phi, rho <= 1.0


def pupil_phase(rho, phi, wavelength, NA, defocus, n_image, zernike, exact):
    """aerial_image.py:142-166 -- defocus OPD + sum 2 pi c_j Z_j."""
    W = defocus_opd(rho, wavelength, NA, defocus, n_image, exact=exact)
    if zernike:
        W = np.array(W, dtype=np.float64, copy=True)
        for j, c in zernike.items():
            W += 2.0 * np.pi * c * zernike_noll(j, rho, phi)
    return W


def build_fft_pupil(rho, phi, inside, wavelength, NA, defocus, n_image, zernike, exact=False):
    """aerial_image.py:169-209: exp(i W) on the in-pupil bins, zeros elsewhere, in the
    layout of the geometry arrays (FFT order when they come from make_freq_grid)."""
    idx = np.nonzero(inside)
    W = pupil_phase(rho[idx], None if phi is None else phi[idx], wavelength, NA, defocus,
                    n_image, zernike, exact)
    pupil = np.zeros(rho.shape, dtype=np.complex128)
    pupil[idx] = np.exp(1j * W)
    return pupil


def normalisation_scale(mode, peak, clear, dose=1.0):
    """aerial_image.py:217-241."""
    m = mode.lower()
    if m == "peak":
        return dose / peak if peak > 0 else dose
    if m == "clear":
        return dose / clear if clear > 0 else dose
    if m == "none":
        return dose
    raise ValueError(f"normalisation must be 'peak', 'clear' or 'none', got '{mode}'")


# --------------------------------------------------------------------------
# The Abbe sum (aerial_image.py:312-494)
# --------------------------------------------------------------------------

FFT_BATCH_BYTES = 64 * 2 ** 20  # aerial_image.py:78


class _Stack:
    """aerial_image.py:312-348: spectra batched into one scipy ifft2 call."""

    def __init__(self, aerial, n):
        self.aerial = aerial
        self.limit = max(1, FFT_BATCH_BYTES // (n * n * 16))
        self.specs, self.planes, self.weights = [], [], []

    def push(self, spec, plane, weight):
        self.specs.append(spec)
        self.planes.append(plane)
        self.weights.append(weight)
        if len(self.specs) >= self.limit:
            self.flush()

    def flush(self):
        if not self.specs:
            return
        fields = sfft.ifft2(np.stack(self.specs), axes=(-2, -1), workers=-1)
        inten = fields.real ** 2 + fields.imag ** 2
        planes = np.asarray(self.planes)
        w = np.asarray(self.weights, dtype=np.float64)
        for p in np.unique(planes):
            sel = planes == p
            self.aerial[p] += np.tensordot(w[sel], inten[sel], axes=1)
        self.specs, self.planes, self.weights = [], [], []


# --------------------------------------------------------------------------
# Reflective (EUV multilayer) mask spectra (expose/m3d/provider.py:190-293,
# multilayer.py, stack.py, materials.py, coat/films.py:88-213)
# --------------------------------------------------------------------------

_EUV_INDEX = {  # materials.py:45-56, n + ik at 13.5 nm
    "vacuum": 1.0 + 0.0j, "Si": 0.99931 + 0.00182j, "Mo": 0.92108 + 0.00644j,
    "Ru": 0.88651 + 0.01710j, "TaBN": 0.94900 + 0.03120j, "TaBO": 0.97000 + 0.02500j,
    "Ta": 0.94270 + 0.04170j, "SiO2": 0.97800 + 0.01070j,
}


def blank_films(optics):
    """stack.py:106-125 with multilayer.py:95-120: (index, thickness) from the ambient
    down to the substrate of the EUV blank the optics name (default 40 x Mo/Si, Ru cap)."""
    spec = optics.mask_stack or {}
    if spec.get("regime", "euv_reflective") != "euv_reflective":
        raise ValueError("multilayer mask model needs a reflective stack")
    ml = dict(n_bilayers=40, period=6.94e-9, gamma=0.4, cap_material="Ru",
              cap_thickness=2.5e-9, substrate="Si")
    ml.update(spec.get("multilayer") or {})
    t_mo = ml["period"] * ml["gamma"]
    films = [(_EUV_INDEX["vacuum"], 0.0), (_EUV_INDEX[ml["cap_material"]], ml["cap_thickness"])]
    for _ in range(ml["n_bilayers"]):
        films += [(_EUV_INDEX["Si"], ml["period"] - t_mo), (_EUV_INDEX["Mo"], t_mo)]
    films.append((_EUV_INDEX[ml["substrate"]], 0.0))
    return films


def film_reflection(films, lam, theta0, pol):
    """films.py:168-213: characteristic-matrix reflection seen from the ambient, one angle."""
    n = [np.conj(complex(f[0])) for f in films]
    sin0 = n[0] * np.sin(theta0)
    cos_t = [np.sqrt(1.0 - (sin0 / nj) ** 2) for nj in n]
    eta = [nj * c if pol == "s" else nj / c for nj, c in zip(n, cos_t)]
    Mt = np.eye(2, dtype=complex)
    for j in range(1, len(films) - 1):
        dlt = 2.0 * np.pi * n[j] * cos_t[j] * films[j][1] / lam
        Mt = Mt @ np.array([[np.cos(dlt), 1j * np.sin(dlt) / eta[j]],
                            [1j * eta[j] * np.sin(dlt), np.cos(dlt)]])
    b, c = Mt @ np.array([1.0, eta[-1]])
    y = c / b
    return complex((eta[0] - y) / (eta[0] + y))


def reflection_table(optics):
    """provider.py:226-250: 501 angles over [0, 25 deg]; |r| and unwrapped arg r of the
    mean of the s and p reflections."""
    films = blank_films(optics)
    theta = np.linspace(0.0, np.deg2rad(25.0), 501)
    r = np.array([0.5 * (film_reflection(films, optics.wavelength, t, "s")
                         + film_reflection(films, optics.wavelength, t, "p")) for t in theta])
    return theta, np.abs(r), np.unwrap(np.angle(r))


def blank_reflection_at(sin_theta, table):
    """provider.py:252-266: r at each order's direction sine, zero when evanescent."""
    theta, mag, phase = table
    s = np.asarray(sin_theta, dtype=np.float64)
    ang = np.arcsin(np.clip(s, 0.0, 1.0))
    r = np.interp(ang, theta, mag) * np.exp(1j * np.interp(ang, theta, phase))
    return np.where(s <= 1.0, r, 0.0)


def mask_side_sines(fs_x, fs_y, optics):
    """provider.py:72-122 (SOURCE_TILT_SIGN = -1, aerial_image.py:73)."""
    sx = -optics.wavelength * fs_x / optics.reduction
    sy = -optics.wavelength * fs_y / optics.reduction
    cra = math.sin(math.radians(optics.chief_ray_deg))
    if optics.chief_ray_axis.lower() == "x":
        sx += cra
    else:
        sy += cra
    return sx, sy


def abbe_sum(mask, optics, grid, planes, src=None):
    """aerial_image.py:351-494: raw (P,n,n), peak (P,), clear (P,); thin mask, or the
    reflective multilayer model (per-source spectrum and clear scale, :421-438)."""
    n = grid.n_pixels
    lam, NA = optics.wavelength, optics.NA
    model = optics.imaging_model.lower()
    if model not in ("scalar", "vector"):
        raise ValueError(f"imaging_model must be 'scalar' or 'vector', got '{optics.imaging_model}'")
    vector = model == "vector"
    exact = bool(optics.exact_defocus) or vector
    zern = optics.zernike_coeffs or None
    need_phi = vector or bool(zern)
    if src is None:
        src = source_points(optics)
    M = np.fft.fft2(np.asarray(mask).astype(np.complex128))  # provider.py:177-178
    FX, FY = make_freq_grid(n, grid.pixel_size)
    multilayer = getattr(optics, "mask_model", "thin") == "multilayer"
    if multilayer:
        table = reflection_table(optics)
        dx = optics.wavelength * FX / optics.reduction
        dy = optics.wavelength * FY / optics.reduction
    P = len(planes)
    aerial = np.zeros((P, n, n))
    clear = np.zeros(P)
    stack = _Stack(aerial, n)
    for s in range(len(src)):
        w = float(src.weight[s])
        rho, phi, inside = pupil_geometry(FX, FY, lam, NA, float(src.fx[s]), float(src.fy[s]),
                                          need_phi=need_phi)
        idx = np.nonzero(inside)
        rho_in = rho[idx]
        phi_in = None if phi is None else phi[idx]
        scale_clear = 1.0
        if multilayer:
            sx, sy = mask_side_sines(float(src.fx[s]), float(src.fy[s]), optics)
            M_s = M * blank_reflection_at(np.sqrt((sx + dx) ** 2 + (sy + dy) ** 2), table)
            scale_clear = float(abs(blank_reflection_at(np.array([math.hypot(sx, sy)]),
                                                        table)[0]) ** 2)
            M_in = M_s[idx]
        else:
            M_in = M[idx]
        dc_in = float(inside[0, 0])
        states = jones_states(optics.polarisation, float(src.phi[s])) if vector else None
        for p, (dz, n_img) in enumerate(planes):
            Wp = pupil_phase(rho_in, phi_in, lam, NA, dz, n_img, zern, exact)
            coh = M_in * np.exp(1j * Wp)
            if not vector:
                spec = np.zeros((n, n), dtype=np.complex128)
                spec[idx] = coh
                stack.push(spec, p, w)
                clear[p] += w * dc_in * scale_clear
                continue
            for jx, jy, wp in states:
                V = vector_coefficients(rho_in, phi_in, jx, jy, NA / n_img,
                                        obliquity=optics.obliquity)
                for comp in V:
                    spec = np.zeros((n, n), dtype=np.complex128)
                    spec[idx] = coh * comp
                    stack.push(spec, p, w * wp)
                if dc_in:
                    at = np.nonzero((idx[0] == 0) & (idx[1] == 0))[0][0]
                    clear[p] += (w * wp) * sum(abs(c[at]) ** 2 for c in V) * scale_clear
    stack.flush()
    return aerial, aerial.reshape(P, -1).max(axis=1), clear


def compute_aerial_image(mask, optics, grid, dose=1.0, return_raw=False):
    """aerial_image.py:502-548."""
    raw, peak, clear = abbe_sum(mask, optics, grid, [(optics.defocus, _image_index(optics))])
    if return_raw:
        return raw[0], float(peak[0]), float(clear[0])
    return raw[0] * normalisation_scale(optics.normalisation, float(peak[0]), float(clear[0]), dose)


def compute_aerial_planes(mask, optics, grid, planes, dose=1.0):
    """aerial_image.py:551-587."""
    raw, peak, clear = abbe_sum(mask, optics, grid, planes)
    for p in range(raw.shape[0]):
        raw[p] *= normalisation_scale(optics.normalisation, float(peak[p]), float(clear[p]), dose)
    return raw


def abbe_direct_dft(mask, optics, grid, dz=None, n_img=None):
    """Independent brute force: I(x) = sum_s w_s |(1/N^2) sum_f M(f) P_s(f) e^{2 pi i f.x}|^2.

    Direct double sums with explicit exponentials -- no FFT anywhere -- so it
    pins the FFT conventions and normalisation of ``abbe_sum``.  O(N^4 S): tiny
    grids only.  Scalar model.
    """
    n = grid.n_pixels
    dz = optics.defocus if dz is None else dz
    n_img = _image_index(optics) if n_img is None else n_img
    m = np.asarray(mask, dtype=np.complex128)
    j = np.arange(n)
    F = np.exp(-2j * np.pi * np.outer(j, j) / n)  # forward DFT matrix
    Mf = F @ m @ F.T
    Fi = np.conj(F) / n
    src = source_points(optics)
    FX, FY = make_freq_grid(n, grid.pixel_size)
    zern = optics.zernike_coeffs or None
    out = np.zeros((n, n))
    for s in range(len(src)):
        rho, phi, inside = pupil_geometry(FX, FY, optics.wavelength, optics.NA,
                                          src.fx[s], src.fy[s], need_phi=True)
        W = pupil_phase(rho, phi, optics.wavelength, optics.NA, dz, n_img, zern,
                        bool(optics.exact_defocus))
        Pp = np.where(inside, np.exp(1j * W), 0.0)
        E = Fi @ (Mf * Pp) @ Fi.T
        out += src.weight[s] * np.abs(E) ** 2
    return out


# --------------------------------------------------------------------------
# SOCS (expose/hopkins.py:103-223)
# --------------------------------------------------------------------------


@dataclass
class SOCS:
    kernels: np.ndarray
    weights: np.ndarray
    clear: float
    energy_kept: float
    n_band: int
    n_source: int
    n: int


def socs_build(optics, grid, tol=1e-8, max_kernels=None):
    """hopkins.py:103-198: TCC = A^T conj(A) on the 2 NA/lambda band, eigh, truncate."""
    n = grid.n_pixels
    lam, NA = optics.wavelength, optics.NA
    k_max = NA / lam
    src = source_points(optics)
    FX, FY = make_freq_grid(n, grid.pixel_size)
    band = np.nonzero(np.hypot(FX, FY) <= 2.0 * k_max + 1e-9 * k_max)
    fxb, fyb = FX[band], FY[band]
    zern = optics.zernike_coeffs or None
    A = np.zeros((len(src), fxb.size), dtype=np.complex128)
    for s in range(len(src)):
        dx, dy = fxb - src.fx[s], fyb - src.fy[s]
        rho = np.hypot(dx, dy) / k_max
        ins = rho <= 1.0
        if not ins.any():
            continue
        phi = np.arctan2(dy[ins], dx[ins]) if zern else None
        W = pupil_phase(rho[ins], phi, lam, NA, optics.defocus, _image_index(optics), zern,
                        optics.exact_defocus)
        A[s, ins] = np.sqrt(src.weight[s]) * np.exp(1j * W)
    tcc = A.T @ A.conj()
    ev, evec = np.linalg.eigh(tcc)
    order = np.argsort(ev)[::-1]
    ev, evec = ev[order], evec[:, order]
    trace = float(np.sum(np.clip(ev, 0.0, None)))
    keep = ev > (tol * ev[0] if ev.size else 0.0)
    keep &= ev > 0.0
    if max_kernels is not None:
        keep[max_kernels:] = False
    lam_k, vec_k = ev[keep], evec[:, keep]
    K = int(lam_k.size)
    kern = np.zeros((K, n, n), dtype=np.complex128)
    for k in range(K):
        kern[k][band] = vec_k[:, k]
    dc = np.nonzero((band[0] == 0) & (band[1] == 0))[0]
    clear = float(np.sum(lam_k * np.abs(vec_k[dc[0], :]) ** 2)) if dc.size else 0.0
    kept = float(lam_k.sum() / trace) if trace > 0 else 1.0
    return SOCS(kern, lam_k, clear, kept, int(fxb.size), len(src), n)


def socs_raw(k: SOCS, mask):
    """hopkins.py:205-218."""
    Mf = np.fft.fft2(np.asarray(mask, dtype=np.complex128))
    f = sfft.ifft2(k.kernels * Mf[None], axes=(-2, -1), workers=-1)
    inten = np.tensordot(k.weights, f.real ** 2 + f.imag ** 2, axes=1)
    return inten, float(inten.max()), k.clear


# --------------------------------------------------------------------------
# ILT forward / adjoint / objective (opc/ilt.py:208-389)
# --------------------------------------------------------------------------


def socs_fields(k: SOCS, mask):
    """ilt.py:208-232."""
    spec = np.fft.fft2(np.asarray(mask, dtype=np.complex128))
    fields = np.fft.ifft2(k.kernels * spec[None], axes=(-2, -1))
    return fields, np.tensordot(k.weights, fields.real ** 2 + fields.imag ** 2, axes=1)


def socs_gradient(k: SOCS, fields, sensitivity):
    """ilt.py:235-266: 2 Re sum_k lam_k ifft2(conj(phi_k) fft2(g E_k))."""
    g = np.asarray(sensitivity, dtype=np.float64)
    back = np.fft.fft2(g[None] * fields)
    adj = np.fft.ifft2(np.conj(k.kernels) * back, axes=(-2, -1))
    return 2.0 * np.real(np.tensordot(k.weights, adj, axes=1))


def abbe_gradient(mask, optics, grid, sensitivity, planes=None, src=None):
    """d/dm of sum_x g(x) I_raw(x) for the scalar/vector Abbe sum.

    The same adjoint as ilt.py:262-266 with the source points in place of the
    eigen-kernels: 2 Re sum_j w_j ifft2(conj(H_j) fft2(g E_j)), summed over
    planes when ``sensitivity`` is (P, n, n).  Pinned by central finite
    differences, and at full rank by ``socs_gradient``.  ``src``: the source points
    (default: the optics' own).
    """
    n = grid.n_pixels
    if planes is None:
        planes = [(optics.defocus, _image_index(optics))]
    g_all = np.asarray(sensitivity, dtype=np.float64).reshape(len(planes), n, n)
    lam, NA = optics.wavelength, optics.NA
    vector = optics.imaging_model.lower() == "vector"
    exact = bool(optics.exact_defocus) or vector
    zern = optics.zernike_coeffs or None
    if src is None:
        src = source_points(optics)
    Mf = np.fft.fft2(np.asarray(mask, dtype=np.complex128))
    FX, FY = make_freq_grid(n, grid.pixel_size)
    acc = np.zeros((n, n), dtype=np.complex128)
    for s in range(len(src)):
        rho, phi, inside = pupil_geometry(FX, FY, lam, NA, src.fx[s], src.fy[s], need_phi=True)
        states = jones_states(optics.polarisation, float(src.phi[s])) if vector else [(0, 0, 1.0)]
        for p, (dz, n_img) in enumerate(planes):
            W = pupil_phase(rho, phi, lam, NA, dz, n_img, zern, exact)
            base = np.where(inside, np.exp(1j * W), 0.0)
            for jx, jy, wp in states:
                if vector:
                    comps = vector_coefficients(rho, phi, jx, jy, NA / n_img, optics.obliquity)
                    Hs = [np.where(inside, base * c, 0.0) for c in comps]
                else:
                    Hs = [base]
                for H in Hs:
                    E = np.fft.ifft2(Mf * H)
                    acc += src.weight[s] * wp * np.conj(H) * np.fft.fft2(g_all[p] * E)
    return 2.0 * np.real(np.fft.ifft2(acc))


def sigmoid_clipped(z):
    return 1.0 / (1.0 + np.exp(-np.clip(z, -60.0, 60.0)))


def blur(a, sigma_px):
    """ilt.py:274-289: scipy.ndimage.gaussian_filter(a, sigma) (mode "reflect",
    truncate 4.0, scipy 1.15.3 here), the input itself below 1e-3 px."""
    from scipy.ndimage import gaussian_filter
    if sigma_px < 1e-3:
        return a
    return gaussian_filter(a, sigma=sigma_px)


def blur_direct(a, sigma):
    """Independent restatement of that filter: taps exp(-k^2 / 2 sigma^2) / sum for
    |k| <= int(4 sigma + 0.5), half-sample symmetric extension, rows then columns,
    explicit index arithmetic (no scipy)."""
    a = np.asarray(a, dtype=np.float64)
    n0, n1 = a.shape
    R = int(4.0 * sigma + 0.5)
    k = np.arange(-R, R + 1)
    w = np.exp(-0.5 / (sigma * sigma) * k.astype(np.float64) ** 2)
    w /= w.sum()

    def refl(i, n):
        r = np.mod(i, 2 * n)
        return np.where(r < n, r, 2 * n - 1 - r)

    out = np.zeros_like(a)
    for t, wk in zip(k, w):
        out += wk * a[refl(np.arange(n0) + t, n0), :]
    res = np.zeros_like(a)
    for t, wk in zip(k, w):
        res += wk * out[:, refl(np.arange(n1) + t, n1)]
    return res


def relaxed_objective(raw, target, scale, level, steepness, sign=1.0, sigma_px=0.0):
    """ilt.py:311-337 without penalties.

    Returns (cost, dcost/draw, signed, pattern_error).
    """
    aerial = raw * scale
    fld = blur(aerial, sigma_px)
    signed = sign * (fld - level)
    a = steepness / max(abs(level), 1e-12)
    z = sigmoid_clipped(a * signed)
    resid = z - target
    npix = float(resid.size)
    cost = float(np.sum(resid ** 2) / npix)
    dz = (2.0 / npix) * resid * (a * z * (1.0 - z))
    d_raw = blur(sign * dz, sigma_px) * scale
    pe = int(np.count_nonzero((signed > 0.0) != (target > 0.5)))
    return cost, d_raw, signed, pe


def mask_penalties(mask, weight_tv=0.0, weight_discrete=0.0):
    """ilt.py:339-355: (cost, gradient) of the TV and discreteness penalties."""
    npix = float(mask.size)
    cost, grad = 0.0, np.zeros_like(mask)
    if weight_tv > 0.0:
        lap = (np.roll(mask, 1, 0) + np.roll(mask, -1, 0) + np.roll(mask, 1, 1)
               + np.roll(mask, -1, 1) - 4.0 * mask)
        gx = mask - np.roll(mask, 1, 1)
        gy = mask - np.roll(mask, 1, 0)
        cost += weight_tv * float(np.sum(gx ** 2 + gy ** 2) / npix)
        grad += weight_tv * (-2.0 / npix) * lap
    if weight_discrete > 0.0:
        cost += weight_discrete * float(np.sum(mask * (1.0 - mask)) / npix)
        grad += weight_discrete * (1.0 - 2.0 * mask) / npix
    return cost, grad


def objective(k: "SOCS", mask, target, scale, level, steepness, sign=1.0, sigma_px=0.0,
              weight_tv=0.0, weight_discrete=0.0):
    """ilt.py:292-360 (_Objective.__call__): the relaxed-resist cost of the SOCS image of
    ``mask`` with its gradient through socs_gradient, plus the TV / discreteness penalties.
    Returns (cost, grad, signed, pattern_error)."""
    fields, raw = socs_fields(k, mask)
    cost, d_raw, signed, perr = relaxed_objective(raw, target, scale, level, steepness, sign,
                                                  sigma_px)
    grad = socs_gradient(k, fields, d_raw)
    c2, g2 = mask_penalties(mask, weight_tv, weight_discrete)
    return cost + c2, grad + g2, signed, perr


def mask_of(theta, beta, smooth_px=0.0):
    """ilt.py:368-378."""
    return sigmoid_clipped(beta * blur(theta, smooth_px))


def grad_theta(grad_mask, mask, beta, smooth_px=0.0):
    """ilt.py:381-389."""
    return blur(grad_mask * (beta * mask * (1.0 - mask)), smooth_px)


# --------------------------------------------------------------------------
# Inputs (mask/patterns.py:22-66)
# --------------------------------------------------------------------------


def lines_and_spaces(n, pixel_size, pitch, cd, orientation="vertical"):
    """mask/patterns.py:22-66."""
    coords = (np.arange(n) - n // 2) * pixel_size
    line = (np.mod(coords + pitch / 2, pitch) < cd).astype(np.float64)
    if orientation == "vertical":
        return np.tile(line[None, :], (n, 1))
    if orientation == "horizontal":
        return np.tile(line[:, None], (1, n))
    raise ValueError(orientation)


# --------------------------------------------------------------------------
# Depth-resolved image through the resist film (develop/resist3d.py:66-185)
# --------------------------------------------------------------------------


def focus_reference_depth(resist):
    """resist3d.py:66-77."""
    ref = resist.focus_reference
    if ref == "top":
        return 0.0
    if ref == "mid":
        return resist.thickness / 2.0
    if ref == "bottom":
        return resist.thickness
    raise ValueError(f"focus_reference must be 'top', 'mid', or 'bottom', got '{ref}'")


def effective_defocus(depth, optics, resist):
    """resist3d.py:80-110."""
    d_ref = focus_reference_depth(resist)
    return optics.defocus - (depth - d_ref) * (optics.n_immersion / resist.n_resist)


def exposure_volume(mask, optics, grid, resist, dose=1.0):
    """resist3d.py:113-185: planes through the film, per-pixel np.interp onto the voxel
    grid, bottom-up."""
    nz = max(int(round(resist.thickness / grid.dz)), 1)
    n_slices = int(np.clip(grid.n_z_slices, 2, nz))
    depths = np.linspace(0.0, resist.thickness, n_slices)
    ratio = resist.n_resist / optics.n_immersion
    planes = compute_aerial_planes(
        mask, optics, grid,
        [(effective_defocus(float(d), optics, resist) * ratio, resist.n_resist) for d in depths],
        dose=dose)
    z_vox_depth = ((np.arange(nz) + 0.5) * grid.dz)[::-1]
    flat = planes.reshape(n_slices, -1)
    out = np.empty((nz, flat.shape[1]), dtype=np.float64)
    for k in range(flat.shape[1]):
        out[:, k] = np.interp(z_vox_depth, depths, flat[:, k])
    return out.reshape(nz, grid.n_pixels, grid.n_pixels), (np.arange(nz) + 0.5) * grid.dz


# --------------------------------------------------------------------------
# Print and solve (bake/peb.py:24-70, develop/resist.py:452-510,
# opc/ilt.py:397-648)
# --------------------------------------------------------------------------


def apply_peb(pac_map, diffusion_sigma, pixel_size):
    """peb.py:24-70: Gaussian of diffusion_sigma / pixel_size px, clipped to [0, 1]."""
    sigma_px = diffusion_sigma / pixel_size
    if sigma_px < 1e-3:
        return np.array(pac_map, dtype=np.float64, copy=True)
    from scipy.ndimage import gaussian_filter
    return np.clip(gaussian_filter(np.asarray(pac_map, dtype=np.float64), sigma=sigma_px), 0.0, 1.0)


def develop_field(aerial, resist, grid):
    """resist.py:452-503, model "threshold": (PEB-diffused aerial, threshold)."""
    return apply_peb(aerial, resist.diffusion_sigma, grid.pixel_size), float(resist.threshold)


def print_mask(mask, optics, grid, resist, dose=1.0, socs_tol=1e-8):
    """ilt.py:624-648 through SOCS kernels: the aerial image at ``dose`` (normalised as the
    optics say) and its threshold-model field.  Returns (aerial, field, level)."""
    k = socs_build(optics, grid, tol=socs_tol)
    raw, peak, clear = socs_raw(k, mask)
    aerial = raw * normalisation_scale(optics.normalisation, peak, clear, dose)
    fld, level = develop_field(aerial, resist, grid)
    return aerial, fld, level


def dose_for_target(target, optics, grid, resist, dose=1.0, tone="clear", mask=None,
                    socs_tol=1e-8):
    """ilt.py:651-721: the dose at which ``mask`` (default: the target) prints as many
    pixels as the target has, by 60 geometric bisection steps on [1e-3, 1e3] over the
    unit-dose blurred field (the image is computed once)."""
    n = grid.n_pixels
    tgt = np.asarray(target, dtype=np.float64)
    m = tgt if mask is None else np.asarray(mask, dtype=np.float64)
    if m.shape != (n, n):
        raise ValueError(f"mask must be {(n, n)} for this grid, got {m.shape}")
    want = int(np.count_nonzero(tgt > 0.5))
    if want == 0 or want == tgt.size:
        return float(dose)
    k = socs_build(optics, grid, tol=socs_tol)
    raw, _peak, clear = socs_raw(k, m)
    sigma_px = float(resist.diffusion_sigma / grid.pixel_size)
    unit = blur(raw * normalisation_scale(optics.normalisation, 1.0, clear, 1.0), sigma_px)
    level = float(resist.threshold)
    sign = 1.0 if tone == "clear" else -1.0

    def area(d):
        return int(np.count_nonzero(sign * (d * unit - level) > 0.0))

    lo, hi = 1e-3, 1e3
    rising = area(hi) > area(lo)
    for _ in range(60):
        mid = np.sqrt(lo * hi)
        if (area(mid) < want) == rising:
            lo = mid
        else:
            hi = mid
    return float(np.sqrt(lo * hi))


def adam_step(theta, mom1, mom2, grad_t, it, step, b1=0.9, b2=0.999, eps=1e-8):
    """ilt.py:587-591: one Adam update; returns (theta, mom1, mom2)."""
    mom1 = b1 * mom1 + (1.0 - b1) * grad_t
    mom2 = b2 * mom2 + (1.0 - b2) * grad_t ** 2
    m_hat = mom1 / (1.0 - b1 ** it)
    v_hat = mom2 / (1.0 - b2 ** it)
    return theta - step * m_hat / (np.sqrt(v_hat) + eps), mom1, mom2


def run_ilt(target, optics, grid, resist, dose=1.0, tone="clear", *, max_iter=120, step=0.4,
            tol=1e-7, beta=4.0, beta_final=12.0, initial_grey=0.9, steepness=12.0,
            weight_tv=0.0, weight_discrete=0.0, smooth=0.0, socs_tol=1e-8):
    """ilt.py:397-621 through SOCS kernels (tol socs_tol): Adam on theta, mask =
    sigmoid(beta G(theta)), beta annealed geometrically, stop when the cost moves by
    less than tol.  Returns (mask, theta, history rows, iterations, converged)."""
    k = socs_build(optics, grid, tol=socs_tol)
    tgt = np.asarray(target, dtype=np.float64)
    smooth_px = float(smooth / grid.pixel_size)
    scale = normalisation_scale(optics.normalisation, 1.0, k.clear, dose)
    level = float(resist.threshold)
    sigma_px = float(resist.diffusion_sigma / grid.pixel_size)
    sign = 1.0 if tone == "clear" else -1.0

    def obj(mask):
        return objective(k, mask, tgt, scale, level, steepness, sign, sigma_px, weight_tv,
                         weight_discrete)

    p = float(np.clip(initial_grey, 0.5 + 1e-3, 1.0 - 1e-6))
    m0 = 0.5 + (np.clip(tgt, 0.0, 1.0) - 0.5) * (2.0 * p - 1.0)
    theta = np.log(m0 / (1.0 - m0)) / beta
    growth = (1.0 if beta_final is None or max_iter <= 1
              else (beta_final / beta) ** (1.0 / (max_iter - 1)))
    mom1 = np.zeros_like(theta)
    mom2 = np.zeros_like(theta)
    b1, b2, eps = 0.9, 0.999, 1e-8
    beta_k = float(beta)
    mask = mask_of(theta, beta_k, smooth_px)
    cost, grad_m, signed, perr = obj(mask)
    rows = [{"iteration": 0, "cost": cost, "pattern_error": perr, "beta": beta_k,
             "grad_norm": float(np.linalg.norm(grad_m))}]
    prev_cost, converged, taken = np.inf, False, 0
    for it in range(1, max_iter + 1):
        grad_t = grad_theta(grad_m, mask, beta_k, smooth_px)
        theta, mom1, mom2 = adam_step(theta, mom1, mom2, grad_t, it, step, b1, b2, eps)
        beta_k *= growth
        mask = mask_of(theta, beta_k, smooth_px)
        cost, grad_m, signed, perr = obj(mask)
        taken = it
        rows.append({"iteration": it, "cost": cost, "pattern_error": perr, "beta": beta_k,
                     "grad_norm": float(np.linalg.norm(grad_m))})
        if abs(prev_cost - cost) < tol:
            converged = True
            break
        prev_cost = cost
    return mask, theta, rows, taken, converged
