"""Write the golden vectors under tests/golden/ from the reference implementation.

Run once, in a container that has the reference checkout at /root/reference:

    PYTHONDONTWRITEBYTECODE=1 python tests/golden/make_golden.py

Only results that no closed form, invariant or brute-force check pins are
written here (vector images for every polarisation state, defocused and
aberrated scalar images on several planes, SOCS eigenvalues / truncated
images), plus the source tables that feed them.  Each file records its inputs,
the optics as a dict, and the numpy / scipy versions.  The outputs are data
only; nothing of the reference's code is stored.
"""
from __future__ import annotations

import json
import os
import sys

import numpy as np
import scipy

REF_SRC = "/root/reference/src"
HERE = os.path.dirname(os.path.abspath(__file__))


def manhattan(n, px, seed, n_rect=(6, 14), wmin=40e-9, wmax=200e-9):
    """Seeded union of axis-aligned rectangles (0/1), snapped to pixels."""
    rng = np.random.default_rng(seed)
    m = np.zeros((n, n))
    for _ in range(int(rng.integers(n_rect[0], n_rect[1] + 1))):
        w = max(1, int(round(rng.uniform(wmin, wmax) / px)))
        h = max(1, int(round(rng.uniform(wmin, wmax) / px)))
        y0 = int(rng.integers(0, n))
        x0 = int(rng.integers(0, n))
        ys = np.arange(y0, y0 + h) % n
        xs = np.arange(x0, x0 + w) % n
        m[np.ix_(ys, xs)] = 1.0
    return m


def main():
    sys.path.insert(0, REF_SRC)
    from litho_sim.core.config import GridConfig, OpticsConfig
    from litho_sim.expose import SOCSKernels, compute_aerial_image, compute_aerial_planes
    from litho_sim.expose.aerial_image import source_points

    meta = {"numpy": np.__version__, "scipy": scipy.__version__}

    def optics_dict(o):
        keys = ["wavelength", "NA", "n_immersion", "defocus", "sigma_outer", "sigma_inner",
                "source_type", "source_grid", "source_kwargs", "zernike_coeffs",
                "exact_defocus", "n_image", "imaging_model", "polarisation", "obliquity",
                "normalisation"]
        d = {k: getattr(o, k) for k in keys}
        d["zernike_coeffs"] = {str(k): v for k, v in d["zernike_coeffs"].items()}
        return d

    # --- source tables for the benchmark sources -----------------------------
    srcs = {
        "annular_0.5_0.8_g41_na0.85": OpticsConfig(NA=0.85, source_type="annular",
                                                   sigma_outer=0.8, sigma_inner=0.5,
                                                   source_grid=41),
        "quad_0.85_0.55_g47_na0.93": OpticsConfig(NA=0.93, source_type="quadrupole",
                                                  sigma_outer=0.85, sigma_inner=0.55,
                                                  source_grid=47),
        "annular_0.9_0.6_g21_na1.35": OpticsConfig(NA=1.35, n_immersion=1.44,
                                                   source_type="annular", sigma_outer=0.9,
                                                   sigma_inner=0.6, source_grid=21),
        "dipole_y_0.9_0.6_g11": OpticsConfig(NA=0.93, source_type="dipole",
                                             source_kwargs={"axis": "y"}, sigma_outer=0.9,
                                             sigma_inner=0.6, source_grid=11),
        "conventional_0.8_g21": OpticsConfig(),
    }
    arrs = {}
    for name, o in srcs.items():
        s = source_points(o)
        for f in ("weight", "fx", "fy", "phi"):
            arrs[f"{name}__{f}"] = s.__dict__[f]
        arrs[f"{name}__optics"] = np.array(json.dumps(optics_dict(o)))
    np.savez_compressed(os.path.join(HERE, "sources.npz"), meta=json.dumps(meta), **arrs)

    # --- vector images, every polarisation, aberrated + defocused -----------
    grid = GridConfig(n_pixels=64, pixel_size=4e-9)
    mask = manhattan(64, 4e-9, seed=7)
    zern = {4: 0.012, 5: -0.015, 6: 0.008, 7: 0.02, 8: -0.01, 9: 0.017, 10: -0.006, 11: 0.011}
    arrs = {"mask": mask, "n": 64, "pixel_size": 4e-9}
    cases = []
    for pol in ("te", "tm", "x", "y", "unpolarised"):
        for obl in (True, False):
            if not obl and pol != "tm":
                continue
            o = OpticsConfig(NA=1.35, n_immersion=1.44, imaging_model="vector",
                             polarisation=pol, obliquity=obl, zernike_coeffs=zern,
                             defocus=-40e-9, source_type="quadrupole", sigma_outer=0.85,
                             sigma_inner=0.55, source_grid=11, normalisation="clear")
            raw, peak, clear = compute_aerial_image(mask, o, grid, return_raw=True)
            key = f"{pol}_obl{int(obl)}"
            arrs[f"{key}__raw"] = raw
            arrs[f"{key}__peak"] = peak
            arrs[f"{key}__clear"] = clear
            arrs[f"{key}__optics"] = np.array(json.dumps(optics_dict(o)))
            cases.append(key)
    np.savez_compressed(os.path.join(HERE, "vector64.npz"), meta=json.dumps(meta),
                        cases=np.array(cases), **arrs)

    # --- scalar, quadrupole, several planes, paraxial and exact -------------
    grid = GridConfig(n_pixels=128, pixel_size=4e-9)
    mask = manhattan(128, 4e-9, seed=11)
    planes = [(-100e-9, 1.0), (0.0, 1.0), (100e-9, 1.0)]
    arrs = {"mask": mask, "n": 128, "pixel_size": 4e-9, "planes": np.array(planes)}
    cases = []
    for name, extra in (("paraxial", {}),
                        ("exact_zernike", {"exact_defocus": True,
                                           "zernike_coeffs": {7: 0.03, 11: -0.02, 5: 0.01}})):
        o = OpticsConfig(NA=0.93, source_type="quadrupole", sigma_outer=0.85,
                         sigma_inner=0.55, source_grid=47, normalisation="none", **extra)
        stack = compute_aerial_planes(mask, o, grid, planes)
        arrs[f"{name}__raw"] = stack
        arrs[f"{name}__optics"] = np.array(json.dumps(optics_dict(o)))
        cases.append(name)
    # a complex (phase-shift) mask, one plane, clear normalised
    pmask = np.where(manhattan(128, 4e-9, seed=12) > 0, -0.245 + 0.0j, 1.0 + 0.0j) * \
        np.exp(0.3j * manhattan(128, 4e-9, seed=13))
    o = OpticsConfig(NA=0.93, source_type="annular", sigma_outer=0.8, sigma_inner=0.5,
                     source_grid=15, defocus=35e-9, normalisation="clear")
    raw, peak, clear = compute_aerial_image(pmask, o, grid, return_raw=True)
    arrs["psm_mask"] = pmask
    arrs["psm__raw"] = raw
    arrs["psm__peak"] = peak
    arrs["psm__clear"] = clear
    arrs["psm__optics"] = np.array(json.dumps(optics_dict(o)))
    np.savez_compressed(os.path.join(HERE, "scalar128.npz"), meta=json.dumps(meta),
                        cases=np.array(cases), **arrs)

    # --- SOCS: eigenvalues, clear, truncated and full images ----------------
    grid = GridConfig(n_pixels=64, pixel_size=8e-9)
    mask = manhattan(64, 8e-9, seed=21)
    o = OpticsConfig(NA=1.35, n_immersion=1.44, source_type="annular", sigma_outer=0.9,
                     sigma_inner=0.6, source_grid=21)
    full = SOCSKernels.build(o, grid, tol=0.0)
    k24 = SOCSKernels.build(o, grid, tol=0.0, max_kernels=24)
    dflt = SOCSKernels.build(o, grid)
    arrs = {"mask": mask, "n": 64, "pixel_size": 8e-9,
            "optics": np.array(json.dumps(optics_dict(o)))}
    for name, k in (("full", full), ("k24", k24), ("default", dflt)):
        raw, peak, clear = k.raw(mask)
        arrs[f"{name}__weights"] = k.weights
        arrs[f"{name}__clear"] = k.clear
        arrs[f"{name}__energy_kept"] = k.energy_kept
        arrs[f"{name}__n_band"] = k.n_band
        arrs[f"{name}__n_source"] = k.n_source
        arrs[f"{name}__raw"] = raw
        arrs[f"{name}__peak"] = peak
    np.savez_compressed(os.path.join(HERE, "socs64.npz"), meta=json.dumps(meta), **arrs)
    print("golden vectors written to", HERE)


def main_multilayer():
    """The reflective-mask (EUV multilayer) images: written to multilayer64.npz by
    ``make_golden.py multilayer``."""
    sys.path.insert(0, REF_SRC)
    from litho_sim.core.config import GridConfig, OpticsConfig
    from litho_sim.expose import compute_aerial_image, compute_aerial_planes
    from litho_sim.expose.m3d.provider import MultilayerSpectra
    from litho_sim.expose.m3d.stack import MaskStack

    meta = {"numpy": np.__version__, "scipy": scipy.__version__}
    keys = ["wavelength", "NA", "n_immersion", "defocus", "sigma_outer", "sigma_inner",
            "source_type", "source_grid", "source_kwargs", "zernike_coeffs", "exact_defocus",
            "n_image", "imaging_model", "polarisation", "obliquity", "normalisation",
            "mask_model", "reduction", "chief_ray_deg", "chief_ray_axis", "mask_stack"]

    def optics_dict(o):
        d = {k: getattr(o, k) for k in keys}
        d["zernike_coeffs"] = {str(k): v for k, v in d["zernike_coeffs"].items()}
        return d

    grid = GridConfig(n_pixels=64, pixel_size=4e-9)
    mask = manhattan(64, 4e-9, seed=31, wmin=16e-9, wmax=60e-9)
    arrs = {"mask": mask, "n": 64, "pixel_size": 4e-9}
    base = dict(wavelength=13.5e-9, NA=0.33, n_immersion=1.0, source_type="annular",
                sigma_outer=0.8, sigma_inner=0.4, source_grid=11, reduction=4.0,
                chief_ray_deg=6.0, mask_model="multilayer", normalisation="clear")
    thin = OpticsConfig(**{**base, "mask_model": "thin"})
    cases = {
        "scalar": OpticsConfig(**base, defocus=20e-9),
        "scalar_y_30bl": OpticsConfig(**{**base, "chief_ray_axis": "y"},
                                      mask_stack={"regime": "euv_reflective",
                                                  "multilayer": {"n_bilayers": 30,
                                                                 "period": 7.0e-9}}),
        "vector_te": OpticsConfig(**base, imaging_model="vector", polarisation="te",
                                  zernike_coeffs={5: 0.01, 7: -0.015}),
    }
    names = []
    for name, o in cases.items():
        raw, peak, clear = compute_aerial_image(mask, o, grid, return_raw=True)
        arrs[f"{name}__raw"] = raw
        arrs[f"{name}__peak"] = peak
        arrs[f"{name}__clear"] = clear
        arrs[f"{name}__optics"] = np.array(json.dumps(optics_dict(o)))
        names.append(name)
        st = MaskStack.from_spec(o.mask_stack) if o.mask_stack else \
            MaskStack.for_wavelength(o.wavelength)
        ms = MultilayerSpectra(mask, o, grid, st)
        arrs[f"{name}__theta"] = ms._theta
        arrs[f"{name}__mag"] = ms._mag
        arrs[f"{name}__phase"] = ms._phase
    planes = [(-30e-9, 1.0), (0.0, 1.0), (30e-9, 1.0)]
    arrs["planes"] = np.array(planes)
    arrs["planes__raw"] = compute_aerial_planes(mask, cases["scalar"], grid, planes)
    raw, peak, clear = compute_aerial_image(mask, thin, grid, return_raw=True)
    arrs["thin__raw"] = raw
    arrs["thin__clear"] = clear
    # the blank's reflectance at 6 degrees, 40 bilayers (the materials table's check)
    st = MaskStack.for_wavelength(13.5e-9)
    arrs["blank_r_s_6deg"] = np.array(st.blank_reflection(13.5e-9, np.deg2rad(6.0), "s"))
    arrs["blank_r_p_6deg"] = np.array(st.blank_reflection(13.5e-9, np.deg2rad(6.0), "p"))
    np.savez_compressed(os.path.join(HERE, "multilayer64.npz"), meta=json.dumps(meta),
                        cases=np.array(names), **arrs)
    print("multilayer golden vectors written to", HERE)


if __name__ == "__main__":
    if sys.argv[1:] == ["multilayer"]:
        main_multilayer()
    else:
        main()
