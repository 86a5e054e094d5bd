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


def main_ilt():
    """The ILT objective family, written to ilt32.npz by ``make_golden.py ilt``:
    ``_Objective.__call__`` (opc/ilt.py:292-360) for both signs, with and without the
    PEB blur and the penalties; ``_mask_of`` / ``_grad_theta`` (:368-389) with and
    without smoothing; two short ``run_ilt`` solves (:397-621, one per tone);
    ``dose_for_target`` (:651-721) for both tones and an explicit mask; and
    ``exposure_volume`` (develop/resist3d.py:113-185) for every focus reference."""
    sys.path.insert(0, REF_SRC)
    from litho_sim.core.config import GridConfig, OpticsConfig, ResistConfig
    from litho_sim.develop.resist3d import exposure_volume
    from litho_sim.expose import SOCSKernels
    from litho_sim.expose.aerial_image import normalisation_scale
    from litho_sim.opc import ilt as I
    from litho_sim.opc.model import PrintModel

    meta = {"numpy": np.__version__, "scipy": scipy.__version__}
    keys = ["wavelength", "NA", "n_immersion", "defocus", "sigma_outer", "sigma_inner",
            "source_type", "source_grid", "source_kwargs", "zernike_coeffs", "exact_defocus",
            "n_image", "imaging_model", "polarisation", "obliquity", "normalisation"]

    def optics_dict(o):
        d = {k: getattr(o, k) for k in keys}
        d["zernike_coeffs"] = {str(k): v for k, v in d["zernike_coeffs"].items()}
        return d

    n, px = 32, 8e-9
    grid = GridConfig(n_pixels=n, pixel_size=px)
    optics = OpticsConfig(NA=1.35, n_immersion=1.44, source_type="quadrupole",
                          sigma_outer=0.85, sigma_inner=0.55, source_grid=21,
                          normalisation="clear")
    arrs = {"n": n, "pixel_size": px, "optics": np.array(json.dumps(optics_dict(optics)))}

    # --- the objective on a sigmoid(8 theta) mask ------------------------------
    k = SOCSKernels.build(optics, grid, tol=1e-8)
    rng = np.random.default_rng(3000)
    theta = rng.standard_normal((n, n))
    mask = I._mask_of(theta, 8.0, 0.0)
    target = manhattan(n, px, seed=3001, n_rect=(3, 6))
    dose = 1.0
    scale = normalisation_scale("clear", 1.0, k.clear, dose)
    arrs.update(obj_theta=theta, obj_mask=mask, obj_target=target, obj_scale=scale,
                obj_socs_tol=1e-8, obj_clear=k.clear, obj_weights=k.weights)
    cases = {
        "plain": dict(sign=1.0, sigma_px=0.0, steepness=50.0, level=0.3, wtv=0.0, wd=0.0),
        "dark_blur": dict(sign=-1.0, sigma_px=1.5, steepness=50.0, level=0.3, wtv=0.0,
                          wd=0.0),
        "blur_penalties": dict(sign=1.0, sigma_px=1.5, steepness=12.0, level=0.25,
                               wtv=0.05, wd=0.02),
    }
    for name, c in cases.items():
        obj = I._Objective(kernels=k, target=target, scale=scale, level=c["level"],
                           sigma_px=c["sigma_px"], sign=c["sign"], steepness=c["steepness"],
                           weight_tv=c["wtv"], weight_discrete=c["wd"])
        cost, grad, signed, perr = obj(mask)
        arrs[f"obj_{name}__params"] = np.array(json.dumps(c))
        arrs[f"obj_{name}__cost"] = cost
        arrs[f"obj_{name}__grad"] = grad
        arrs[f"obj_{name}__signed"] = signed
        arrs[f"obj_{name}__pattern_error"] = perr
    # socs_gradient on the fields socs_fields returns, for a plain sensitivity
    fields, inten = I.socs_fields(k, mask)
    sens = rng.standard_normal((n, n))
    arrs["sg_sensitivity"] = sens
    arrs["sg_intensity"] = inten
    arrs["sg_grad"] = I.socs_gradient(k, fields, sens)

    # --- the parameterisation ---------------------------------------------------
    gm = rng.standard_normal((n, n))
    arrs["par_grad_mask"] = gm
    for name, beta, smooth_px in (("b8_s0", 8.0, 0.0), ("b4_s1.3", 4.0, 1.3),
                                  ("b12_s2.5", 12.0, 2.5)):
        m = I._mask_of(theta, beta, smooth_px)
        arrs[f"par_{name}__beta"] = beta
        arrs[f"par_{name}__smooth_px"] = smooth_px
        arrs[f"par_{name}__mask"] = m
        arrs[f"par_{name}__grad_theta"] = I._grad_theta(gm, m, beta, smooth_px)

    # --- dose_for_target and two short solves -------------------------------------
    resist = ResistConfig(threshold=0.3, diffusion_sigma=8e-9)
    ilt_target = manhattan(n, px, seed=3003, n_rect=(2, 4), wmin=48e-9, wmax=120e-9)
    other = manhattan(n, px, seed=3004, n_rect=(2, 4), wmin=48e-9, wmax=120e-9)
    arrs["ilt_target"] = ilt_target
    arrs["dose_other_mask"] = other
    arrs["resist_threshold"] = resist.threshold
    arrs["resist_diffusion_sigma"] = resist.diffusion_sigma
    for tone in ("clear", "dark"):
        model = PrintModel(optics=optics, resist=resist, grid=grid, tone=tone)
        arrs[f"dose_{tone}"] = I.dose_for_target(ilt_target, model)
        arrs[f"dose_{tone}__other"] = I.dose_for_target(ilt_target, model, mask=other)
    solves = {
        "clear": dict(tone="clear", max_iter=5, smooth=8e-9, weight_tv=1e-3,
                      weight_discrete=1e-3, beta=4.0, beta_final=12.0, steepness=12.0),
        "dark": dict(tone="dark", max_iter=5, smooth=0.0, weight_tv=0.0,
                     weight_discrete=0.0, beta=4.0, beta_final=None, steepness=4.0),
    }
    for name, s in solves.items():
        # the dark-tone solve prints the complement (its features are the dark side)
        tgt = ilt_target if s["tone"] == "clear" else 1.0 - ilt_target
        model = PrintModel(optics=optics, resist=resist, grid=grid, tone=s["tone"])
        model.dose = I.dose_for_target(tgt, model)
        kw = {key: v for key, v in s.items() if key != "tone"}
        res = I.run_ilt(tgt, model, **kw)
        h = res.history
        arrs[f"run_{name}__target"] = tgt
        arrs[f"run_{name}__params"] = np.array(json.dumps(s))
        arrs[f"run_{name}__dose"] = model.dose
        for col in ("cost", "pattern_error", "beta", "grad_norm"):
            arrs[f"run_{name}__{col}"] = h[col].to_numpy()
        arrs[f"run_{name}__mask"] = res.mask
        arrs[f"run_{name}__theta"] = res.theta
        arrs[f"run_{name}__iterations"] = res.iterations
        arrs[f"run_{name}__converged"] = res.converged
        arrs[f"run_{name}__printed_aerial"] = res.printed.aerial
        arrs[f"run_{name}__printed_field"] = res.printed.field

    # --- exposure_volume for every focus reference ----------------------------------
    vgrid = GridConfig(n_pixels=16, pixel_size=px, dz=4e-9, n_z_slices=6)
    voptics = {
        "none": OpticsConfig(NA=0.93, source_type="annular", sigma_outer=0.8,
                             sigma_inner=0.5, source_grid=7, defocus=20e-9,
                             normalisation="none"),
        "clear": OpticsConfig(NA=1.35, n_immersion=1.44, source_type="quadrupole",
                              sigma_outer=0.85, sigma_inner=0.55, source_grid=11,
                              defocus=-30e-9, normalisation="clear"),
    }
    vmask = manhattan(16, px, seed=3005, n_rect=(2, 3), wmin=16e-9, wmax=48e-9)
    arrs["vol_mask"] = vmask
    arrs["vol_dz"] = vgrid.dz
    arrs["vol_n_z_slices"] = vgrid.n_z_slices
    vcases = []
    for oname, o in voptics.items():
        arrs[f"vol_optics_{oname}"] = np.array(json.dumps(optics_dict(o)))
        for ref in ("top", "mid", "bottom"):
            r = ResistConfig(thickness=90e-9, n_resist=1.7, focus_reference=ref)
            vol, z = exposure_volume(vmask, o, vgrid, r, dose=1.3)
            key = f"vol_{oname}_{ref}"
            arrs[f"{key}__intensity"] = vol
            arrs[f"{key}__z"] = z
            vcases.append(key)
    arrs["vol_thickness"] = 90e-9
    arrs["vol_n_resist"] = 1.7
    arrs["vol_dose"] = 1.3
    np.savez_compressed(os.path.join(HERE, "ilt32.npz"), meta=json.dumps(meta),
                        obj_cases=np.array(list(cases)), vol_cases=np.array(vcases), **arrs)
    print("ILT golden vectors written to", HERE)


if __name__ == "__main__":
    if sys.argv[1:] == ["multilayer"]:
        main_multilayer()
    elif sys.argv[1:] == ["ilt"]:
        main_ilt()
    else:
        main()
