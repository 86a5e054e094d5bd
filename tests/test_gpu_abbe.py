"""Device Abbe sum vs the float64 oracle and the reference's golden images.

FP64 path: max |dI| <= 1e-12 * peak.  FP32 path: max |dI| / max I <= 1e-5.
"""
import json
import os

import numpy as np
import pytest

from oracle import litho_oracle as O

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")


def _cfg(d):
    from lithobat_amd.config import OpticsConfig
    d = dict(d)
    d["zernike_coeffs"] = {int(k): v for k, v in (d.get("zernike_coeffs") or {}).items()}
    return OpticsConfig(**d)


def _rel(a, b):
    return float(np.abs(np.asarray(a) - np.asarray(b)).max() / np.abs(np.asarray(b)).max())


CASES = {
    # name: (optics kwargs, n, pixel)  -- coarse-grid and direct paths both covered
    "conv_64_4nm": (dict(NA=0.93, sigma_outer=0.8, source_grid=9), 64, 4e-9),
    "annular_defocus_zern_64_8nm": (dict(NA=0.93, source_type="annular", sigma_outer=0.85,
                                         sigma_inner=0.5, source_grid=11, defocus=120e-9,
                                         zernike_coeffs={7: 0.03, 11: -0.02}), 64, 8e-9),
    "direct_tilt_64_32nm": (dict(NA=0.93, sigma_outer=0.5, source_grid=5,
                                 zernike_coeffs={4: 0.05}, exact_defocus=True), 64, 32e-9),
    "dipole_exact_128_4nm": (dict(NA=1.35, n_immersion=1.44, n_image=1.7, source_type="dipole",
                                  source_kwargs={"axis": "x"}, sigma_outer=0.9, sigma_inner=0.6,
                                  source_grid=11, defocus=-60e-9, exact_defocus=True), 128, 4e-9),
    "quad_256_5nm": (dict(NA=0.85, source_type="quadrupole", sigma_outer=0.85,
                          sigma_inner=0.55, source_grid=15, defocus=40e-9), 256, 5e-9),
}


@pytest.mark.parametrize("name", list(CASES))
@pytest.mark.parametrize("precision", [64, 32])
def test_scalar_raw_matches_oracle(device, name, precision):
    from lithobat_amd import compute_aerial_image
    from lithobat_amd.config import GridConfig, OpticsConfig
    kw, n, px = CASES[name]
    optics = OpticsConfig(normalisation="none", **kw)
    grid = GridConfig(n_pixels=n, pixel_size=px)
    rng = np.random.default_rng(n)
    mask = (rng.random((n, n)) > 0.5).astype(np.float64)
    ref, rpk, rcl = O.compute_aerial_image(mask, optics, grid, return_raw=True)
    raw, pk, cl = compute_aerial_image(mask, optics, grid, return_raw=True, precision=precision)
    tol = 1e-12 if precision == 64 else 1e-5
    assert _rel(raw, ref) <= tol, (name, _rel(raw, ref))
    assert pk == pytest.approx(rpk, rel=tol * 10)
    assert cl == pytest.approx(rcl, rel=1e-12)


def test_golden_scalar_planes(device):
    from lithobat_amd import compute_aerial_planes
    from lithobat_amd.config import GridConfig
    z = np.load(os.path.join(GOLD, "scalar128.npz"))
    grid = GridConfig(n_pixels=128, pixel_size=4e-9)
    planes = [tuple(p) for p in z["planes"]]
    for c in z["cases"]:
        optics = _cfg(json.loads(str(z[f"{c}__optics"])))
        for prec, tol in ((64, 1e-12), (32, 1e-5)):
            st = compute_aerial_planes(z["mask"], optics, grid, planes, precision=prec)
            assert _rel(st, z[f"{c}__raw"]) <= tol, (c, prec, _rel(st, z[f"{c}__raw"]))


def test_golden_phase_shift_mask(device):
    from lithobat_amd import compute_aerial_image
    from lithobat_amd.config import GridConfig
    z = np.load(os.path.join(GOLD, "scalar128.npz"))
    optics = _cfg(json.loads(str(z["psm__optics"])))
    grid = GridConfig(n_pixels=128, pixel_size=4e-9)
    raw, pk, cl = compute_aerial_image(z["psm_mask"], optics, grid, return_raw=True)
    assert _rel(raw, z["psm__raw"]) <= 1e-12
    assert cl == pytest.approx(float(z["psm__clear"]), rel=1e-12)


def test_golden_vector_every_polarisation(device):
    from lithobat_amd import compute_aerial_image
    from lithobat_amd.config import GridConfig
    z = np.load(os.path.join(GOLD, "vector64.npz"))
    grid = GridConfig(n_pixels=64, pixel_size=4e-9)
    for c in z["cases"]:
        optics = _cfg(json.loads(str(z[f"{c}__optics"])))
        for prec, tol in ((64, 1e-12), (32, 1e-5)):
            raw, pk, cl = compute_aerial_image(z["mask"], optics, grid, return_raw=True,
                                               precision=prec)
            assert _rel(raw, z[f"{c}__raw"]) <= tol, (c, prec, _rel(raw, z[f"{c}__raw"]))
            assert cl == pytest.approx(float(z[f"{c}__clear"]), rel=1e-12)


def test_batched_masks_and_planes_equal_single_calls(device):
    import torch
    from lithobat_amd import compute_aerial_batch, compute_aerial_image
    from lithobat_amd.config import GridConfig, OpticsConfig
    from lithobat_amd.masks import manhattan
    grid = GridConfig(n_pixels=128, pixel_size=4e-9)
    optics = OpticsConfig(NA=0.93, source_type="quadrupole", sigma_outer=0.85, sigma_inner=0.55,
                          source_grid=21, normalisation="none")
    masks = np.stack([manhattan(128, 4e-9, seed=s, n_rect=(5, 15)) for s in range(3)])
    planes = [(-50e-9, 1.0), (0.0, 1.0), (70e-9, 1.0)]
    mt = torch.tensor(masks, device=device, dtype=torch.float32)
    batch = compute_aerial_batch(mt, optics, grid, planes).cpu().numpy()
    for b in range(3):
        for p, (dz, ni) in enumerate(planes):
            o = OpticsConfig(**{**optics.__dict__, "defocus": dz, "n_image": ni})
            ref = O.compute_aerial_image(masks[b], o, grid, return_raw=True)[0]
            assert _rel(batch[b, p], ref) <= 1e-5
            one = compute_aerial_image(mt[b], o, grid, return_raw=True)[0].cpu().numpy()
            assert _rel(batch[b, p], one) <= 1e-6


def test_boxes_tightened_to_pupil_support(device):
    """NA 0.93 at 1024 x 4 nm: pupil radius 19.74 bins covers at most 40 bins per axis, so the
    boxes shrink from floor(2r) + 3 = 42 to 40 (coarse grid 96; 80 on request)."""
    from lithobat_amd.config import GridConfig, OpticsConfig
    from lithobat_amd.plan import ImagingPlan
    optics = OpticsConfig(NA=0.93, source_type="quadrupole", sigma_outer=0.85, sigma_inner=0.55,
                          source_grid=47, normalisation="none")
    plan = ImagingPlan(optics, GridConfig(n_pixels=1024, pixel_size=4e-9), precision=32)
    assert plan.info["box"] == 40 and plan.info["m"] == 96
    small = ImagingPlan(optics, GridConfig(n_pixels=1024, pixel_size=4e-9), precision=32,
                        coarse_m=80)
    assert small.info["box"] == 40 and small.info["m"] == 80
    forced = ImagingPlan(optics, GridConfig(n_pixels=1024, pixel_size=4e-9), precision=32,
                         coarse_m=128)
    assert forced.info["box"] == 40 and forced.info["m"] == 128


@pytest.mark.parametrize("precision", [32, 64])
def test_1024_quadrupole_defocus_matches_oracle(device, precision):
    """The batched-benchmark geometry at one mask and one plane."""
    from lithobat_amd import compute_aerial_image
    from lithobat_amd.config import GridConfig, OpticsConfig
    from lithobat_amd.masks import manhattan
    grid = GridConfig(n_pixels=1024, pixel_size=4e-9)
    optics = OpticsConfig(NA=0.93, source_type="quadrupole", sigma_outer=0.85, sigma_inner=0.55,
                          source_grid=47, defocus=-100e-9, normalisation="none")
    mask = manhattan(1024, 4e-9, seed=1000)
    ref = O.compute_aerial_image(mask, optics, grid, return_raw=True)[0]
    raw = compute_aerial_image(mask, optics, grid, return_raw=True, precision=precision)[0]
    tol = 1e-12 if precision == 64 else 1e-5
    assert _rel(raw, ref) <= tol, _rel(raw, ref)
    from lithobat_amd.plan import ImagingPlan
    for m in (80, 128):  # the same image on the radix-5 and on a larger coarse grid
        plan = ImagingPlan(optics, grid, precision=precision, coarse_m=m)
        import torch
        t = torch.tensor(mask, device=device,
                         dtype=torch.float64 if precision == 64 else torch.float32)
        img = plan.forward(t)[0][0, 0].double().cpu().numpy()
        assert _rel(img, ref) <= tol, (m, _rel(img, ref))


@pytest.mark.parametrize("precision", [64, 32])
def test_dense_reference_dataflow_equals_coarse_grid(device, precision):
    """One dense N x N transform per source point gives the coarse-grid image."""
    import torch
    from lithobat_amd.config import GridConfig, OpticsConfig
    from lithobat_amd.masks import manhattan
    from lithobat_amd.plan import ImagingPlan
    grid = GridConfig(n_pixels=256, pixel_size=4e-9)
    optics = OpticsConfig(NA=0.93, source_type="quadrupole", sigma_outer=0.85, sigma_inner=0.55,
                          source_grid=15, defocus=50e-9, zernike_coeffs={8: 0.02},
                          normalisation="none")
    plan = ImagingPlan(optics, grid, precision=precision)
    dt = torch.float64 if precision == 64 else torch.float32
    m = torch.tensor(manhattan(256, 4e-9, seed=5, n_rect=(5, 12)), device=device, dtype=dt)
    coarse = plan.forward(m)[0][0, 0].double().cpu().numpy()
    dense = plan.forward_dense(m, batch=5).double().cpu().numpy()
    ref = O.compute_aerial_image(m.double().cpu().numpy(), optics, grid, return_raw=True)[0]
    tol = 1e-12 if precision == 64 else 1e-5
    assert _rel(dense, ref) <= tol
    assert _rel(coarse, ref) <= tol


def test_plain_fp64_lines_and_spaces_488_points(device):
    """The CPU-runnable configuration: vertical lines/spaces (pitch 200 nm, CD 100 nm) at 256^2,
    5 nm, NA 0.85, annular 0.5-0.8 on a 41-point grid (488 points), peak normalisation, FP64."""
    from lithobat_amd import compute_aerial_image
    from lithobat_amd.config import GridConfig, OpticsConfig
    grid = GridConfig(n_pixels=256, pixel_size=5e-9)
    optics = OpticsConfig(wavelength=193e-9, NA=0.85, source_type="annular", sigma_outer=0.8,
                          sigma_inner=0.5, source_grid=41)
    mask = O.lines_and_spaces(256, 5e-9, 200e-9, 100e-9)
    assert len(O.source_points(optics)) == 488
    ref = O.compute_aerial_image(mask, optics, grid)
    img = compute_aerial_image(mask, optics, grid)
    assert img.dtype == np.float64 and img.shape == (256, 256)
    assert _rel(img, ref) <= 1e-12, _rel(img, ref)



_SIDE_REF = {}


@pytest.mark.parametrize("n", [48, 80, 96, 192, 384, 768, 1536, 2048])
@pytest.mark.parametrize("precision", [32, 64])
def test_every_grid_family_matches_oracle(device, n, precision):
    """Grid sides 3 * 2^k and 80 (radix-3 and radix-5 band plans, the 1536-point one in four
    stages) and the largest power of two, 4 nm pixels, NA 0.85 annular on a 7-point source
    grid: the band transforms' block shapes change with the side and the precision."""
    from lithobat_amd import compute_aerial_image
    from lithobat_amd.config import GridConfig, OpticsConfig
    from lithobat_amd.masks import manhattan
    grid = GridConfig(n_pixels=n, pixel_size=4e-9)
    optics = OpticsConfig(NA=0.85, source_type="annular", sigma_outer=0.8, sigma_inner=0.5,
                          source_grid=7, defocus=30e-9, normalisation="none")
    if n not in _SIDE_REF:
        mask = manhattan(n, 4e-9, seed=n, n_rect=(max(2, n // 64), max(4, n // 32)))
        _SIDE_REF[n] = (mask, O.compute_aerial_image(mask, optics, grid, return_raw=True)[0])
    mask, ref = _SIDE_REF[n]
    raw = compute_aerial_image(mask, optics, grid, return_raw=True, precision=precision)[0]
    tol = 1e-12 if precision == 64 else 1e-5
    assert _rel(raw, ref) <= tol, (n, precision, _rel(raw, ref))
