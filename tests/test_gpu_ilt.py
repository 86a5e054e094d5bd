"""Device adjoint (recomputing backward) vs the oracle's adjoint and finite differences."""
import numpy as np
import pytest

from oracle import litho_oracle as O

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.abs(np.asarray(a) - np.asarray(b)).max() / np.abs(np.asarray(b)).max())


GRAD_CASES = {
    "scalar_quad_coarse": (dict(NA=1.35, n_immersion=1.44, source_type="quadrupole",
                                sigma_outer=0.85, sigma_inner=0.55, source_grid=11,
                                defocus=30e-9, zernike_coeffs={8: 0.02}), 64, 4e-9),
    "scalar_direct": (dict(NA=0.93, sigma_outer=0.8, source_grid=7), 32, 8e-9),
    # pupil radius 19.7 bins: 40-bin boxes (coarse grid 96; 80 in the forced case below)
    "scalar_quad_w40": (dict(NA=0.93, source_type="quadrupole", sigma_outer=0.85,
                             sigma_inner=0.55, source_grid=15, defocus=-40e-9,
                             zernike_coeffs={7: 0.02}), 512, 8e-9),
    "vector_te": (dict(NA=1.35, n_immersion=1.44, imaging_model="vector", polarisation="te",
                       source_type="quadrupole", sigma_outer=0.85, sigma_inner=0.55,
                       source_grid=9, defocus=-50e-9, zernike_coeffs={5: 0.02, 9: -0.01}),
                  64, 4e-9),
    "vector_unpolarised_direct": (dict(NA=1.35, n_immersion=1.44, imaging_model="vector",
                                       polarisation="unpolarised", source_type="annular",
                                       sigma_outer=0.9, sigma_inner=0.6, source_grid=7),
                                  32, 8e-9),
}


@pytest.mark.parametrize("name", list(GRAD_CASES))
@pytest.mark.parametrize("precision", [64, 32])
def test_backward_matches_oracle_adjoint(device, name, precision):
    import torch
    from lithobat_amd.config import GridConfig, OpticsConfig
    from lithobat_amd.plan import ImagingPlan
    kw, n, px = GRAD_CASES[name]
    optics = OpticsConfig(**kw)
    grid = GridConfig(n_pixels=n, pixel_size=px)
    rng = np.random.default_rng(7)
    mask = rng.random((n, n))
    sens = rng.standard_normal((n, n))
    ref = O.abbe_gradient(mask, optics, grid, sens)
    plan = ImagingPlan(optics, grid, precision=precision)
    dt = torch.float64 if precision == 64 else torch.float32
    g = plan.backward(torch.tensor(mask[None], device=device, dtype=dt),
                      torch.tensor(sens[None, None], device=device, dtype=dt))
    tol = 1e-10 if precision == 64 else 2e-5
    assert _rel(g[0].cpu().numpy(), ref) <= tol, _rel(g[0].cpu().numpy(), ref)


@pytest.mark.parametrize("precision", [64, 32])
def test_backward_on_radix5_coarse_grid(device, precision):
    """The adjoint on the 80-point (radix 10 x 8) coarse grid equals the oracle's."""
    import torch
    from lithobat_amd.config import GridConfig, OpticsConfig
    from lithobat_amd.plan import ImagingPlan
    kw, n, px = GRAD_CASES["scalar_quad_w40"]
    optics = OpticsConfig(**kw)
    grid = GridConfig(n_pixels=n, pixel_size=px)
    rng = np.random.default_rng(8)
    mask = rng.random((n, n))
    sens = rng.standard_normal((n, n))
    ref = O.abbe_gradient(mask, optics, grid, sens)
    plan = ImagingPlan(optics, grid, precision=precision, coarse_m=80)
    assert plan.info["m"] == 80
    dt = torch.float64 if precision == 64 else torch.float32
    g = plan.backward(torch.tensor(mask[None], device=device, dtype=dt),
                      torch.tensor(sens[None, None], device=device, dtype=dt))
    tol = 1e-10 if precision == 64 else 2e-5
    assert _rel(g[0].cpu().numpy(), ref) <= tol, _rel(g[0].cpu().numpy(), ref)


@pytest.mark.parametrize("precision", [64, 32])
def test_dense_adjoint_equals_coarse_adjoint_and_oracle(device, precision):
    """The per-system dense adjoint (the reference's dataflow) gives the same gradient."""
    import torch
    from lithobat_amd.config import GridConfig, OpticsConfig
    from lithobat_amd.plan import ImagingPlan
    optics = OpticsConfig(NA=1.35, n_immersion=1.44, source_type="quadrupole", sigma_outer=0.85,
                          sigma_inner=0.55, source_grid=11, defocus=25e-9,
                          zernike_coeffs={6: 0.02})
    grid = GridConfig(n_pixels=128, pixel_size=4e-9)
    rng = np.random.default_rng(17)
    mask = rng.random((128, 128))
    sens = rng.standard_normal((128, 128))
    ref = O.abbe_gradient(mask, optics, grid, sens)
    plan = ImagingPlan(optics, grid, precision=precision)
    dt = torch.float64 if precision == 64 else torch.float32
    m = torch.tensor(mask, device=device, dtype=dt)
    g = torch.tensor(sens, device=device, dtype=dt)
    dense = plan.backward_dense(m, g, batch=3).double().cpu().numpy()
    coarse = plan.backward(m[None], g[None, None])[0].double().cpu().numpy()
    tol = 1e-10 if precision == 64 else 2e-5
    assert _rel(dense, ref) <= tol, _rel(dense, ref)
    assert _rel(coarse, ref) <= tol, _rel(coarse, ref)


def test_autograd_ilt_step_vs_central_differences(device):
    """sigmoid(8 theta) -> Abbe (FP64 on the device) -> relaxed resist L2; FD on 16 pixels."""
    import torch
    from lithobat_amd.config import GridConfig, OpticsConfig
    from lithobat_amd.ilt import abbe_intensity, mask_of, relaxed_resist_loss
    from lithobat_amd.masks import manhattan
    from lithobat_amd.plan import ImagingPlan
    n = 64
    grid = GridConfig(n_pixels=n, pixel_size=8e-9)
    optics = OpticsConfig(NA=1.35, n_immersion=1.44, source_type="quadrupole", sigma_outer=0.85,
                          sigma_inner=0.55, source_grid=11, normalisation="clear")
    plan = ImagingPlan(optics, grid, precision=64)
    scale = 1.0 / float(plan.clear[0])
    rng = np.random.default_rng(3000)
    theta = torch.tensor(rng.standard_normal((n, n)), device=device, requires_grad=True)
    target = torch.tensor(manhattan(n, 8e-9, seed=3001, n_rect=(3, 8)), device=device)

    def loss_of(th):
        raw = abbe_intensity(mask_of(th, 8.0)[None], plan)[:, 0]
        return relaxed_resist_loss(raw, target[None], scale, 0.3, 50.0)

    loss = loss_of(theta)
    loss.backward()
    grad = theta.grad.detach().cpu().numpy()
    idx = rng.choice(n * n, 16, replace=False)
    h = 1e-6
    fd = []
    with torch.no_grad():
        for i in idx:
            tp = theta.detach().clone()
            tm = theta.detach().clone()
            tp.view(-1)[i] += h
            tm.view(-1)[i] -= h
            fd.append((float(loss_of(tp)) - float(loss_of(tm))) / (2 * h))
    fd = np.array(fd)
    err = np.abs(grad.flat[idx] - fd)
    assert np.all(err <= 1e-4 * np.abs(fd).max() + 1e-12), (err, fd)


@pytest.mark.parametrize("precision", [64, 32])
def test_spectrum_entry_points_equal_mask_entry_points(device, precision):
    """Imaging and the adjoint from band spectra computed once (the autograd path) give the
    same bits as the entry points that analyse the masks themselves."""
    import torch
    from lithobat_amd.config import GridConfig, OpticsConfig
    from lithobat_amd.plan import ImagingPlan
    dt = torch.float64 if precision == 64 else torch.float32
    optics = OpticsConfig(NA=1.35, n_immersion=1.44, source_type="quadrupole", sigma_outer=0.85,
                          sigma_inner=0.55, source_grid=11, defocus=20e-9)
    plan = ImagingPlan(optics, GridConfig(n_pixels=128, pixel_size=4e-9),
                       [(-30e-9, 1.0), (10e-9, 1.0)], precision=precision)
    rng = np.random.default_rng(17)
    m = torch.tensor(rng.random((3, 128, 128)), device=device, dtype=dt)
    g = torch.tensor(rng.standard_normal((3, 2, 128, 128)), device=device, dtype=dt)
    spec = plan.mask_spectrum(m)
    a, pa = plan.forward(m)
    b, pb = plan.forward_spectrum(spec)
    assert torch.equal(a, b) and torch.equal(pa, pb)
    assert torch.equal(plan.backward(m, g), plan.backward_spectrum(spec, g))
    with pytest.raises(ValueError, match="spectrum must be"):
        plan.forward_spectrum(spec[:, :, :-1])
