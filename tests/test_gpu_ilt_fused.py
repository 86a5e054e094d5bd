"""ILTObjective: theta -> the relaxed-resist cost of the Abbe image of the mask
sigmoid(clip(beta theta, +-60)) in one autograd function (the sigmoid inside the masks' band
analysis, the loss gradient applied on the device with the sigmoid derivative), against the
composed autograd path relaxed_resist_loss(abbe_intensity(mask_of(theta, beta))), which
test_gpu_cfg4.py / test_gpu_ilt.py pin to the oracle and to central differences
(opc/ilt.py:311-389).

The loss is scaled by 2.5 before the backward, so the incoming gradient is not 1.  n = 100
has no fused analysis row pass and takes the sigmoid-then-analysis route; n = 512 at 4 nm
(image band inside |nu| < N/16) evaluates the objective where the images are synthesised
and runs the adjoint from the band spectrum of the intensity derivative, n = 128 at 8 nm
(band too wide for it) writes the images and runs the objective pass.  Tolerances: FP64
<= 1e-12 (value) / 1e-10 (gradient, relative to its largest entry), FP32 <= 1e-6 / 2e-5.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _rel(a, b):
    a, b = a.double(), b.double()
    return float((a - b).abs().max() / b.abs().max().clamp_min(1e-300))


@pytest.mark.parametrize("precision,n,px,tol_v,tol_g", [(64, 128, 8e-9, 1e-12, 1e-10),
                                                        (32, 128, 8e-9, 1e-6, 2e-5),
                                                        (64, 100, 8e-9, 1e-12, 1e-10),
                                                        (64, 512, 4e-9, 1e-12, 1e-10),
                                                        (32, 512, 4e-9, 1e-6, 2e-5)])
def test_fused_objective_matches_the_composed_path(device, precision, n, px, tol_v, tol_g):
    import torch
    from lithobat_amd.config import GridConfig, OpticsConfig
    from lithobat_amd.ilt import abbe_intensity, ilt_objective, mask_of, relaxed_resist_loss
    from lithobat_amd.plan import ImagingPlan
    print(f"fused ILT objective FP{precision} n={n} px={px}", flush=True)
    optics = OpticsConfig(wavelength=193e-9, NA=1.35, n_immersion=1.44, source_type="quadrupole",
                          sigma_outer=0.85, sigma_inner=0.55, source_grid=15)
    plan = ImagingPlan(optics, GridConfig(n_pixels=n, pixel_size=px), precision=precision)
    dt = plan.real_dtype
    rng = np.random.default_rng(40 + n)
    theta = torch.tensor(rng.standard_normal((3, n, n)), device=device, dtype=dt)
    target = torch.tensor((rng.random((3, n, n)) > 0.6).astype(np.float64), device=device,
                          dtype=dt)
    scale = 1.0 / float(plan.clear[0])
    # which objective pass runs: the one inside the synthesis exactly when the band allows
    spec = plan.mask_spectrum_sigmoid(theta, 8.0)
    assert (plan.forward_objective(spec, target, scale, 0.3, 50.0) is not None) == (n == 512)
    th1 = theta.clone().requires_grad_(True)
    loss1 = relaxed_resist_loss(abbe_intensity(mask_of(th1, 8.0), plan), target[:, None], scale,
                                0.3, 50.0)
    (2.5 * loss1).backward()
    th2 = theta.clone().requires_grad_(True)
    loss2 = ilt_objective(th2, plan, target, scale, 0.3, 50.0, beta=8.0)
    (2.5 * loss2).backward()
    assert _rel(loss2.detach(), loss1.detach()) <= tol_v
    assert _rel(th2.grad, th1.grad) <= tol_g
    # a single 2-D theta keeps its shape
    th3 = theta[0].clone().requires_grad_(True)
    ilt_objective(th3, plan, target[0], scale, 0.3, 50.0, beta=8.0).backward()
    assert th3.grad.shape == (n, n)
    ref = th1.grad[0] / 2.5  # the per-image terms are independent
    assert _rel(th3.grad, ref) <= tol_g


def test_fused_objective_replays_in_a_graph(device):
    """The fused step (forward, loss, backward) captured into a graph replays like eager."""
    import torch
    from bench import optics_cfg4
    from lithobat_amd.config import GridConfig
    from lithobat_amd.ilt import ilt_objective
    from lithobat_amd.plan import ImagingPlan
    plan = ImagingPlan(optics_cfg4(), GridConfig(n_pixels=512, pixel_size=4e-9), precision=32)
    rng = np.random.default_rng(3)
    theta = torch.tensor(rng.standard_normal((4, 512, 512)), device=device, dtype=torch.float32,
                         requires_grad=True)
    target = torch.tensor((rng.random((4, 512, 512)) > 0.5).astype(np.float32), device=device)
    scale = 1.0 / float(plan.clear[0])

    def step():
        theta.grad = None
        ilt_objective(theta, plan, target, scale, 0.3, 50.0).backward()

    step()
    torch.cuda.synchronize()
    eager = theta.grad.clone()
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        step()
    torch.cuda.current_stream().wait_stream(side)
    torch.cuda.synchronize()
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        step()
    for _ in range(3):
        graph.replay()
        torch.cuda.synchronize()
        assert torch.equal(theta.grad, eager)
        step()  # eager work between replays
        torch.cuda.synchronize()


def test_fused_objective_over_focus_planes(device):
    """A plan with three focus planes: the objective sums the planes' costs against one
    target per mask (broadcast over the planes), inside the synthesis (512^2 at 4 nm) as in
    the composed path; FP64 <= 1e-12 (value) / 1e-10 (gradient)."""
    import torch
    from lithobat_amd.config import GridConfig, OpticsConfig
    from lithobat_amd.ilt import abbe_intensity, ilt_objective, mask_of, relaxed_resist_loss
    from lithobat_amd.plan import ImagingPlan
    optics = OpticsConfig(wavelength=193e-9, NA=1.35, n_immersion=1.44, source_type="quadrupole",
                          sigma_outer=0.85, sigma_inner=0.55, source_grid=11)
    planes = [(dz, 1.44) for dz in (-40e-9, 0.0, 40e-9)]
    plan = ImagingPlan(optics, GridConfig(n_pixels=512, pixel_size=4e-9), planes, precision=64)
    rng = np.random.default_rng(77)
    theta = torch.tensor(rng.standard_normal((2, 512, 512)), device=device)
    target = torch.tensor((rng.random((2, 512, 512)) > 0.5).astype(np.float64), device=device)
    scale = 1.0 / float(plan.clear[0])
    spec = plan.mask_spectrum_sigmoid(theta, 8.0)
    assert plan.forward_objective(spec, target, scale, 0.3, 50.0) is not None
    th1 = theta.clone().requires_grad_(True)
    loss1 = relaxed_resist_loss(abbe_intensity(mask_of(th1, 8.0), plan), target[:, None], scale,
                                0.3, 50.0)
    (1.7 * loss1).backward()
    th2 = theta.clone().requires_grad_(True)
    loss2 = ilt_objective(th2, plan, target, scale, 0.3, 50.0, beta=8.0)
    (1.7 * loss2).backward()
    assert _rel(loss2.detach(), loss1.detach()) <= 1e-12
    assert _rel(th2.grad, th1.grad) <= 1e-10
