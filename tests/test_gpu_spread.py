"""The spread work layout (LB_CHUNK_SPREAD): a call's (image, system) pairs, in image order,
cut into one equal run per resident block, each run's image segments summed afterwards in
block order.  Same images and gradients as the per-image chunks to FP64 rounding, in the
three regimes of run length: shorter than one image's systems (runs straddle at most one
image boundary), several images per run, and fewer systems than blocks (empty runs).
"""
import numpy as np
import pytest

from oracle import litho_oracle as O

pytestmark = pytest.mark.gpu

SPREAD = -2


def _rel(a, b):
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    return float(np.abs(a - b).max() / np.abs(b).max())


def _optics(**kw):
    from lithobat_amd.config import OpticsConfig
    base = dict(NA=1.35, n_immersion=1.44, source_type="annular", sigma_outer=0.8,
                sigma_inner=0.5, source_grid=9, defocus=20e-9, normalisation="none")
    base.update(kw)
    return OpticsConfig(**base)


def _both(plan, masks, sens, nt, cs):
    """(images, gradients) with automatic work items and with the spread layout."""
    from lithobat_amd import autotune
    out = []
    for ch in (0, SPREAD):
        autotune.set_shape(plan, False, nt, cs, ch)
        autotune.set_shape(plan, True, nt, cs, ch)
        img = plan.forward(masks, with_peak=True)
        grad = plan.backward(masks, sens)
        out.append((img[0].cpu().numpy(), img[1].cpu().numpy(), grad.cpu().numpy()))
    autotune.unpin(plan)
    return out


@pytest.mark.parametrize("n_masks,planes", [(1, 1), (6, 3), (300, 1)])
def test_spread_runs_match_per_image_chunks(device, n_masks, planes):
    """256^2 grid on the 128-point coarse grid (FP64, one 256-thread block per CU): 1 mask has
    fewer systems than blocks, 6 x 3 planes straddles image boundaries, 300 masks (more
    images than blocks) put several images in each run."""
    import torch
    from lithobat_amd.config import GridConfig
    from lithobat_amd.plan import ImagingPlan
    grid = GridConfig(n_pixels=256, pixel_size=4e-9)
    optics = _optics(source_grid=5 if n_masks == 300 else 9)
    pl = [(dz, 1.0) for dz in (-30e-9, 0.0, 30e-9)[:planes]]
    plan = ImagingPlan(optics, grid, pl if planes > 1 else None, precision=64, coarse_m=128)
    rng = np.random.default_rng(n_masks)
    masks = torch.tensor(rng.random((n_masks, 256, 256)), device=device)
    sens = torch.tensor(rng.standard_normal((n_masks, planes, 256, 256)), device=device)
    spp = plan.info["n_systems"]
    total = n_masks * planes * spp
    from lithobat_amd import autotune
    autotune.set_shape(plan, False, 256, 1, SPREAD)
    chunk, items = plan.work_items(n_masks)
    runs = plan.info["grid_blocks"]
    autotune.unpin(plan)
    print(f"spread: {n_masks} masks x {planes} planes, {spp} systems/image, {runs} runs, "
          f"{items // runs} segments/run", flush=True)
    assert chunk == SPREAD and items % runs == 0
    if n_masks == 1:
        assert total < runs
    if n_masks == 300:
        assert total > runs * spp, (total, runs, spp)
    (i0, p0, g0), (i1, p1, g1) = _both(plan, masks, sens, 256, 1)
    assert _rel(i1, i0) <= 1e-12
    assert np.array_equal(p1 > 0, p0 > 0) and _rel(p1, p0) <= 1e-12
    assert _rel(g1, g0) <= 1e-12
    if n_masks == 1:
        ref = O.compute_aerial_image(masks[0].cpu().numpy(), optics, grid, return_raw=True)[0]
        assert _rel(i1[0, 0], ref) <= 1e-12


def test_spread_fp32_headline_shape(device):
    """1024^2 at 4 nm, NA 0.93 quadrupole (148 points), 16 masks x 5 planes, FP32: spread vs
    chunked images and gradients within the FP32 tolerance (1e-5 of the largest value; the
    two sum the systems in different orders), the NA 1.35 plan's adjoint likewise."""
    import torch
    from bench import DEFOCUS, N, PIXEL, optics_cfg2, optics_cfg4
    from lithobat_amd.config import GridConfig
    from lithobat_amd.plan import ImagingPlan
    grid = GridConfig(n_pixels=N, pixel_size=PIXEL)
    plan = ImagingPlan(optics_cfg2(), grid, [(dz, 1.0) for dz in DEFOCUS], precision=32)
    rng = np.random.default_rng(77)
    masks = torch.tensor(rng.random((16, N, N)), device=device, dtype=torch.float32)
    sens = torch.tensor(rng.standard_normal((16, 5, N, N)), device=device, dtype=torch.float32)
    print("spread FP32: 16 masks x 5 planes at 1024^2", flush=True)
    (i0, _, g0), (i1, _, g1) = _both(plan, masks, sens, 256, 1)
    assert _rel(i1, i0) <= 1e-5
    assert _rel(g1, g0) <= 1e-5
    ilt = ImagingPlan(optics_cfg4(), grid, precision=32)
    (_, _, g0), (_, _, g1) = _both(ilt, masks, sens[:, :1], 512, 1)
    assert _rel(g1, g0) <= 1e-5
