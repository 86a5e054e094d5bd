"""The staged coherent-system path (row tables through device memory, column
transforms in a second launch; the adjoint adds a transposed row pass per row chunk)
against the single-launch kernels, the oracle and central differences.

Same arithmetic stage by stage, so FP64 agrees to <= 1e-12 and FP32 to <= 1e-5 of the
peak (adjoint 2e-5); covers several passes of row tables, chunked work items, systems
with their own boxes and both block sizes.
"""
import numpy as np
import pytest

from oracle import litho_oracle as O

pytestmark = pytest.mark.gpu


def _rel(a, b):
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    return float(np.abs(a - b).max() / np.abs(b).max())


def _euv(**kw):
    from lithobat_amd.config import OpticsConfig
    return OpticsConfig(wavelength=13.5e-9, NA=0.33, source_type="annular", sigma_outer=0.8,
                        sigma_inner=0.4, source_grid=11, normalisation="none", **kw)


def _forward(plan, masks, shape):
    from lithobat_amd import autotune
    autotune.set_shape(plan, False, *shape)
    assert plan.info["column_splits"] == shape[1]
    return plan.forward(masks)


def _single_launch(plan):
    """The first single-launch kernel shape that fits this plan (column splits > 0)."""
    from lithobat_amd import autotune
    nt, cs = next((nt, cs) for nt, cs in autotune.shapes(plan, False) if cs > 0)
    return nt, cs, 0


@pytest.mark.parametrize("precision", [64, 32])
def test_staged_matches_single_launch_and_oracle_m256(device, precision):
    import torch
    from lithobat_amd.config import GridConfig
    from lithobat_amd.masks import manhattan
    from lithobat_amd.plan import ImagingPlan
    grid = GridConfig(n_pixels=512, pixel_size=4e-9)
    optics = _euv()
    planes = [(-40e-9, 1.0), (25e-9, 1.0)]
    plan = ImagingPlan(optics, grid, planes, precision=precision)
    assert plan.info["m"] == 256
    dt = torch.float64 if precision == 64 else torch.float32
    masks_np = np.stack([manhattan(512, 4e-9, seed=70 + i) for i in range(12)])
    masks = torch.tensor(masks_np, device=device, dtype=dt)
    assert plan.info["column_splits"] == 0  # staged by default on this grid
    ref, ref_pk = _forward(plan, masks, _single_launch(plan))
    tol = 1e-12 if precision == 64 else 1e-5
    # 24 images: several passes of row tables in FP64, one in FP32
    for shape in [(256, 0, 0), (256, 0, 5), (512, 0, 0)]:
        img, pk = _forward(plan, masks, shape)
        assert _rel(img.cpu(), ref.cpu()) <= tol, shape
        assert _rel(pk.cpu(), ref_pk.cpu()) <= tol, shape
    if precision == 64:
        o, _, _ = O.compute_aerial_image(masks_np[3], _euv(defocus=25e-9), grid, return_raw=True)
        assert _rel(img[3, 1].cpu(), o) <= 1e-12


def _passes(plan, n_masks):
    """Passes of row tables a staged forward of n_masks masks runs (160 MiB per pass)."""
    info = plan.info
    per_img = info["box"] * info["m"] * (16 if info["precision"] == 64 else 8) * info["n_systems"]
    n_img = n_masks * info["n_planes"]
    return -(-n_img // max(1, min(n_img, (160 << 20) // per_img)))


@pytest.mark.parametrize("precision", [64, 32])
def test_staged_several_passes_m256(device, precision):
    """A batch of several passes of row tables (the last partly filled) against calls of
    2 masks (one pass each) and the single-launch kernel on 2 of the masks, with whole
    items (each column block runs its image's every system, inputs loaded ahead) and
    items of 3 systems, and one item per image (whose partials are the coarse intensities,
    taken by the band analysis as they are); the last against the single-launch kernel."""
    import torch
    from lithobat_amd.config import GridConfig
    from lithobat_amd.masks import manhattan
    from lithobat_amd.plan import ImagingPlan
    grid = GridConfig(n_pixels=512, pixel_size=4e-9)
    plan = ImagingPlan(_euv(), grid, [(-40e-9, 1.0), (25e-9, 1.0)], precision=precision)
    assert plan.info["m"] == 256
    n = 19 if precision == 64 else 37
    print(f"staged passes: FP{precision}, {n} masks, {_passes(plan, n)} passes", flush=True)
    assert _passes(plan, n) >= 2 and _passes(plan, 2) == 1
    dt = torch.float64 if precision == 64 else torch.float32
    masks_np = np.stack([manhattan(512, 4e-9, seed=400 + i) for i in range(n)])
    masks = torch.tensor(masks_np, device=device, dtype=dt)
    tol = 1e-12 if precision == 64 else 1e-5
    spp = plan.info["n_systems"]
    for shape in [(256, 0, 0), (512, 0, 3), (256, 0, spp)]:
        big, big_pk = _forward(plan, masks, shape)
        if shape[2] == spp:
            assert plan.work_items(n) == (spp, 2 * n)  # one work item per image
        parts = [_forward(plan, masks[i:i + 2], shape) for i in range(0, n, 2)]
        small = torch.cat([p[0] for p in parts]).cpu()
        small_pk = torch.cat([p[1] for p in parts]).cpu()
        assert _rel(big.cpu(), small) <= tol, shape
        assert _rel(big_pk.cpu(), small_pk) <= tol, shape
    ref, _ = _forward(plan, masks[n - 2:], _single_launch(plan))
    assert _rel(big[n - 2:].cpu(), ref.cpu()) <= tol


def test_staged_m128_and_single_image(device):
    """The staged shape on a 128-point coarse grid (radix 16 x 8, two stage-2 tasks per
    thread) and a single mask (one pass, chunked items reduced by the split sum)."""
    import torch
    from lithobat_amd.config import GridConfig, OpticsConfig
    from lithobat_amd.masks import manhattan
    from lithobat_amd.plan import ImagingPlan
    grid = GridConfig(n_pixels=512, pixel_size=8e-9)
    optics = OpticsConfig(wavelength=193e-9, NA=1.35, n_immersion=1.44, source_type="annular",
                          sigma_outer=0.9, sigma_inner=0.6, source_grid=15,
                          normalisation="none")
    plan = ImagingPlan(optics, grid, precision=64)
    assert plan.info["m"] == 128
    m_np = manhattan(512, 8e-9, seed=5)
    m = torch.tensor(m_np[None], device=device)
    ref, _ = _forward(plan, m, _single_launch(plan))
    for chunk in (0, 1, 7):
        img, _ = _forward(plan, m, (256, 0, chunk))
        assert _rel(img.cpu(), ref.cpu()) <= 1e-12, chunk
    o, _, _ = O.compute_aerial_image(m_np, optics, grid, return_raw=True)
    assert _rel(img[0, 0].cpu(), o) <= 1e-12


def _backward(plan, masks, g, shape):
    from lithobat_amd import autotune
    autotune.set_shape(plan, True, *shape)
    return plan.backward(masks, g)


def _single_launch_adj(plan):
    from lithobat_amd import autotune
    nt, cs = next((nt, cs) for nt, cs in autotune.shapes(plan, True) if cs > 0)
    return nt, cs, 0


@pytest.mark.parametrize("precision", [64, 32])
def test_staged_adjoint_matches_single_launch_m256(device, precision):
    """Abbe systems with their own boxes (the per-row-chunk band-gradient partials must not
    race), several passes of row tables in FP64, chunked items, both block sizes."""
    import torch
    from lithobat_amd.config import GridConfig
    from lithobat_amd.masks import manhattan
    from lithobat_amd.plan import ImagingPlan
    grid = GridConfig(n_pixels=512, pixel_size=4e-9)
    plan = ImagingPlan(_euv(), grid, [(-40e-9, 1.0), (25e-9, 1.0)], precision=precision)
    dt = torch.float64 if precision == 64 else torch.float32
    masks = torch.tensor(np.stack([manhattan(512, 4e-9, seed=90 + i) for i in range(12)]),
                         device=device, dtype=dt)
    rng = np.random.default_rng(3)
    g = torch.tensor(rng.standard_normal((12, 2, 512, 512)), device=device, dtype=dt)
    ref = _backward(plan, masks, g, _single_launch_adj(plan)).cpu()
    tol = 1e-12 if precision == 64 else 2e-5
    for shape in [(256, 0, 0), (256, 0, 5), (512, 0, 0)]:
        got = _backward(plan, masks, g, shape).cpu()
        assert _rel(got, ref) <= tol, shape


def test_staged_adjoint_socs_and_central_differences(device):
    """The SOCS adjoint at 512^2 / 8 nm (K = 12, one shared box) through the staged path,
    against the single-launch adjoint and a central difference of sum(g * I) (the image is
    quadratic in the mask, so the difference is exact)."""
    import torch
    from lithobat_amd.config import GridConfig, OpticsConfig
    from lithobat_amd.hopkins import SOCSKernels
    from lithobat_amd.masks import manhattan
    optics = OpticsConfig(wavelength=193e-9, NA=1.35, n_immersion=1.44, source_type="annular",
                          sigma_outer=0.9, sigma_inner=0.6, source_grid=11,
                          normalisation="none")
    k = SOCSKernels.build(optics, GridConfig(n_pixels=512, pixel_size=8e-9), tol=0.0,
                          max_kernels=12, precision=64)
    plan = k.plan
    assert plan.info["m"] == 256
    m = torch.tensor(manhattan(512, 8e-9, seed=12)[None], device=device)
    g = torch.tensor(np.random.default_rng(4).standard_normal((1, 1, 512, 512)), device=device)
    ref = _backward(plan, m, g, _single_launch_adj(plan))[0].cpu().numpy()
    got = _backward(plan, m, g, (256, 0, 0))[0].cpu().numpy()
    assert _rel(got, ref) <= 1e-12

    def obj(x):
        return float((plan.forward(x, with_peak=False)[0] * g).sum())

    for (i, j) in [(0, 0), (100, 311), (511, 7)]:
        mp, mm = m.clone(), m.clone()
        mp[0, i, j] += 0.5
        mm[0, i, j] -= 0.5
        assert got[i, j] == pytest.approx(obj(mp) - obj(mm), rel=1e-9,
                                          abs=1e-12 * np.abs(got).max())


def test_staged_shapes_listed(device):
    from lithobat_amd import autotune
    from lithobat_amd.config import GridConfig
    from lithobat_amd.plan import ImagingPlan
    plan = ImagingPlan(_euv(), GridConfig(n_pixels=512, pixel_size=4e-9), precision=32)
    for adjoint in (False, True):
        shapes = autotune.shapes(plan, adjoint)
        assert shapes[0] == (256, 0) and any(cs > 0 for _, cs in shapes)
