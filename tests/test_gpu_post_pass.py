"""The forward post-pass of a large batch: when a call holds at least half as many images as
the device has CUs, each image's coarse partials are summed and taken to their band spectrum
in one kernel per image (the fused pass), and a real-output synthesis over a band symmetric
about 0 transforms only its columns nu >= 0.  Smaller calls run the separate partial-sum and
band-analysis passes, so a large batch is checked against the same masks imaged a few at a
time, in both work layouts (per-image chunks and the spread runs), and one image of an FP64
batch against the oracle (aerial_image.py:551-587).  Tolerances: FP64 1e-12, FP32 1e-5 of
the largest value (the two paths add the same terms in different orders).
"""
import numpy as np
import pytest

from oracle import litho_oracle as O

pytestmark = pytest.mark.gpu

SPREAD = -2


def _rel(a, b):
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    return float(np.abs(a - b).max() / np.abs(b).max())


def _images_for_fused(planes: int) -> int:
    """Masks whose call holds at least half a CU count of images (plus a margin)."""
    import torch
    cus = torch.cuda.get_device_properties(0).multi_processor_count
    return -(-(cus // 2 + 2) // planes)


def _optics(**kw):
    from lithobat_amd.config import OpticsConfig
    base = dict(NA=0.93, source_type="quadrupole", sigma_outer=0.85, sigma_inner=0.55,
                source_grid=47, normalisation="none")
    base.update(kw)
    return OpticsConfig(**base)


@pytest.mark.parametrize("layout", [0, SPREAD])
def test_large_batch_matches_small_calls(device, layout):
    """1024^2 at 4 nm, NA 0.93 quadrupole (148 points) on the 96-point coarse grid, 5 focus
    planes, FP32: one call of enough masks for the fused pass against calls of 4 masks."""
    import torch
    from bench import DEFOCUS
    from lithobat_amd import autotune
    from lithobat_amd.config import GridConfig
    from lithobat_amd.plan import ImagingPlan
    grid = GridConfig(n_pixels=1024, pixel_size=4e-9)
    plan = ImagingPlan(_optics(), grid, [(dz, 1.0) for dz in DEFOCUS], precision=32)
    assert plan.info["m"] == 96
    n = _images_for_fused(len(DEFOCUS))
    print(f"post-pass FP32 layout={layout}: {n} masks x {len(DEFOCUS)} planes", flush=True)
    rng = np.random.default_rng(500 + n)
    masks = torch.tensor(rng.random((n, 1024, 1024)), device=device, dtype=torch.float32)
    autotune.set_shape(plan, False, plan.info["block_threads"], plan.info["column_splits"],
                       layout)
    try:
        big, big_pk = plan.forward(masks, with_peak=True)
        big, big_pk = big.cpu().numpy(), big_pk.cpu().numpy()
        parts = [plan.forward(masks[i:i + 4], with_peak=True) for i in range(0, n, 4)]
    finally:
        autotune.unpin(plan)
    small = np.concatenate([p[0].cpu().numpy() for p in parts])
    small_pk = np.concatenate([p[1].cpu().numpy() for p in parts])
    assert big.shape == small.shape == (n, len(DEFOCUS), 1024, 1024)
    assert _rel(big, small) <= 1e-5, _rel(big, small)
    assert _rel(big_pk, small_pk) <= 1e-5


def test_large_fp64_batch_matches_oracle(device):
    """FP64, the same geometry at two focus planes: the fused pass's images of one mask
    against the oracle, and the whole batch against calls of 4 masks."""
    import torch
    from lithobat_amd.config import GridConfig
    from lithobat_amd.masks import manhattan
    from lithobat_amd.plan import ImagingPlan
    grid = GridConfig(n_pixels=1024, pixel_size=4e-9)
    dz = (-100e-9, 60e-9)
    plan = ImagingPlan(_optics(), grid, [(d, 1.0) for d in dz], precision=64)
    n = _images_for_fused(len(dz))
    print(f"post-pass FP64: {n} masks x {len(dz)} planes", flush=True)
    rng = np.random.default_rng(9)
    host = rng.random((n, 1024, 1024))
    host[n // 2] = manhattan(1024, 4e-9, seed=1000)
    masks = torch.tensor(host, device=device)
    big = plan.forward(masks, with_peak=False)[0].cpu().numpy()
    small = np.concatenate([plan.forward(masks[i:i + 4], with_peak=False)[0].cpu().numpy()
                            for i in range(0, n, 4)])
    assert _rel(big, small) <= 1e-12
    ref = O.abbe_sum(host[n // 2], _optics(), grid, [(d, 1.0) for d in dz])[0]
    for k in range(len(dz)):
        assert _rel(big[n // 2, k], ref[k]) <= 1e-12, (k, _rel(big[n // 2, k], ref[k]))
