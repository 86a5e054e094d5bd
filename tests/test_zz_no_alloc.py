"""lb_forward_static (the forward that allocates nothing, on the library's fixed device
scratch) against lb_forward on the same plans and masks, through the C ABI.

Where lb_forward's own workspace fits the scratch, both run the same launches on the
same layout: images and peaks are identical.  Where it does not, the static call takes
more systems per work item (a different summation order): FP32 images agree to 1e-5 of
the peak.  Two calls in a row, the second with fewer masks, and a call too large for the
scratch (LB_ERR_WORKSPACE).
"""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _masks(n, px, count, seed0):
    from lithobat_amd.masks import manhattan
    return np.stack([manhattan(n, px, seed=seed0 + k, n_rect=(6, 16)) for k in range(count)])


def _static(plan, m, with_peak=True):
    import torch
    from lithobat_amd import _lib
    B = m.shape[0]
    P = len(plan.planes)
    out = torch.empty((B, P, plan.n, plan.n), dtype=plan.real_dtype, device=m.device)
    peak = torch.empty((B, P), dtype=plan.real_dtype, device=m.device) if with_peak else None
    st = torch.cuda.current_stream(m.device).cuda_stream
    rc = _lib.load().lb_forward_static(plan._h, m.data_ptr(), 0, B, out.data_ptr(),
                                       peak.data_ptr() if peak is not None else None, st)
    return rc, out, peak


def _case(device, precision, n, px, optics, planes, counts, seed0):
    import torch
    from lithobat_amd import _lib
    from lithobat_amd.config import GridConfig
    from lithobat_amd.plan import ImagingPlan
    grid = GridConfig(n_pixels=n, pixel_size=px)
    plan = ImagingPlan(optics, grid, planes, precision=precision)
    dt = plan.real_dtype
    results = []
    for k, B in enumerate(counts):
        m = torch.tensor(_masks(n, px, B, seed0 + 100 * k), device=device, dtype=dt)
        want, wpk = plan.forward(m)           # lb_forward (selects the batch's kernel shapes)
        rc, got, gpk = _static(plan, m)
        torch.cuda.synchronize()
        assert rc == _lib.LB_OK, _lib.load().lb_last_error()
        fits = plan.workspace_bytes(B) <= _lib.STATIC_SCRATCH_BYTES
        results.append(fits)
        if fits:
            assert torch.equal(got, want) and torch.equal(gpk, wpk), (B, precision)
        else:
            tol = 1e-12 if precision == 64 else 1e-5
            err = float((got - want).abs().max() / want.abs().max())
            assert err <= tol, (B, err)
            assert float((gpk - wpk).abs().max() / wpk.abs().max()) <= tol
    return results


@pytest.mark.parametrize("precision", [32, 64])
def test_static_forward_equals_forward_two_calls(device, precision):
    from lithobat_amd.config import OpticsConfig
    print(f"static forward: 256^2 quadrupole, precision {precision}, 3 then 1 masks",
          flush=True)
    optics = OpticsConfig(NA=0.93, source_type="quadrupole", sigma_outer=0.85,
                          sigma_inner=0.55, source_grid=21, normalisation="none")
    planes = [(-40e-9, 1.0), (30e-9, 1.0)]
    fits = _case(device, precision, 256, 4e-9, optics, planes, [3, 1], 500)
    assert fits == [True, True]


def test_static_forward_large_batch_takes_bigger_work_items(device):
    """24 masks at 1024^2 (148 source points): lb_forward's workspace exceeds the scratch,
    the static call coarsens its work items and still matches; then 2 masks (fits)."""
    from lithobat_amd.config import OpticsConfig
    print("static forward: 1024^2 quadrupole 148 points, 24 then 2 masks", flush=True)
    optics = OpticsConfig(NA=0.93, source_type="quadrupole", sigma_outer=0.85,
                          sigma_inner=0.55, source_grid=47, normalisation="none")
    fits = _case(device, 32, 1024, 4e-9, optics, None, [24, 2], 700)
    assert fits == [False, True]


def test_static_forward_refuses_what_the_scratch_cannot_hold(device):
    import torch
    from lithobat_amd import _lib
    from lithobat_amd.config import GridConfig, OpticsConfig
    from lithobat_amd.plan import ImagingPlan
    print("static forward: 64 masks at 1024^2 (too large)", flush=True)
    optics = OpticsConfig(NA=0.93, source_type="quadrupole", sigma_outer=0.85,
                          sigma_inner=0.55, source_grid=47, normalisation="none")
    plan = ImagingPlan(optics, GridConfig(n_pixels=1024, pixel_size=4e-9), precision=32)
    m = torch.zeros((64, 1024, 1024), device=device, dtype=torch.float32)
    rc, _, _ = _static(plan, m)
    assert rc == _lib.LB_ERR_WORKSPACE
    assert "static scratch" in _lib.load().lb_last_error().decode()
    lib = _lib.load()
    assert lib.lb_forward_static(None, m.data_ptr(), 0, 1, m.data_ptr(), None, None) == \
        _lib.LB_ERR_ARG
    # the refused call left the scratch usable
    small = torch.tensor(_masks(1024, 4e-9, 1, 900), device=device, dtype=torch.float32)
    want, _ = plan.forward(small)
    rc, got, _ = _static(plan, small, with_peak=False)
    torch.cuda.synchronize()
    assert rc == _lib.LB_OK and torch.equal(got, want)
