"""Kernel-shape selection follows the batch size: a plan alternating between batch sizes
runs each with the shapes recorded for it (or its defaults), not the last size's."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_shapes_follow_the_batch_size(device):
    import torch
    from lithobat_amd import autotune
    from lithobat_amd.config import GridConfig, OpticsConfig
    from lithobat_amd.plan import ImagingPlan
    # the 1024^2 / NA 1.35 quadrupole ILT plan has recorded shapes for 1 and 16 masks
    optics = OpticsConfig(wavelength=193e-9, NA=1.35, n_immersion=1.44,
                          source_type="quadrupole", sigma_outer=0.85, sigma_inner=0.55,
                          source_grid=47)
    plan = ImagingPlan(optics, GridConfig(n_pixels=1024, pixel_size=4e-9), precision=32)
    table = autotune.load_table()
    k16, k1 = (autotune.key_of(plan, b, False) for b in (16, 1))
    if k16 not in table or k1 not in table:
        pytest.skip("no recorded shapes for this plan")
    m = torch.rand((16, 1024, 1024), device=device)

    def shape():
        i = plan.info
        return i["block_threads"], i["column_splits"], i["chunk"]

    want16 = (table[k16]["threads"], table[k16]["splits"], table[k16]["chunk"])
    want1 = (table[k1]["threads"], table[k1]["splits"], table[k1]["chunk"])
    assert want16 != want1
    ref1 = plan.forward(m[:1])[0].clone()
    assert shape() == want1
    big = plan.forward(m)[0]
    assert shape() == want16
    one = plan.forward(m[:1])[0]
    assert shape() == want1
    assert torch.equal(one, ref1)
    assert float((big[0] - ref1[0]).abs().max() / ref1.abs().max()) <= 1e-5
    # shapes set by hand stay until handed back
    autotune.set_shape(plan, False, *want16)
    plan.forward(m[:1])
    assert shape() == want16
    autotune.unpin(plan)
    plan.forward(m[:1])
    assert shape() == want1
