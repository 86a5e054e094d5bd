"""Depth-resolved image through the resist film: device vs the float64 oracle."""
import numpy as np
import pytest

from oracle import litho_oracle as O

pytestmark = pytest.mark.gpu


def _rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.abs(a - b).max() / np.abs(b).max())


@pytest.mark.parametrize("ref", ["top", "mid", "bottom"])
def test_exposure_volume_matches_oracle(device, ref):
    import torch
    from lithobat_amd import exposure_volume
    from lithobat_amd.config import GridConfig, OpticsConfig, ResistConfig
    from lithobat_amd.masks import manhattan
    grid = GridConfig(n_pixels=64, pixel_size=4e-9)           # dz 2 nm, 21 slices (defaults)
    optics = OpticsConfig(NA=1.35, n_immersion=1.44, source_type="quadrupole", sigma_outer=0.85,
                          sigma_inner=0.55, source_grid=11, defocus=15e-9,
                          zernike_coeffs={5: 0.02}, exact_defocus=True)
    resist = ResistConfig(thickness=100e-9, n_resist=1.7, focus_reference=ref)
    mask = manhattan(64, 4e-9, seed=40, n_rect=(4, 10))
    ref_vol, ref_z = O.exposure_volume(mask, optics, grid, resist, dose=1.3)
    vol, z = exposure_volume(mask, optics, grid, resist, dose=1.3)
    assert vol.shape == ref_vol.shape == (50, 64, 64) and vol.dtype == np.float64
    assert np.array_equal(z, ref_z)
    assert _rel(vol, ref_vol) <= 1e-12, _rel(vol, ref_vol)
    t = torch.tensor(mask, device=device, dtype=torch.float32)
    vol32, z32 = exposure_volume(t, optics, grid, resist, dose=1.3)
    assert vol32.dtype == torch.float32 and vol32.device.type == "cuda"
    assert _rel(vol32.double().cpu().numpy(), ref_vol) <= 1e-5


@pytest.mark.parametrize("norm", ["peak", "none"])
def test_exposure_volume_normalisations_and_batch(device, norm):
    """Each plane's own peak (or no) normalisation folded into the interpolation pass, for
    a single NumPy mask (FP64) and a batch of device masks (one call, FP32)."""
    import torch
    from lithobat_amd import exposure_volume
    from lithobat_amd.config import GridConfig, OpticsConfig, ResistConfig
    from lithobat_amd.masks import manhattan
    grid = GridConfig(n_pixels=64, pixel_size=4e-9, dz=3e-9, n_z_slices=9)
    optics = OpticsConfig(NA=1.35, n_immersion=1.44, source_type="annular", sigma_outer=0.9,
                          sigma_inner=0.6, source_grid=9, defocus=-20e-9, normalisation=norm)
    resist = ResistConfig(thickness=80e-9, n_resist=1.7, focus_reference="mid")
    masks = [manhattan(64, 4e-9, seed=70 + k, n_rect=(3, 8)) for k in range(3)]
    refs = [O.exposure_volume(m, optics, grid, resist, dose=0.8)[0] for m in masks]
    vol, _ = exposure_volume(masks[0], optics, grid, resist, dose=0.8)
    assert _rel(vol, refs[0]) <= 1e-12
    t = torch.tensor(np.stack(masks), device=device, dtype=torch.float32)
    vb, _ = exposure_volume(t, optics, grid, resist, dose=0.8)
    assert tuple(vb.shape) == (3,) + refs[0].shape
    for b in range(3):
        assert _rel(vb[b].double().cpu().numpy(), refs[b]) <= 1e-5
