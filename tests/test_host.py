"""Host-side logic (no GPU): source tables, configuration, the C-ABI library's exports."""
import ctypes
import json
import os
import re

import numpy as np
import pytest

from lithobat_amd import source as S
from lithobat_amd.config import GridConfig, OpticsConfig

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden")


def test_package_source_tables_equal_the_reference_bit_for_bit():
    z = np.load(os.path.join(GOLD, "sources.npz"))
    names = sorted({k.split("__")[0] for k in z.files if "__" in k})
    for name in names:
        d = json.loads(str(z[f"{name}__optics"]))
        d["zernike_coeffs"] = {}
        s = S.source_points(OpticsConfig(**d))
        for f in ("weight", "fx", "fy", "phi"):
            assert np.array_equal(getattr(s, f), z[f"{name}__{f}"]), (name, f)


@pytest.mark.parametrize("kind", ["monopole", "quasar", "cquad"])
def test_sector_sources_are_normalised_and_inside_the_pupil(kind):
    src = S.build_source(31, kind, 0.9, 0.6)
    assert src.sum() == pytest.approx(1.0)
    rho = S.pupil_grid(31)[0]
    assert np.all(src[rho > 1.0] == 0)
    assert np.all(src[(src > 0)] > 0)


def test_rotating_a_dipole_transposes_it():
    """rows = eta (expose/source.py conventions, tests/test_source.py:200-209)."""
    x = S.build_source(21, "dipole", 0.9, 0.5, axis="x")
    y = S.build_source(21, "dipole", 0.9, 0.5, axis="y")
    assert np.array_equal(x.T, y)


def test_config_validation_messages():
    with pytest.raises(ValueError, match="exceeds the image-space index"):
        OpticsConfig(NA=1.35)
    with pytest.raises(ValueError, match="imaging_model"):
        OpticsConfig(imaging_model="full")
    with pytest.raises(ValueError, match="polarisation"):
        OpticsConfig(polarisation="circular")
    with pytest.raises(ValueError, match="Noll"):
        OpticsConfig(zernike_coeffs={0: 0.1})
    with pytest.raises(TypeError):
        GridConfig(n_pixels=12.5)
    assert OpticsConfig(NA=1.35, n_immersion=1.44).image_index == 1.44


def test_normalisation_scale_modes():
    from lithobat_amd.expose import normalisation_scale
    assert normalisation_scale("peak", 2.0, 4.0, 1.5) == 0.75
    assert normalisation_scale("clear", 2.0, 4.0, 1.5) == 0.375
    assert normalisation_scale("none", 2.0, 4.0, 1.5) == 1.5
    with pytest.raises(ValueError, match="normalisation must be"):
        normalisation_scale("max", 1, 1)


def test_library_exports_every_declared_symbol():
    from lithobat_amd import _lib
    lib = _lib.load()
    header = open(os.path.join(ROOT, "include", "lithobat.h")).read()
    declared = set(re.findall(r"^\s*(?:lb_status|int32_t|const char\*|size_t)\s+(lb_\w+)\(",
                              header, re.M))
    assert len(declared) >= 16
    for name in declared:
        assert hasattr(lib, name), name
        assert isinstance(getattr(lib, name), ctypes._CFuncPtr)
    assert set(_lib.EXPORTS) == declared
    assert lib.lb_version() >= 100


def test_plan_creation_refuses_without_a_device():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a device is present")
    from lithobat_amd.plan import ImagingPlan
    with pytest.raises(RuntimeError, match="HIP device"):
        ImagingPlan(OpticsConfig(), GridConfig(n_pixels=64))


def test_the_package_never_imports_the_oracle():
    pkg = os.path.join(ROOT, "lithobat_amd")
    for dirpath, _, files in os.walk(pkg):
        for f in files:
            if f.endswith(".py"):
                text = open(os.path.join(dirpath, f)).read()
                assert "oracle" not in re.sub(r"#.*", "", text).replace('"""', ""), f


def test_thick_mask_refusals_before_any_launch():
    """The reflective mask reaches the Abbe path only: SOCS kernels, a SOCS print model and
    the ILT driver refuse it with the reference's reasons (hopkins.py:130-135,
    ilt.py:490-498), and a multilayer on a transmissive wavelength has no mirror."""
    from lithobat_amd.config import GridConfig, OpticsConfig, ResistConfig
    from lithobat_amd.hopkins import SOCSKernels
    from lithobat_amd.opc import PrintModel, run_ilt
    o = OpticsConfig(wavelength=13.5e-9, NA=0.33, mask_model="multilayer")
    grid = GridConfig(n_pixels=64, pixel_size=4e-9)
    assert not SOCSKernels.supports(o)
    with pytest.raises(ValueError, match="thin-mask process only"):
        SOCSKernels.build(o, grid)
    with pytest.raises(ValueError, match="scalar thin-mask process"):
        PrintModel(optics=o, resist=ResistConfig(), grid=grid, imaging="socs")
    with pytest.raises(ValueError, match="no thick-mask form"):
        run_ilt(np.zeros((64, 64)), PrintModel(optics=o, resist=ResistConfig(), grid=grid))
