"""Boundary behaviour through the C ABI and the Python mirror: complex-mask gradients, the
plan workspace under graph capture, the library's error statuses, a NaN mask, and the
reference's device-path edge cases (clear field, piston, Noll 12-21 and unknown indices).
"""
import ctypes as C
import logging

import numpy as np
import pytest

from oracle import litho_oracle as O

pytestmark = pytest.mark.gpu


def _rel(a, b):
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    return float(np.abs(a - b).max() / np.abs(b).max())


def _small(NA=0.93, **kw):
    from lithobat_amd.config import GridConfig, OpticsConfig
    grid = GridConfig(n_pixels=64, pixel_size=8e-9)
    optics = OpticsConfig(NA=NA, source_type="annular", sigma_outer=0.8, sigma_inner=0.5,
                          source_grid=9, normalisation="none", **kw)
    return optics, grid


def test_complex_mask_gradient(device):
    """A complex (phase-shift) mask through the autograd form: the real part of its
    gradient is the reference's projection 2 Re(adjoint) (the oracle's Abbe adjoint of the
    complex mask), the imaginary part the derivative along i (central differences)."""
    import torch
    from lithobat_amd.ilt import abbe_intensity
    from lithobat_amd.masks import manhattan
    from lithobat_amd.plan import ImagingPlan
    optics, grid = _small(defocus=30e-9, zernike_coeffs={7: 0.02})
    plan = ImagingPlan(optics, grid, precision=64)
    rng = np.random.default_rng(4)
    base = manhattan(64, 8e-9, seed=8, n_rect=(3, 8))
    mnp = np.where(base > 0, -0.245 + 0.1j, 1.0 + 0.0j) * np.exp(0.2j * rng.random((64, 64)))
    sens = rng.standard_normal((64, 64))
    m = torch.tensor(mnp, device=device, requires_grad=True)
    st = torch.tensor(sens, device=device)
    (abbe_intensity(m[None], plan)[0, 0] * st).sum().backward()
    g = m.grad.cpu().numpy()
    ref = O.abbe_gradient(mnp, optics, grid, sens)
    assert _rel(g.real, ref) <= 1e-10

    def I(mm):
        return O.compute_aerial_image(mm, optics, grid, return_raw=True)[0]
    # I is quadratic in the mask, so a central difference is exact for any step up to
    # rounding: a large step keeps that rounding small; the difference of J = sum(sens * I)
    # is formed pixel by pixel before it is summed
    h = 1e-3
    for i, j in rng.integers(0, 64, size=(4, 2)):
        up, dn = mnp.copy(), mnp.copy()
        up[i, j] += 1j * h
        dn[i, j] -= 1j * h
        fd = float(np.sum(sens * (I(up) - I(dn)))) / (2 * h)
        assert g.imag[i, j] == pytest.approx(fd, rel=1e-6, abs=1e-9 * np.abs(g).max())
    # FP32: the same against the FP64 gradient
    p32 = ImagingPlan(optics, grid, precision=32)
    m32 = torch.tensor(mnp, device=device, dtype=torch.complex64, requires_grad=True)
    (abbe_intensity(m32[None], p32)[0, 0] * st.float()).sum().backward()
    assert m32.grad.dtype == torch.complex64
    g32 = m32.grad.cpu().numpy().astype(np.complex128)
    assert float(np.abs(g32 - g).max() / np.abs(g).max()) <= 2e-5


def test_workspace_of_a_captured_graph_survives_a_larger_call(device):
    """A forward captured on the plan's 1-mask workspace still replays correctly after a
    16-mask call moved the plan to a larger workspace, and allocations made in between do
    not land in the memory the graph writes."""
    import torch
    from lithobat_amd.masks import manhattan
    from lithobat_amd.plan import ImagingPlan
    optics, grid = _small()
    plan = ImagingPlan(optics, grid, precision=32)
    m1 = torch.tensor(manhattan(64, 8e-9, seed=1, n_rect=(3, 8)), device=device,
                      dtype=torch.float32)
    want, _ = plan.forward(m1)
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        plan.forward(m1)
    torch.cuda.current_stream().wait_stream(s)
    torch.cuda.synchronize()
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        out, _ = plan.forward(m1)
    small_ws = plan._ws.numel()
    m16 = torch.rand((16, 64, 64), device=device)
    plan.forward(m16)
    assert plan._ws.numel() > small_ws
    fillers = [torch.full((small_ws // 4 + 64,), 7.0, device=device) for _ in range(4)]
    for rep in range(3):
        graph.replay()
        torch.cuda.synchronize()
        assert torch.equal(out, want), rep
        assert all(bool((f == 7.0).all()) for f in fillers), rep


def test_error_statuses_through_the_c_abi(device):
    """LB_ERR_WORKSPACE (workspace too small), LB_ERR_UNSUPPORTED (fields of an Abbe plan),
    LB_ERR_ARG (null plan, no masks, a null plan inside a batch); the message is set."""
    import torch
    from lithobat_amd import _lib
    from lithobat_amd.plan import ImagingPlan
    lib = _lib.load()
    optics, grid = _small()
    plan = ImagingPlan(optics, grid, precision=32)
    m = torch.zeros((1, 64, 64), device=device)
    out = torch.empty((1, 1, 64, 64), device=device)
    need = plan.workspace_bytes(1)
    ws = torch.empty(need, dtype=torch.uint8, device=device)
    st = torch.cuda.current_stream().cuda_stream
    assert lib.lb_forward(plan._h, m.data_ptr(), 0, 1, out.data_ptr(), None, ws.data_ptr(),
                          need - 256, st) == _lib.LB_ERR_WORKSPACE
    assert str(need) in lib.lb_last_error().decode()
    assert lib.lb_forward(plan._h, m.data_ptr(), 0, 1, out.data_ptr(), None, None, 0, st) == \
        _lib.LB_ERR_WORKSPACE
    f = torch.empty((1, 1, 64, 64), dtype=torch.complex64, device=device)
    fw = int(lib.lb_fields_workspace_bytes(plan._h, 1))
    fws = torch.empty(max(fw, 256), dtype=torch.uint8, device=device)
    assert lib.lb_fields(plan._h, m.data_ptr(), 0, 1, f.data_ptr(), fws.data_ptr(), fws.numel(),
                         st) == _lib.LB_ERR_UNSUPPORTED
    assert "SOCS" in lib.lb_last_error().decode()
    assert lib.lb_forward(None, m.data_ptr(), 0, 1, out.data_ptr(), None, ws.data_ptr(), need,
                          st) == _lib.LB_ERR_ARG
    assert lib.lb_forward(plan._h, m.data_ptr(), 0, 0, out.data_ptr(), None, ws.data_ptr(),
                          need, st) == _lib.LB_ERR_ARG
    hs = (C.c_void_p * 2)(plan._h.value, None)
    h = C.c_void_p()
    assert lib.lb_batch_create(hs, 2, C.byref(h)) == _lib.LB_ERR_ARG
    assert "null plan" in lib.lb_last_error().decode() and not h.value
    with pytest.raises(ValueError, match="mask must be"):
        plan.forward(torch.zeros((1, 32, 32), device=device))
    # the plan still images after the refusals
    good, _ = plan.forward(torch.ones((1, 64, 64), device=device))
    torch.cuda.synchronize()
    assert torch.isfinite(good).all()


@pytest.mark.parametrize("precision", [64, 32])
def test_nan_mask_gives_nan_images_and_leaves_the_plan_clean(device, precision):
    """A NaN pixel spreads through the spectrum into every image pixel, as the reference's
    FFTs spread it; the next call on the same plan and workspace is unaffected."""
    import torch
    from lithobat_amd.masks import manhattan
    from lithobat_amd.plan import ImagingPlan
    optics, grid = _small()
    plan = ImagingPlan(optics, grid, precision=precision)
    dt = plan.real_dtype
    m = torch.tensor(manhattan(64, 8e-9, seed=3, n_rect=(3, 8)), device=device, dtype=dt)
    want, _ = plan.forward(m)
    bad = m.clone()
    bad[5, 7] = float("nan")
    img, _ = plan.forward(bad)
    torch.cuda.synchronize()
    assert torch.isnan(img).all()
    g = plan.backward(bad[None], torch.ones((1, 1, 64, 64), device=device, dtype=dt))
    assert torch.isnan(g).all()
    again, _ = plan.forward(m)
    assert torch.equal(again, want)
    mnp = m.double().cpu().numpy()
    mnp[5, 7] = np.nan
    assert np.isnan(O.compute_aerial_image(mnp, optics, grid, return_raw=True)[0]).all()


@pytest.mark.parametrize("kw", [
    {}, dict(imaging_model="vector", polarisation="te"),
    dict(imaging_model="vector", polarisation="tm"), dict(imaging_model="vector", polarisation="x"),
    dict(imaging_model="vector", polarisation="y"),
    dict(imaging_model="vector", polarisation="unpolarised")])
def test_clear_field_prints_one_and_flat(device, kw):
    """tests/test_vector_imaging.py:429-439 of the reference on the device path: an open
    frame under clear normalisation prints exactly the dose, flat, for every polarisation."""
    from lithobat_amd import compute_aerial_image
    from lithobat_amd.config import GridConfig, OpticsConfig
    grid = GridConfig(n_pixels=64, pixel_size=4e-9)
    o = OpticsConfig(NA=1.35, n_immersion=1.44, n_image=1.70, source_grid=11,
                     normalisation="clear", **kw)
    a = compute_aerial_image(np.ones((64, 64)), o, grid)
    assert a.mean() == pytest.approx(1.0, rel=1e-9)
    assert np.ptp(a) == pytest.approx(0.0, abs=1e-12)


def test_piston_leaves_the_image_invariant(device):
    """tests/test_api_expose_pupil.py:181-186 on the device: Z1 is a constant phase."""
    from lithobat_amd import compute_aerial_image
    from lithobat_amd.config import GridConfig, OpticsConfig
    grid = GridConfig(n_pixels=64, pixel_size=8e-9)
    mask = O.lines_and_spaces(64, 8e-9, 200e-9, 100e-9)
    base = compute_aerial_image(mask, OpticsConfig(NA=0.93, sigma_outer=0.7), grid)
    pist = compute_aerial_image(mask, OpticsConfig(NA=0.93, sigma_outer=0.7,
                                                   zernike_coeffs={1: 0.25}), grid)
    assert np.allclose(pist, base, atol=1e-9)
    coma = compute_aerial_image(mask, OpticsConfig(NA=0.93, sigma_outer=0.7,
                                                   zernike_coeffs={7: 0.05}), grid)
    assert np.abs(coma - base).max() > 1e-3


@pytest.mark.parametrize("vector", [False, True])
def test_high_noll_terms_match_the_oracle(device, vector):
    """Noll 12-21 (n = 4, 5 radial orders) on the device against the oracle, scalar and
    vector, FP64 <= 1e-12 of the peak."""
    from lithobat_amd import compute_aerial_image
    from lithobat_amd.config import GridConfig, OpticsConfig
    from lithobat_amd.masks import manhattan
    rng = np.random.default_rng(21)
    zern = {j: float(rng.uniform(-0.03, 0.03)) for j in range(12, 22)}
    grid = GridConfig(n_pixels=64, pixel_size=4e-9)
    kw = dict(imaging_model="vector", polarisation="tm") if vector else {}
    o = OpticsConfig(NA=1.35, n_immersion=1.44, source_type="quadrupole", sigma_outer=0.85,
                     sigma_inner=0.55, source_grid=11, defocus=-25e-9, zernike_coeffs=zern,
                     normalisation="none", **kw)
    mask = manhattan(64, 4e-9, seed=22, n_rect=(4, 10))
    got = compute_aerial_image(mask, o, grid, return_raw=True)[0]
    ref = O.compute_aerial_image(mask, o, grid, return_raw=True)[0]
    assert _rel(got, ref) <= 1e-12
    plain = O.compute_aerial_image(mask, OpticsConfig(**{**o.__dict__, "zernike_coeffs": {}}),
                                   grid, return_raw=True)[0]
    assert np.abs(ref - plain).max() > 1e-3 * ref.max()  # the terms do act


def test_unknown_noll_indices_contribute_zero_with_a_warning(device, caplog):
    """tests/test_api_expose_pupil.py:135-143: indices outside 1..21 return zeros (the image
    is the unaberrated one) and log a warning."""
    from lithobat_amd import compute_aerial_image
    from lithobat_amd.config import GridConfig, OpticsConfig
    grid = GridConfig(n_pixels=64, pixel_size=8e-9)
    mask = O.lines_and_spaces(64, 8e-9, 200e-9, 100e-9)
    base = compute_aerial_image(mask, OpticsConfig(NA=0.93, defocus=20e-9), grid)
    with caplog.at_level(logging.WARNING):
        odd = compute_aerial_image(mask, OpticsConfig(NA=0.93, defocus=20e-9,
                                                      zernike_coeffs={22: 0.05, 40: 0.1,
                                                                      999: -0.2}), grid)
    assert np.array_equal(odd, base)
    msgs = [r.getMessage() for r in caplog.records]
    assert sum("not implemented" in s for s in msgs) >= 3, msgs


@pytest.mark.parametrize("n,px,shift,defocus,zern,exact", [
    (64, 32e-9, (3, -2), 0.0, {4: 0.05}, True),
    (100, 8e-9, (0, 0), 60e-9, {7: 0.03, 12: -0.02}, False),
    (37, 20e-9, (-1, 2), -40e-9, {}, True),
])
def test_dense_pupil_matches_the_oracle(device, n, px, shift, defocus, zern, exact):
    """_build_fft_pupil (aerial_image.py:169-209) on the device against the oracle's
    restatement, FFT order, for on- and off-axis source points, any grid side."""
    from lithobat_amd.config import GridConfig, OpticsConfig
    from lithobat_amd.expose import build_fft_pupil
    optics = OpticsConfig(NA=0.93, defocus=defocus, zernike_coeffs=zern, exact_defocus=exact)
    grid = GridConfig(n_pixels=n, pixel_size=px)
    fsx, fsy = shift[0] / (n * px), shift[1] / (n * px)
    got = build_fft_pupil(optics, grid, fsx, fsy)
    FX, FY = O.make_freq_grid(n, px)
    rho, phi, inside = O.pupil_geometry(FX, FY, optics.wavelength, optics.NA, fsx, fsy,
                                        need_phi=True)
    want = O.build_fft_pupil(rho, phi, inside, optics.wavelength, optics.NA, defocus,
                             optics.n_immersion, zern, exact)
    assert got.shape == (n, n) and got.dtype == np.complex128
    assert np.array_equal(got != 0, want != 0)
    assert np.abs(got - want).max() <= 1e-12


def test_dense_pupil_tilts_the_mask_the_other_way(device):
    """tests/test_aerial_image.py:136-202 of the reference with the device pupil: a source
    point at +f_s images the mask as the on-axis pupil does the mask tilted by
    exp(-i 2 pi f_s x)."""
    from lithobat_amd.config import GridConfig, OpticsConfig
    from lithobat_amd.expose import build_fft_pupil
    n, dx = 64, 32e-9
    mask = np.random.default_rng(1).random((n, n))
    sx, sy = 3, -2
    optics = OpticsConfig(NA=0.93, source_grid=3, normalisation="none",
                          zernike_coeffs={4: 0.05}, exact_defocus=True)
    grid = GridConfig(n_pixels=n, pixel_size=dx)
    eng = np.abs(np.fft.ifft2(np.fft.fft2(mask) * build_fft_pupil(
        optics, grid, sx / (n * dx), sy / (n * dx)))) ** 2
    on_axis = build_fft_pupil(optics, grid)
    X, Y = np.meshgrid(np.arange(n), np.arange(n))

    def imaged(sign):
        t = np.exp(sign * 2j * np.pi * (sx * X + sy * Y) / n)
        return np.abs(np.fft.ifft2(np.fft.fft2(mask * t) * on_axis)) ** 2
    assert np.abs(eng - imaged(-1)).max() < 1e-12 * eng.max()
    assert np.abs(eng - imaged(+1)).max() > 1e-3 * eng.max()


def test_objective_entry_points_refuse_what_they_cannot_run(device):
    """lb_forward_objective: LB_ERR_UNSUPPORTED where the image band reaches |nu| >= N/16
    (64^2 at 8 nm), for a target that is not 16-byte aligned, LB_ERR_WORKSPACE below its
    workspace size (zero when unsupported); lb_backward_sensitivity_band refuses a plan
    imaged directly on its own grid.  Each refusal leaves a message and launches nothing."""
    import torch
    from lithobat_amd import _lib
    from lithobat_amd.config import GridConfig, OpticsConfig
    from lithobat_amd.plan import ImagingPlan
    lib = _lib.load()
    st = torch.cuda.current_stream().cuda_stream
    optics, grid = _small()
    small = ImagingPlan(optics, grid, precision=32)
    spec = small.mask_spectrum(torch.ones((1, 64, 64), device=device))
    tgt = torch.zeros((1, 64, 64), device=device)
    sb = torch.empty((1, 64, 64), dtype=torch.complex64, device=device)
    ws = torch.empty(1 << 20, dtype=torch.uint8, device=device)
    assert lib.lb_forward_objective_workspace_bytes(small._h, 1) == 0
    assert lib.lb_forward_objective(small._h, spec.data_ptr(), 1, tgt.data_ptr(), 1.0, 0.3, 50.0,
                                    1.0, None, None, None, sb.data_ptr(), None, None,
                                    ws.data_ptr(), ws.numel(), st) == _lib.LB_ERR_UNSUPPORTED
    assert "N/16" in lib.lb_last_error().decode()
    assert small.forward_objective(spec, tgt, 1.0, 0.3, 50.0) is None
    if small.info["m"] == small.info["n"]:
        assert lib.lb_backward_sensitivity_band(small._h, spec.data_ptr(), 1, sb.data_ptr(),
                                                tgt.data_ptr(), ws.data_ptr(), ws.numel(),
                                                st) == _lib.LB_ERR_UNSUPPORTED
    o = OpticsConfig(NA=1.35, n_immersion=1.44, source_type="quadrupole", sigma_outer=0.85,
                     sigma_inner=0.55, source_grid=9)
    plan = ImagingPlan(o, GridConfig(n_pixels=512, pixel_size=4e-9), precision=32)
    spec = plan.mask_spectrum(torch.ones((1, 512, 512), device=device))
    need = int(lib.lb_forward_objective_workspace_bytes(plan._h, 1))
    assert need > 0
    wb = 2 * plan.info["box"] - 1
    assert lib.lb_sensitivity_band_bytes(plan._h, 1) == 8 * wb * wb
    sb = torch.empty((1, wb, wb), dtype=torch.complex64, device=device)
    tgt = torch.zeros(512 * 512 + 4, device=device)
    big = torch.empty(need, dtype=torch.uint8, device=device)
    args = (1.0, 0.3, 50.0, 1.0, None, None, None, sb.data_ptr(), None, None)
    assert lib.lb_forward_objective(plan._h, spec.data_ptr(), 1, tgt.data_ptr() + 4, *args,
                                    big.data_ptr(), need, st) == _lib.LB_ERR_UNSUPPORTED
    assert "aligned" in lib.lb_last_error().decode()
    assert lib.lb_forward_objective(plan._h, spec.data_ptr(), 1, tgt.data_ptr(), *args,
                                    big.data_ptr(), need - 256, st) == _lib.LB_ERR_WORKSPACE
    assert str(need) in lib.lb_last_error().decode()
    assert lib.lb_forward_objective(plan._h, spec.data_ptr(), 1, tgt.data_ptr(), *args,
                                    big.data_ptr(), need, st) == _lib.LB_OK
    torch.cuda.synchronize()
    assert torch.isfinite(torch.view_as_real(sb)).all()
    # the theta-gradient pass: unaligned theta refused, aligned theta runs
    bw = torch.empty(plan.workspace_bytes(1, True), dtype=torch.uint8, device=device)
    gt = torch.empty(512 * 512 + 4, device=device)
    assert lib.lb_backward_objective_theta(plan._h, spec.data_ptr(), 1, sb.data_ptr(),
                                           tgt.data_ptr() + 4, 8.0, None, gt.data_ptr(),
                                           bw.data_ptr(), bw.numel(), st) == _lib.LB_ERR_UNSUPPORTED
    assert "aligned" in lib.lb_last_error().decode()
    assert lib.lb_backward_objective_theta(plan._h, spec.data_ptr(), 1, sb.data_ptr(),
                                           tgt.data_ptr(), 8.0, None, gt.data_ptr(),
                                           bw.data_ptr(), bw.numel(), st) == _lib.LB_OK
    torch.cuda.synchronize()
    assert torch.isfinite(gt[:512 * 512]).all()
    assert lib.lb_backward_objective_theta(small._h, spec.data_ptr(), 1, sb.data_ptr(),
                                           tgt.data_ptr(), 8.0, None, gt.data_ptr(),
                                           bw.data_ptr(), bw.numel(), st) == _lib.LB_ERR_UNSUPPORTED
