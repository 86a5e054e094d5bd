"""Pins the float64 oracle (oracle/litho_oracle.py) before anything is checked against it.

Each oracle function is held to something independent of its own formula:
the golden vectors written by the reference (tests/golden), closed forms,
invariants, or a brute-force direct DFT.  A few tests plant a plausible
mistake in the oracle and check that the pins catch it.
"""
import json
import math
import os

import numpy as np
import pytest

from oracle import litho_oracle as O

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def _optics(d):
    d = dict(d)
    d["zernike_coeffs"] = {int(k): v for k, v in (d.get("zernike_coeffs") or {}).items()}
    return O.Optics(**d)


# ------------------------------------------------------------------ grids

def test_freq_grid_is_the_fftfreq_definition():
    n, px = 12, 5e-9
    FX, FY = O.make_freq_grid(n, px)
    nu = np.array([k if k <= (n - 1) // 2 else k - n for k in range(n)], dtype=float)
    assert np.array_equal(FX[0], nu * (1.0 / (n * px)))
    assert np.array_equal(FY[:, 0], nu * (1.0 / (n * px)))
    assert np.all(FX == FX[0][None, :]) and np.all(FY == FY[:, 0][:, None])


def test_pupil_grid_rows_are_eta():
    rho, phi, xi, eta = O.pupil_grid(5)
    assert xi[0, 0] == -1 and xi[0, -1] == 1 and eta[0, 0] == -1 and eta[-1, 0] == 1
    assert rho[2, 2] == 0 and phi[-1, 2] == pytest.approx(np.pi / 2)


# ----------------------------------------------------------------- pupil

def test_zernike_low_orders_closed_forms():
    rng = np.random.default_rng(0)
    rho, phi = rng.random(50), rng.uniform(-np.pi, np.pi, 50)
    assert np.allclose(O.zernike_noll(1, rho, phi), 1.0, atol=1e-14)
    assert np.allclose(O.zernike_noll(2, rho, phi), 2 * rho * np.cos(phi), atol=1e-13)
    assert np.allclose(O.zernike_noll(3, rho, phi), 2 * rho * np.sin(phi), atol=1e-13)
    assert np.allclose(O.zernike_noll(4, rho, phi), np.sqrt(3) * (2 * rho ** 2 - 1), atol=1e-13)
    assert np.allclose(O.zernike_noll(11, rho, phi),
                       np.sqrt(5) * (6 * rho ** 4 - 6 * rho ** 2 + 1), atol=1e-12)
    assert np.allclose(O.zernike_noll(22, rho, phi), 0.0)


def test_zernike_noll_set_is_orthonormal_on_the_disc():
    # polar quadrature: Gauss-Legendre in rho (exact for these polynomials), uniform in phi
    x, w = np.polynomial.legendre.leggauss(16)
    r = 0.5 * (x + 1)
    wr = 0.5 * w * r
    phi = np.linspace(0, 2 * np.pi, 64, endpoint=False)
    R, P = np.meshgrid(r, phi, indexing="ij")
    Wt = wr[:, None] * (2 * np.pi / 64) / np.pi
    Z = np.stack([O.zernike_noll(j, R, P) for j in range(1, 22)])
    G = np.einsum("aij,bij,ij->ab", Z, Z, Wt)
    assert np.allclose(G, np.eye(21), atol=1e-10)


def test_defocus_paraxial_closed_form_and_zero():
    rho = np.linspace(0, 1, 11)
    lam, NA, n, dz = 193e-9, 0.93, 1.0, 150e-9
    assert np.allclose(O.defocus_opd(rho, lam, NA, dz, n),
                       np.pi * NA ** 2 * rho ** 2 * dz / (lam * n), rtol=1e-12)
    for exact in (False, True):
        assert np.array_equal(O.defocus_opd(rho, lam, 1.35, 0.0, 1.44, exact), np.zeros(11))


def test_exact_defocus_understates_a_third_at_immersion():
    par = O.defocus_opd(np.array([1.0]), 193e-9, 1.35, 100e-9, 1.44, exact=False)[0]
    exa = O.defocus_opd(np.array([1.0]), 193e-9, 1.35, 100e-9, 1.44, exact=True)[0]
    assert 30.0 < (exa - par) / exa * 100 < 35.0
    # exact = (2 pi / lam) n dz (1 - cos theta), sin theta = NA/n
    assert exa == pytest.approx(2 * np.pi / 193e-9 * 1.44 * 100e-9
                                * (1 - np.sqrt(1 - (1.35 / 1.44) ** 2)), rel=1e-13)


def test_vector_two_beam_laws():
    """TE overlap 1, TM overlap cos 2 theta, |v_z| = sin theta (TM), 0 (TE)."""
    stm = 1.35 / 1.44
    for rho in (0.3, 0.7, 1.0):
        th = np.arcsin(rho * stm)
        # two beams at +-theta along x; TE polarisation is along y
        a = [np.array(O.vector_coefficients(rho, p, 0.0, 1.0, stm, obliquity=False)).real
             for p in (0.0, np.pi)]
        assert np.dot(a[0], a[1]) == pytest.approx(1.0, abs=1e-12)
        assert abs(a[0][2]) < 1e-15
        b = [np.array(O.vector_coefficients(rho, p, 1.0, 0.0, stm, obliquity=False)).real
             for p in (0.0, np.pi)]
        assert np.dot(b[0], b[1]) == pytest.approx(np.cos(2 * th), abs=1e-12)
        assert abs(b[0][2]) == pytest.approx(np.sin(th), abs=1e-12)
        # obliquity scales power by cos theta
        c = np.array(O.vector_coefficients(rho, 0.4, 0.6, 0.8, stm, obliquity=True)).real
        assert np.sum(c ** 2) == pytest.approx(np.cos(th), abs=1e-12)


# ---------------------------------------------------------------- sources

def test_source_tables_match_the_reference():
    z = np.load(os.path.join(GOLD, "sources.npz"))
    names = sorted({k.split("__")[0] for k in z.files if "__" in k})
    assert len(names) == 5
    for name in names:
        o = _optics(json.loads(str(z[f"{name}__optics"])))
        s = O.source_points(o)
        for f in ("weight", "fx", "fy", "phi"):
            assert np.array_equal(getattr(s, f), z[f"{name}__{f}"]), (name, f)
    counts = {n: z[f"{n}__weight"].size for n in names}
    assert counts["annular_0.5_0.8_g41_na0.85"] == 488
    assert counts["quad_0.85_0.55_g47_na0.93"] == 148
    assert counts["annular_0.9_0.6_g21_na1.35"] == 142


def test_sources_are_normalised_and_bounded():
    o = O.Optics(source_type="annular", sigma_outer=0.8, sigma_inner=0.5, source_grid=15)
    s = O.source_points(o)
    k = o.NA / o.wavelength
    assert s.weight.sum() == pytest.approx(1.0)
    r = np.hypot(s.fx, s.fy)
    assert np.all(r <= k * (1 + 1e-12)) and np.all(r >= 0.5 * k * (1 - 1e-12))


# ------------------------------------------------------------- Abbe sum

def test_abbe_sum_equals_the_direct_dft():
    """Independent of any FFT: explicit exponentials, defocus and aberrations on."""
    g = O.Grid(16, 25e-9)
    rng = np.random.default_rng(3)
    mask = (rng.random((16, 16)) > 0.6).astype(float)
    o = O.Optics(NA=0.93, sigma_outer=0.7, source_grid=7, defocus=60e-9,
                 zernike_coeffs={7: 0.03, 5: -0.02}, exact_defocus=True, normalisation="none")
    raw, peak, _ = O.compute_aerial_image(mask, o, g, return_raw=True)
    ref = O.abbe_direct_dft(mask, o, g)
    assert np.abs(raw - ref).max() < 1e-13 * ref.max()


def test_source_point_tilts_the_mask_the_other_way():
    """aerial_image.py:60-73 / tests/test_aerial_image.py:136-202, restated."""
    n, dx = 64, 32e-9
    mask = np.random.default_rng(1).random((n, n))
    sx, sy = 3, -2
    lam, NA = 193e-9, 0.93
    fsx, fsy = sx / (n * dx), sy / (n * dx)
    FX, FY = O.make_freq_grid(n, dx)

    def pupil(fx0, fy0):
        rho, phi, inside = O.pupil_geometry(FX, FY, lam, NA, fx0, fy0, need_phi=True)
        W = O.pupil_phase(rho, phi, lam, NA, 0.0, 1.0, {4: 0.05}, True)
        P = O.build_fft_pupil(rho, phi, inside, lam, NA, 0.0, 1.0, {4: 0.05}, True)
        assert np.array_equal(P, np.where(inside, np.exp(1j * W), 0))
        assert P[0, 0] != 0 and np.count_nonzero(P) == np.count_nonzero(inside)
        return P

    eng = np.abs(np.fft.ifft2(np.fft.fft2(mask) * pupil(fsx, fsy))) ** 2
    X, Y = np.meshgrid(np.arange(n), np.arange(n))

    def tilted(sign):
        t = np.exp(sign * 2j * np.pi * (sx * X + sy * Y) / n)
        return np.abs(np.fft.ifft2(np.fft.fft2(mask * t) * pupil(0.0, 0.0))) ** 2

    assert np.abs(eng - tilted(-1)).max() < 1e-12 * eng.max()
    assert np.abs(eng - tilted(+1)).max() > 1e-3 * eng.max()


@pytest.mark.parametrize("pol", ["te", "tm", "x", "y", "unpolarised"])
def test_clear_field_prints_one_and_flat(pol):
    g = O.Grid(32, 4e-9)
    model = "scalar" if pol == "x" else "vector"
    o = O.Optics(NA=1.35, n_immersion=1.44, imaging_model=model, polarisation=pol,
                 source_type="annular", sigma_outer=0.9, sigma_inner=0.6, source_grid=9,
                 normalisation="clear")
    img = O.compute_aerial_image(np.ones((32, 32)), o, g)
    assert img.mean() == pytest.approx(1.0, rel=1e-9)
    assert np.ptp(img) < 1e-12


def test_planes_equal_single_images():
    g = O.Grid(32, 8e-9)
    mask = O.lines_and_spaces(32, 8e-9, 128e-9, 64e-9)
    o = O.Optics(NA=0.93, source_type="annular", sigma_outer=0.85, sigma_inner=0.5,
                 source_grid=9, zernike_coeffs={7: 0.03})
    planes = [(-100e-9, 1.0), (45e-9, 1.2)]
    stack = O.compute_aerial_planes(mask, o, g, planes)
    for p, (dz, ni) in enumerate(planes):
        one = O.compute_aerial_image(mask, O.Optics(**{**o.__dict__, "defocus": dz,
                                                       "n_image": ni}), g)
        assert np.abs(stack[p] - one).max() < 1e-12 * one.max()


def _gold_vector():
    return np.load(os.path.join(GOLD, "vector64.npz"))


def test_vector_images_match_the_reference():
    z = _gold_vector()
    g = O.Grid(64, 4e-9)
    for c in z["cases"]:
        o = _optics(json.loads(str(z[f"{c}__optics"])))
        raw, pk, cl = O.compute_aerial_image(z["mask"], o, g, return_raw=True)
        assert np.abs(raw - z[f"{c}__raw"]).max() <= 1e-12 * pk, c
        assert cl == pytest.approx(float(z[f"{c}__clear"]), rel=1e-12)


def test_scalar_planes_match_the_reference():
    z = np.load(os.path.join(GOLD, "scalar128.npz"))
    g = O.Grid(128, 4e-9)
    planes = [tuple(p) for p in z["planes"]]
    for c in z["cases"]:
        o = _optics(json.loads(str(z[f"{c}__optics"])))
        st = O.compute_aerial_planes(z["mask"], o, g, planes)
        assert np.abs(st - z[f"{c}__raw"]).max() <= 1e-12 * st.max(), c
    o = _optics(json.loads(str(z["psm__optics"])))
    raw, pk, cl = O.compute_aerial_image(z["psm_mask"], o, g, return_raw=True)
    assert np.abs(raw - z["psm__raw"]).max() <= 1e-12 * pk


# ------------------------------------------------------------------- SOCS

@pytest.fixture(scope="module")
def socs_gold():
    z = np.load(os.path.join(GOLD, "socs64.npz"))
    o = _optics(json.loads(str(z["optics"])))
    return z, o, O.Grid(64, 8e-9)


def test_socs_eigenvalues_and_images_match_the_reference(socs_gold):
    z, o, g = socs_gold
    for name, kw in (("full", dict(tol=0.0)), ("k24", dict(tol=0.0, max_kernels=24)),
                     ("default", {})):
        k = O.socs_build(o, g, **kw)
        ref_w = z[f"{name}__weights"]
        top = min(40, ref_w.size)
        assert np.allclose(k.weights[:top], ref_w[:top], rtol=1e-10, atol=1e-14 * ref_w[0])
        raw, pk, cl = O.socs_raw(k, z["mask"])
        assert np.abs(raw - z[f"{name}__raw"]).max() <= 1e-11 * pk, name
        assert cl == pytest.approx(float(z[f"{name}__clear"]), rel=1e-10)
        assert k.n_band == int(z[f"{name}__n_band"])


def test_full_rank_socs_is_the_abbe_sum(socs_gold):
    z, o, g = socs_gold
    k = O.socs_build(o, g, tol=0.0)
    raw_k, pk_k, cl_k = O.socs_raw(k, z["mask"])
    raw_a, pk_a, cl_a = O.compute_aerial_image(z["mask"], o, g, return_raw=True)
    assert np.abs(raw_k - raw_a).max() < 1e-12 * pk_a
    assert cl_k == pytest.approx(cl_a, rel=1e-12)


# ---------------------------------------------------------------- adjoints

def _fd(f, x, idx, h=1e-6):
    out = []
    for i in idx:
        xp, xm = x.copy(), x.copy()
        xp.flat[i] += h
        xm.flat[i] -= h
        out.append((f(xp) - f(xm)) / (2 * h))
    return np.array(out)


def test_socs_gradient_matches_finite_differences():
    g = O.Grid(32, 8e-9)
    k = O.socs_build(O.Optics(), g, tol=1e-6)
    rng = np.random.default_rng(5)
    mask = rng.random((32, 32))
    sens = rng.standard_normal((32, 32))
    fields, _ = O.socs_fields(k, mask)
    grad = O.socs_gradient(k, fields, sens)
    idx = rng.choice(32 * 32, 8, replace=False)
    fd = _fd(lambda m: float(np.sum(sens * O.socs_fields(k, m)[1])), mask, idx)
    assert np.allclose(grad.flat[idx], fd, rtol=1e-5, atol=1e-9)


@pytest.mark.parametrize("model,pol", [("scalar", "x"), ("vector", "te"), ("vector", "unpolarised")])
def test_abbe_gradient_matches_finite_differences(model, pol):
    g = O.Grid(32, 8e-9)
    o = O.Optics(NA=1.35, n_immersion=1.44, imaging_model=model, polarisation=pol,
                 source_type="quadrupole", sigma_outer=0.85, sigma_inner=0.55, source_grid=9,
                 defocus=30e-9, zernike_coeffs={8: 0.02})
    rng = np.random.default_rng(6)
    mask = rng.random((32, 32))
    sens = rng.standard_normal((32, 32))
    grad = O.abbe_gradient(mask, o, g, sens)
    idx = rng.choice(32 * 32, 6, replace=False)
    fd = _fd(lambda m: float(np.sum(sens * O.compute_aerial_image(m, o, g, return_raw=True)[0])),
             mask, idx)
    assert np.allclose(grad.flat[idx], fd, rtol=1e-5, atol=1e-9)


def test_abbe_gradient_equals_full_rank_socs_gradient():
    g = O.Grid(32, 8e-9)
    o = O.Optics(NA=0.93, source_type="annular", sigma_outer=0.8, sigma_inner=0.5, source_grid=9)
    k = O.socs_build(o, g, tol=0.0)
    rng = np.random.default_rng(8)
    mask = rng.random((32, 32))
    sens = rng.standard_normal((32, 32))
    a = O.abbe_gradient(mask, o, g, sens)
    s = O.socs_gradient(k, O.socs_fields(k, mask)[0], sens)
    assert np.abs(a - s).max() < 1e-10 * np.abs(a).max()


# ------------------------------------------------- planted mistakes are caught

def test_a_wrong_tilt_sign_is_caught(monkeypatch):
    """One off-axis source point (no centrosymmetric partner to cancel the sign)."""
    n, dx = 32, 32e-9
    g = O.Grid(n, dx)
    mask = np.random.default_rng(4).random((n, n))
    # an even aberration: with an odd one and a real mask the two signs image alike
    o = O.Optics(NA=0.93, zernike_coeffs={4: 0.05}, normalisation="none")
    one = O.SourcePts(weight=np.array([1.0]), fx=np.array([3 / (n * dx)]),
                      fy=np.array([-2 / (n * dx)]), phi=np.array([0.0]), grid=3)
    good = O.abbe_sum(mask, o, g, [(0.0, 1.0)], src=one)[0][0]
    real = O.pupil_geometry
    monkeypatch.setattr(O, "pupil_geometry",
                        lambda FX, FY, lam, NA, fx=0.0, fy=0.0, need_phi=False:
                        real(FX, FY, lam, NA, -fx, -fy, need_phi))
    bad = O.abbe_sum(mask, o, g, [(0.0, 1.0)], src=one)[0][0]
    assert np.abs(bad - good).max() > 1e-3 * good.max()


def test_a_dropped_obliquity_fails_the_vector_golden(monkeypatch):
    z = _gold_vector()
    o = _optics(json.loads(str(z["te_obl1__optics"])))
    real = O.vector_coefficients
    monkeypatch.setattr(O, "vector_coefficients",
                        lambda rho, phi, jx, jy, stm, obliquity=True:
                        real(rho, phi, jx, jy, stm, obliquity=False))
    raw = O.compute_aerial_image(z["mask"], o, O.Grid(64, 4e-9), return_raw=True)[0]
    assert np.abs(raw - z["te_obl1__raw"]).max() > 1e-4 * raw.max()


def test_transposed_source_rows_fail_the_golden_table():
    z = np.load(os.path.join(GOLD, "sources.npz"))
    o = _optics(json.loads(str(z["dipole_y_0.9_0.6_g11__optics"])))
    s = O.source_points(o)
    assert not np.array_equal(s.fy, z["dipole_y_0.9_0.6_g11__fx"])
    assert np.array_equal(s.fy, z["dipole_y_0.9_0.6_g11__fy"])


def test_lines_and_spaces_period_and_duty():
    m = O.lines_and_spaces(200, 4e-9, 200e-9, 100e-9)
    row = m[0]
    # 200 nm pitch = 50 px (edge pixels may flip with float rounding of mod)
    assert np.mean(row[:150] == row[50:200]) > 0.97
    assert abs(row.mean() - 0.5) < 0.02                 # 100 nm CD
    assert np.all(m == row[None, :])                    # vertical lines


def test_effective_defocus_closed_forms():
    """resist3d.py:80-110: defocus - (depth - d_ref) n_imm / n_resist, d_ref by reference plane."""
    from lithobat_amd.config import OpticsConfig, ResistConfig
    o = OpticsConfig(NA=1.35, n_immersion=1.44, defocus=30e-9)
    for ref, d_ref in (("top", 0.0), ("mid", 50e-9), ("bottom", 100e-9)):
        r = ResistConfig(thickness=100e-9, n_resist=1.7, focus_reference=ref)
        assert O.focus_reference_depth(r) == d_ref
        for depth in (0.0, 20e-9, 100e-9):
            want = 30e-9 - (depth - d_ref) * 1.44 / 1.7
            assert O.effective_defocus(depth, o, r) == pytest.approx(want, rel=1e-15, abs=1e-24)
    with pytest.raises(ValueError, match="focus_reference"):
        ResistConfig(focus_reference="centre")


def test_exposure_volume_interpolates_between_the_right_planes():
    """Each voxel lies between the two film planes around its depth (bottom-up order), and the
    linear interpolation of the smooth through-focus image stays close to the image evaluated
    directly at the voxel's own depth."""
    from lithobat_amd.config import GridConfig, OpticsConfig, ResistConfig
    grid = GridConfig(n_pixels=32, pixel_size=8e-9, dz=4e-9, n_z_slices=6)
    optics = OpticsConfig(NA=0.93, source_type="annular", sigma_outer=0.8, sigma_inner=0.5,
                          source_grid=7, defocus=20e-9, normalisation="none")
    resist = ResistConfig(thickness=100e-9, n_resist=1.7, focus_reference="top")
    mask = O.lines_and_spaces(32, 8e-9, 128e-9, 64e-9)
    vol, z = O.exposure_volume(mask, optics, grid, resist)
    nz = 25
    assert vol.shape == (nz, 32, 32) and np.allclose(z, (np.arange(nz) + 0.5) * 4e-9)
    ratio = 1.7 / optics.n_immersion
    for iz in (0, 7, nz - 1):  # iz = 0 is the bottom of the film
        depth = 100e-9 - (iz + 0.5) * 4e-9
        direct = O.compute_aerial_planes(
            mask, optics, grid, [(O.effective_defocus(depth, optics, resist) * ratio, 1.7)])[0]
        assert np.abs(vol[iz] - direct).max() <= 2e-3 * np.abs(direct).max(), iz
    # planted mistake: a top-down (unflipped) volume misses the direct image badly
    flipped = vol[::-1]
    depth0 = 100e-9 - 0.5 * 4e-9
    direct0 = O.compute_aerial_planes(
        mask, optics, grid, [(O.effective_defocus(depth0, optics, resist) * ratio, 1.7)])[0]
    assert np.abs(flipped[0] - direct0).max() > 5 * np.abs(vol[0] - direct0).max()


@pytest.mark.parametrize("sigma", [0.6, 2.3, 7.0])
def test_blur_is_the_direct_gaussian_with_reflect_edges(sigma):
    """The oracle's blur (the reference's scipy call) equals an explicit restatement."""
    rng = np.random.default_rng(int(sigma * 10))
    a = rng.standard_normal((24, 40))
    assert np.abs(O.blur(a, sigma) - O.blur_direct(a, sigma)).max() <= 1e-14
    # a planted mistake: periodic instead of reflected borders differs at the edges
    import scipy.ndimage as nd
    wrong = nd.gaussian_filter(a, sigma, mode="wrap")
    assert np.abs(wrong - O.blur(a, sigma)).max() > 1e-3
    assert O.blur(a, 5e-4) is a


def test_blur_is_self_adjoint_and_objective_gradient_matches_fd():
    """<B x, y> = <x, B y> (symmetric taps + symmetric extension), so the chain rule of
    ilt.py:336 is exact; the relaxed objective with a blur against central differences."""
    rng = np.random.default_rng(5)
    x, y = rng.standard_normal((2, 20, 20))
    assert abs(np.sum(O.blur(x, 1.7) * y) - np.sum(x * O.blur(y, 1.7))) <= 1e-12
    raw = rng.random((16, 16)) * 1.5
    target = (rng.random((16, 16)) > 0.5).astype(float)
    cost, d_raw, _, _ = O.relaxed_objective(raw, target, 0.8, 0.3, 5.0, sigma_px=1.3)
    h = 1e-6
    for i, j in [(0, 0), (3, 7), (15, 2)]:
        rp, rm = raw.copy(), raw.copy()
        rp[i, j] += h
        rm[i, j] -= h
        fd = (O.relaxed_objective(rp, target, 0.8, 0.3, 5.0, sigma_px=1.3)[0]
              - O.relaxed_objective(rm, target, 0.8, 0.3, 5.0, sigma_px=1.3)[0]) / (2 * h)
        assert d_raw[i, j] == pytest.approx(fd, rel=1e-5, abs=1e-10)
    m = rng.random((12, 12))
    c, g = O.mask_penalties(m, 0.3, 0.2)
    for i, j in [(1, 1), (11, 5)]:
        mp, mm = m.copy(), m.copy()
        mp[i, j] += h
        mm[i, j] -= h
        fd = (O.mask_penalties(mp, 0.3, 0.2)[0] - O.mask_penalties(mm, 0.3, 0.2)[0]) / (2 * h)
        assert g[i, j] == pytest.approx(fd, rel=1e-6, abs=1e-12)


def test_apply_peb_and_develop_field():
    """peb.py:24-70: blur then clip to [0, 1]; no blur below 1e-3 px; threshold model."""
    from lithobat_amd.config import GridConfig
    rng = np.random.default_rng(3)
    a = rng.random((32, 32)) * 1.4
    out = O.apply_peb(a, 10e-9, 4e-9)
    assert out.min() >= 0.0 and out.max() <= 1.0
    assert np.allclose(out, np.clip(O.blur(a, 2.5), 0.0, 1.0), atol=0, rtol=0)
    same = O.apply_peb(a, 1e-15, 4e-9)
    assert same is not a and np.array_equal(same, a)

    class R:
        diffusion_sigma, threshold = 8e-9, 0.3
    fld, lvl = O.develop_field(a, R(), GridConfig(n_pixels=32, pixel_size=4e-9))
    assert lvl == 0.3 and np.array_equal(fld, O.apply_peb(a, 8e-9, 4e-9))


def test_run_ilt_descends_and_stops():
    """The restated solve lowers the cost, anneals beta geometrically to beta_final, and its
    first step equals a hand-computed Adam step (the bias-corrected first step is -step *
    sign(g) wherever g != 0)."""
    from lithobat_amd.config import GridConfig, OpticsConfig
    grid = GridConfig(n_pixels=32, pixel_size=8e-9)
    optics = OpticsConfig(NA=1.35, n_immersion=1.44, source_type="annular", sigma_outer=0.9,
                          sigma_inner=0.6, source_grid=9, normalisation="clear")

    class R:
        threshold, diffusion_sigma = 0.3, 8e-9
    target = O.lines_and_spaces(32, 8e-9, 128e-9, 64e-9)
    mask, theta, rows, taken, conv = O.run_ilt(target, optics, grid, R(), max_iter=6,
                                               smooth=8e-9, weight_tv=1e-3)
    assert taken == 6 and len(rows) == 7
    assert rows[-1]["cost"] < rows[0]["cost"]
    # beta_k = beta growth^k, growth = (beta_final / beta)^(1 / (max_iter - 1)): beta_final is
    # reached at iteration max_iter - 1 and the last step goes one factor past it (ilt.py:538-541)
    assert rows[5]["beta"] == pytest.approx(12.0, rel=1e-12)
    assert rows[6]["beta"] == pytest.approx(12.0 * 3.0 ** 0.2, rel=1e-12)
    # hand-computed first Adam step from the initial state
    _, theta1, _, _, _ = O.run_ilt(target, optics, grid, R(), max_iter=1, beta_final=None,
                                   smooth=8e-9, weight_tv=1e-3)
    p = 0.9
    m0 = 0.5 + (target - 0.5) * (2 * p - 1)
    th0 = np.log(m0 / (1 - m0)) / 4.0
    # bias-corrected first Adam step: |d theta| = step |g| / (|g| + eps), just under step
    # (without the bias correction it would be step * 0.1 / sqrt(0.001) = 1.26 step)
    d = np.abs(theta1 - th0)
    assert np.all(d <= 0.4) and np.all(d >= 0.399)


@pytest.mark.parametrize("tone", ["clear", "dark"])
def test_dose_for_target_prints_the_target_area(tone):
    """ilt.py:651-721 restated: at the returned dose the printed-pixel count crosses the
    target's count (just below the dose on one side, just above on the other, in the
    direction the tone makes the area move); empty / full targets keep the given dose."""
    from lithobat_amd.config import GridConfig, OpticsConfig
    grid = GridConfig(n_pixels=32, pixel_size=8e-9)
    optics = OpticsConfig(NA=1.35, n_immersion=1.44, source_type="annular", sigma_outer=0.9,
                          sigma_inner=0.6, source_grid=9, normalisation="clear")

    class R:
        threshold, diffusion_sigma = 0.3, 8e-9
    target = O.lines_and_spaces(32, 8e-9, 128e-9, 64e-9)
    d = O.dose_for_target(target, optics, grid, R(), tone=tone)
    assert 1e-3 < d < 1e3
    k = O.socs_build(optics, grid)
    raw, _, clear = O.socs_raw(k, target)
    unit = O.blur(raw / clear, 1.0)
    sign = 1.0 if tone == "clear" else -1.0

    def area(x):
        return int(np.count_nonzero(sign * (x * unit - 0.3) > 0.0))
    want = int(np.count_nonzero(target > 0.5))
    a_lo, a_hi = area(d * (1 - 1e-6)), area(d * (1 + 1e-6))
    if tone == "clear":  # more dose prints more
        assert a_lo <= want <= a_hi + 1 and a_hi >= a_lo
    else:                # more dose clears more of a dark-tone resist
        assert a_hi <= want <= a_lo + 1 and a_lo >= a_hi
    assert O.dose_for_target(np.zeros((32, 32)), optics, grid, R(), dose=1.7) == 1.7


def _ml_optics(z, case):
    from lithobat_amd.config import OpticsConfig
    d = json.loads(str(z[f"{case}__optics"]))
    d["zernike_coeffs"] = {int(k): v for k, v in d["zernike_coeffs"].items()}
    return OpticsConfig(**d)


def test_multilayer_oracle_matches_golden_images():
    """Reflective-mask Abbe sums restated in the oracle vs the reference's images (scalar
    with defocus, a 30-bilayer blank with the chief ray on y, vector TE with Zernike), their
    clear fields and angle tables; planes of one call vs the reference's plane stack."""
    from lithobat_amd.config import GridConfig
    z = np.load(os.path.join(GOLD, "multilayer64.npz"))
    grid = GridConfig(n_pixels=64, pixel_size=4e-9)
    for c in z["cases"]:
        o = _ml_optics(z, c)
        th, mag, ph = O.reflection_table(o)
        assert np.array_equal(th, z[f"{c}__theta"])
        assert np.abs(mag - z[f"{c}__mag"]).max() <= 1e-13
        assert np.abs(ph - z[f"{c}__phase"]).max() <= 1e-12
        raw, pk, cl = O.compute_aerial_image(z["mask"], o, grid, return_raw=True)
        assert np.abs(raw - z[f"{c}__raw"]).max() <= 1e-12 * z[f"{c}__raw"].max()
        assert cl == pytest.approx(float(z[f"{c}__clear"]), rel=1e-12)
    o = _ml_optics(z, "scalar")
    st = O.compute_aerial_planes(z["mask"], o, grid, [tuple(p) for p in z["planes"]])
    assert np.abs(st - z["planes__raw"]).max() <= 1e-12 * z["planes__raw"].max()


def test_film_reflection_closed_forms():
    """The characteristic-matrix reflection: a zero-thickness layer leaves the Fresnel
    coefficients of the ambient / substrate interface (both polarisations, oblique); a
    lossless quarter-wave layer of index sqrt(n_sub) at normal incidence cancels the
    reflection; the 40-bilayer Mo/Si blank reflects 74.7 % at 6 deg (the materials table's
    physical check) as the reference computes it."""
    n0, n2, th = 1.0, 1.5, 0.4
    films = [(n0, 0.0), (1.3, 0.0), (n2, 0.0)]
    c0 = np.cos(th)
    c2 = np.sqrt(1 - (n0 * np.sin(th) / n2) ** 2)
    rs = (n0 * c0 - n2 * c2) / (n0 * c0 + n2 * c2)
    rp = (n0 / c0 - n2 / c2) / (n0 / c0 + n2 / c2)
    assert O.film_reflection(films, 193e-9, th, "s") == pytest.approx(rs, abs=1e-15)
    assert O.film_reflection(films, 193e-9, th, "p") == pytest.approx(rp, abs=1e-15)
    lam = 193e-9
    n1 = np.sqrt(n2)
    qw = [(n0, 0.0), (n1, lam / (4 * n1)), (n2, 0.0)]
    assert abs(O.film_reflection(qw, lam, 0.0, "s")) <= 1e-15
    z = np.load(os.path.join(GOLD, "multilayer64.npz"))
    from lithobat_amd.config import OpticsConfig
    o = OpticsConfig(wavelength=13.5e-9, NA=0.33, mask_model="multilayer", chief_ray_deg=6.0)
    films = O.blank_films(o)
    r_s = O.film_reflection(films, 13.5e-9, np.deg2rad(6.0), "s")
    assert abs(r_s) ** 2 == pytest.approx(0.747, abs=5e-4)
    assert r_s == pytest.approx(complex(z["blank_r_s_6deg"]), abs=1e-13)
    assert O.film_reflection(films, 13.5e-9, np.deg2rad(6.0), "p") == pytest.approx(
        complex(z["blank_r_p_6deg"]), abs=1e-13)


def test_host_reflection_table_matches_oracle():
    """lithobat_amd.m3d (the plan's host-side blank solve, vectorised over angles) against
    the oracle's per-angle restatement and the reference's tables; the reference's errors."""
    from lithobat_amd import m3d
    z = np.load(os.path.join(GOLD, "multilayer64.npz"))
    for c in z["cases"]:
        o = _ml_optics(z, c)
        th, mag, ph = m3d.reflection_table(m3d.stack_for(o), o.wavelength)
        assert np.array_equal(th, z[f"{c}__theta"])
        assert np.abs(mag - z[f"{c}__mag"]).max() <= 1e-13
        assert np.abs(ph - z[f"{c}__phase"]).max() <= 1e-12
        for fx, fy in ((0.0, 0.0), (3.1e6, -1.7e6)):
            assert m3d.mask_side_sin_theta(fx, fy, o) == pytest.approx(
                O.mask_side_sines(fx, fy, o), abs=1e-17)
    with pytest.raises(KeyError, match="Unknown mask material"):
        m3d.mask_material("Au", 13.5e-9)
    with pytest.raises(KeyError, match="No mask-material table"):
        m3d.mask_material("Si", 248e-9)
    with pytest.raises(ValueError, match="needs a multilayer"):
        m3d.MaskStack(multilayer=None)


def test_adam_step_closed_forms():
    """First step from zero moments: the bias-corrected moments are g and g^2, so theta
    moves by -step g / (|g| + eps); a constant gradient keeps m_hat = g, v_hat = g^2 at
    every later step too (the bias corrections cancel the geometric sums exactly)."""
    rng = np.random.default_rng(2)
    g = rng.standard_normal(50)
    th0 = rng.standard_normal(50)
    th, m1, m2 = O.adam_step(th0, np.zeros(50), np.zeros(50), g, 1, 0.4)
    assert np.allclose(th, th0 - 0.4 * g / (np.abs(g) + 1e-8), rtol=0, atol=1e-15)
    assert np.allclose(m1, 0.1 * g, rtol=1e-15) and np.allclose(m2, 0.001 * g ** 2, rtol=1e-13)
    th, a, b = th0, np.zeros(50), np.zeros(50)
    for it in range(1, 8):
        prev = th
        th, a, b = O.adam_step(th, a, b, g, it, 0.05)
        assert np.allclose(th, prev - 0.05 * g / (np.abs(g) + 1e-8), rtol=0, atol=1e-12)


# ------------------------------------------------ ILT family vs the reference's outputs

def _ilt_gold():
    from lithobat_amd.config import GridConfig, OpticsConfig, ResistConfig
    z = np.load(os.path.join(GOLD, "ilt32.npz"))
    d = json.loads(str(z["optics"]))
    d["zernike_coeffs"] = {int(k): v for k, v in d["zernike_coeffs"].items()}
    optics = OpticsConfig(**d)
    grid = GridConfig(n_pixels=int(z["n"]), pixel_size=float(z["pixel_size"]))
    resist = ResistConfig(threshold=float(z["resist_threshold"]),
                          diffusion_sigma=float(z["resist_diffusion_sigma"]))
    return z, optics, grid, resist


def _relmax(a, b):
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def test_objective_matches_the_reference_objective():
    """ilt.py:292-360 (_Objective.__call__) on a sigmoid(8 theta) mask: cost, gradient,
    signed field and pattern error for sign +1 and -1, with and without the PEB blur and
    the TV / discreteness penalties, equal the reference's outputs."""
    z, optics, grid, _ = _ilt_gold()
    k = O.socs_build(optics, grid, tol=float(z["obj_socs_tol"]))
    assert np.abs(k.weights - z["obj_weights"]).max() <= 1e-12 * z["obj_weights"].max()
    assert k.clear == pytest.approx(float(z["obj_clear"]), rel=1e-12)
    scale = O.normalisation_scale("clear", 1.0, k.clear, 1.0)
    assert scale == pytest.approx(float(z["obj_scale"]), rel=1e-14)
    mask, target = z["obj_mask"], z["obj_target"]
    for c in z["obj_cases"]:
        p = json.loads(str(z[f"obj_{c}__params"]))
        cost, grad, signed, perr = O.objective(k, mask, target, scale, p["level"],
                                               p["steepness"], p["sign"], p["sigma_px"],
                                               p["wtv"], p["wd"])
        assert cost == pytest.approx(float(z[f"obj_{c}__cost"]), rel=1e-12), c
        assert _relmax(grad, z[f"obj_{c}__grad"]) <= 1e-10, c
        assert _relmax(signed, z[f"obj_{c}__signed"]) <= 1e-12, c
        assert perr == int(z[f"obj_{c}__pattern_error"]), c


def test_objective_pins_catch_a_wrong_logistic_slope():
    """Planted mistakes: the slope a = steepness (not steepness / |level|), or the
    relaxed cost summed instead of averaged, miss the reference's cost."""
    z, optics, grid, _ = _ilt_gold()
    k = O.socs_build(optics, grid, tol=float(z["obj_socs_tol"]))
    mask, target = z["obj_mask"], z["obj_target"]
    p = json.loads(str(z["obj_plain__params"]))
    want = float(z["obj_plain__cost"])
    wrong_a = O.objective(k, mask, target, 1.0, p["level"], p["steepness"] * abs(p["level"]),
                          p["sign"])[0]
    assert abs(wrong_a - want) > 1e-3 * want
    right = O.objective(k, mask, target, 1.0, p["level"], p["steepness"], p["sign"])[0]
    assert abs(right * mask.size - want) > 1.0


def test_socs_gradient_matches_the_reference_on_plain_fields():
    """opc/ilt.py:208-266: socs_fields' intensity and socs_gradient of a plain sensitivity
    on the fields socs_fields returns, against the reference's outputs."""
    z, optics, grid, _ = _ilt_gold()
    k = O.socs_build(optics, grid, tol=float(z["obj_socs_tol"]))
    fields, inten = O.socs_fields(k, z["obj_mask"])
    assert _relmax(inten, z["sg_intensity"]) <= 1e-12
    g = O.socs_gradient(k, fields, z["sg_sensitivity"])
    assert _relmax(g, z["sg_grad"]) <= 1e-10


def test_mask_parameterisation_matches_the_reference():
    """ilt.py:368-389: _mask_of and _grad_theta with and without the smoothing Gaussian."""
    z, _, _, _ = _ilt_gold()
    for name in ("b8_s0", "b4_s1.3", "b12_s2.5"):
        beta, spx = float(z[f"par_{name}__beta"]), float(z[f"par_{name}__smooth_px"])
        m = O.mask_of(z["obj_theta"], beta, spx)
        assert np.abs(m - z[f"par_{name}__mask"]).max() <= 1e-14, name
        gt = O.grad_theta(z["par_grad_mask"], m, beta, spx)
        assert _relmax(gt, z[f"par_{name}__grad_theta"]) <= 1e-13, name


@pytest.mark.parametrize("tone", ["clear", "dark"])
def test_dose_for_target_matches_the_reference(tone):
    """ilt.py:651-721 for both tones, on the target itself and on another mask."""
    z, optics, grid, resist = _ilt_gold()
    tgt = z["ilt_target"]
    d = O.dose_for_target(tgt, optics, grid, resist, tone=tone)
    assert d == pytest.approx(float(z[f"dose_{tone}"]), rel=1e-12)
    d2 = O.dose_for_target(tgt, optics, grid, resist, tone=tone, mask=z["dose_other_mask"])
    assert d2 == pytest.approx(float(z[f"dose_{tone}__other"]), rel=1e-12)


@pytest.mark.parametrize("name", ["clear", "dark"])
def test_run_ilt_matches_the_reference_solve(name):
    """ilt.py:397-621: five Adam iterations (clear tone with smoothing, PEB, TV and
    discreteness penalties and the beta anneal; dark tone with beta held) give the
    reference's history (cost, pattern error, beta, gradient norm), mask, theta,
    iteration count and final print (ilt.py:624-648)."""
    z, optics, grid, resist = _ilt_gold()
    p = json.loads(str(z[f"run_{name}__params"]))
    tgt = z[f"run_{name}__target"]
    dose = O.dose_for_target(tgt, optics, grid, resist, tone=p["tone"])
    assert dose == pytest.approx(float(z[f"run_{name}__dose"]), rel=1e-12)
    kw = {k: v for k, v in p.items() if k != "tone"}
    mask, theta, rows, taken, conv = O.run_ilt(tgt, optics, grid, resist, dose=dose,
                                               tone=p["tone"], **kw)
    assert taken == int(z[f"run_{name}__iterations"]) and conv == bool(z[f"run_{name}__converged"])
    cost = np.array([r["cost"] for r in rows])
    assert _relmax(cost, z[f"run_{name}__cost"]) <= 1e-10
    assert [r["pattern_error"] for r in rows] == list(z[f"run_{name}__pattern_error"])
    assert np.allclose([r["beta"] for r in rows], z[f"run_{name}__beta"], rtol=1e-14, atol=0)
    assert np.allclose([r["grad_norm"] for r in rows], z[f"run_{name}__grad_norm"], rtol=1e-9,
                       atol=0)
    assert np.abs(mask - z[f"run_{name}__mask"]).max() <= 1e-10
    assert np.abs(theta - z[f"run_{name}__theta"]).max() <= 1e-9
    aerial, fld, level = O.print_mask(mask, optics, grid, resist, dose=dose)
    assert _relmax(aerial, z[f"run_{name}__printed_aerial"]) <= 1e-11
    assert _relmax(fld, z[f"run_{name}__printed_field"]) <= 1e-11 and level == resist.threshold


def test_exposure_volume_matches_the_reference():
    """develop/resist3d.py:113-185 for the top, mid and bottom focus references, without and
    with clear normalisation (NA 1.35 immersion into the resist index): the voxel intensities
    and heights equal the reference's."""
    from lithobat_amd.config import GridConfig, OpticsConfig, ResistConfig
    z = np.load(os.path.join(GOLD, "ilt32.npz"))
    vm = z["vol_mask"]
    grid = GridConfig(n_pixels=vm.shape[0], pixel_size=float(z["pixel_size"]),
                      dz=float(z["vol_dz"]), n_z_slices=int(z["vol_n_z_slices"]))
    for c in z["vol_cases"]:
        _, oname, ref = str(c).split("_")
        d = json.loads(str(z[f"vol_optics_{oname}"]))
        d["zernike_coeffs"] = {int(k): v for k, v in d["zernike_coeffs"].items()}
        o = OpticsConfig(**d)
        r = ResistConfig(thickness=float(z["vol_thickness"]), n_resist=float(z["vol_n_resist"]),
                         focus_reference=ref)
        vol, zz = O.exposure_volume(vm, o, grid, r, dose=float(z["vol_dose"]))
        assert np.array_equal(zz, z[f"{c}__z"]), c
        assert _relmax(vol, z[f"{c}__intensity"]) <= 1e-12, c


def test_exposure_volume_pin_catches_a_wrong_depth_conversion():
    """Planted mistake: the defocus handed to the planes without the n_resist / n_immersion
    conversion (the index applied twice) misses the reference's volume."""
    from lithobat_amd.config import GridConfig, OpticsConfig, ResistConfig
    z = np.load(os.path.join(GOLD, "ilt32.npz"))
    vm = z["vol_mask"]
    grid = GridConfig(n_pixels=vm.shape[0], pixel_size=float(z["pixel_size"]),
                      dz=float(z["vol_dz"]), n_z_slices=int(z["vol_n_z_slices"]))
    d = json.loads(str(z["vol_optics_clear"]))
    d["zernike_coeffs"] = {}
    o = OpticsConfig(**d)
    r = ResistConfig(thickness=float(z["vol_thickness"]), n_resist=float(z["vol_n_resist"]),
                     focus_reference="top")
    nz, ns = 22, int(z["vol_n_z_slices"])
    depths = np.linspace(0.0, r.thickness, ns)
    planes = O.compute_aerial_planes(
        vm, o, grid, [(O.effective_defocus(float(t), o, r), r.n_resist) for t in depths],
        dose=float(z["vol_dose"]))
    zv = ((np.arange(nz) + 0.5) * grid.dz)[::-1]
    flat = planes.reshape(ns, -1)
    wrong = np.stack([np.interp(zv, depths, flat[:, q]) for q in range(flat.shape[1])], 1)
    want = z["vol_clear_top__intensity"].reshape(nz, -1)
    assert _relmax(wrong, want) > 1e-4
