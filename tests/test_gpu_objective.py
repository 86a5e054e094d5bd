"""Fused ILT objective kernels (lb_resist_objective, lb_mask_sigmoid*) vs the oracle.

FP64: <= 1e-12 relative (cost, derivative, signed field, mask, mask adjoint);
mismatch counts exact.  FP32: <= 1e-5 relative to the largest magnitude.
"""
import numpy as np
import pytest

from oracle import litho_oracle as O

pytestmark = pytest.mark.gpu


def _rel(a, b):
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def _fields(n_img, n, seed):
    rng = np.random.default_rng(seed)
    raw = 0.3 + 0.25 * rng.standard_normal((n_img, n, n))  # straddles the 0.3 level
    target = (rng.random((n_img, n, n)) > 0.5).astype(np.float64)
    return raw, target


@pytest.mark.parametrize("precision", [64, 32])
@pytest.mark.parametrize("sign", [1.0, -1.0])
def test_resist_objective_matches_oracle(device, precision, sign):
    import torch
    from lithobat_amd.objective import resist_objective
    dt = torch.float64 if precision == 64 else torch.float32
    tol = 1e-12 if precision == 64 else 1e-5
    raw, target = _fields(3, 96, 7)
    scale, level, steep = 1.3, 0.3, 50.0
    r = resist_objective(torch.tensor(raw, device=device, dtype=dt),
                         torch.tensor(target, device=device, dtype=dt), scale, level, steep, sign,
                         signed=True, mismatch=True)
    tot = 0.0
    for b in range(3):
        cost, d_raw, signed, pe = O.relaxed_objective(raw[b], target[b], scale, level, steep, sign)
        assert float(r["cost"][b]) == pytest.approx(cost, rel=tol)
        assert _rel(r["d_intensity"][b].cpu(), d_raw) <= tol
        assert _rel(r["signed"][b].cpu(), signed) <= tol
        if precision == 64:
            assert int(r["mismatch"][b]) == pe
        else:  # pixels within FP32 rounding of the level may flip
            assert abs(int(r["mismatch"][b]) - pe) <= 2
        tot += cost
    assert float(r["total"]) == pytest.approx(tot, rel=tol)


def test_resist_objective_saturated_and_broadcast_target(device):
    """Fields far beyond +-60 / a saturate the logistic (z = 0 or 1 exactly as the
    reference's clip gives); a single target broadcasts over the leading axes."""
    import torch
    from lithobat_amd.objective import resist_objective
    raw = np.concatenate([np.full((1, 16, 16), 50.0), np.full((1, 16, 16), -50.0)])
    t = np.zeros((16, 16))
    t[:8] = 1.0
    r = resist_objective(torch.tensor(raw, device=device), torch.tensor(t, device=device), 1.0,
                         0.3, 50.0)
    for b in range(2):
        cost, d_raw, _, _ = O.relaxed_objective(raw[b], t, 1.0, 0.3, 50.0)
        assert float(r["cost"][b]) == pytest.approx(cost, rel=1e-12)
        assert _rel(r["d_intensity"][b].cpu(), d_raw) <= 1e-12 or np.abs(d_raw).max() < 1e-20


def test_resist_objective_deterministic(device):
    import torch
    from lithobat_amd.objective import resist_objective
    raw, target = _fields(4, 256, 11)
    f = torch.tensor(raw, device=device, dtype=torch.float32)
    t = torch.tensor(target, device=device, dtype=torch.float32)
    a = resist_objective(f, t, 1.0, 0.3, 50.0)
    b = resist_objective(f, t, 1.0, 0.3, 50.0)
    assert torch.equal(a["cost"], b["cost"]) and torch.equal(a["total"], b["total"])


@pytest.mark.parametrize("precision", [64, 32])
def test_mask_sigmoid_and_adjoint_match_oracle(device, precision):
    import torch
    from lithobat_amd.objective import mask_sigmoid, mask_sigmoid_grad
    dt = torch.float64 if precision == 64 else torch.float32
    tol = 1e-12 if precision == 64 else 1e-6
    rng = np.random.default_rng(3)
    theta = rng.standard_normal((2, 64, 64)) * 3.0
    theta[0, 0, :4] = [20.0, -20.0, 0.0, 7.5]  # clipped at beta * theta = +-60 for beta 8
    g = rng.standard_normal((2, 64, 64))
    m = mask_sigmoid(torch.tensor(theta, device=device, dtype=dt), 8.0)
    ref_m = O.mask_of(theta, 8.0)
    assert _rel(m.cpu(), ref_m) <= tol
    gx = mask_sigmoid_grad(torch.tensor(g, device=device, dtype=dt), m, 8.0)
    assert _rel(gx.cpu(), O.grad_theta(g, m.double().cpu().numpy(), 8.0)) <= tol


def test_autograd_objective_chain_matches_oracle(device):
    """mask_of -> (identity intensity) -> relaxed_resist_loss, backward through both fused
    functions, against the oracle's closed-form chain in float64."""
    import torch
    from lithobat_amd.ilt import mask_of, relaxed_resist_loss
    rng = np.random.default_rng(5)
    theta = rng.standard_normal((2, 48, 48))
    target = (rng.random((2, 48, 48)) > 0.5).astype(np.float64)
    th = torch.tensor(theta, device=device, requires_grad=True)
    m = mask_of(th, 8.0)
    loss = relaxed_resist_loss(m * 0.6, torch.tensor(target, device=device), 1.5, 0.3, 50.0)
    loss.backward()
    cost = 0.0
    for b in range(2):
        mb = O.mask_of(theta[b], 8.0)
        c, d_raw, _, _ = O.relaxed_objective(mb * 0.6, target[b], 1.5, 0.3, 50.0)
        cost += c
        ref = O.grad_theta(d_raw * 0.6, mb, 8.0)
        assert _rel(th.grad[b].cpu(), ref) <= 1e-12
    assert float(loss) == pytest.approx(cost, rel=1e-12)


def test_blurred_objective_matches_oracle(device):
    import torch
    from lithobat_amd.ilt import relaxed_resist_loss
    raw, target = _fields(1, 64, 9)
    r = torch.tensor(raw, device=device, requires_grad=True)
    loss = relaxed_resist_loss(r, torch.tensor(target, device=device), 1.2, 0.3, 50.0,
                               sign=-1.0, sigma_px=2.0)
    loss.backward()
    cost, d_raw, _, _ = O.relaxed_objective(raw[0], target[0], 1.2, 0.3, 50.0, -1.0, 2.0)
    assert float(loss) == pytest.approx(cost, rel=1e-12)
    assert _rel(r.grad[0].cpu(), d_raw) <= 1e-12


def test_odd_sizes_unaligned_views_and_many_images(device):
    """Vector tails (37^2 pixels), 8-byte-offset views (the scalar path) and more images
    than the final reduction's wavefronts (20)."""
    import torch
    from lithobat_amd.objective import mask_sigmoid, mask_sigmoid_grad, resist_objective
    rng = np.random.default_rng(8)
    n, b = 37, 20
    base = rng.standard_normal(b * n * n + 1)
    x = torch.tensor(base, device=device)[1:].reshape(b, n, n)  # 8-byte offset: not 16-aligned
    m = mask_sigmoid(x, 8.0)
    ref_m = O.mask_of(base[1:].reshape(b, n, n), 8.0)
    assert _rel(m.cpu(), ref_m) <= 1e-12
    g = rng.standard_normal((b, n, n))
    out = torch.empty(b * n * n + 1, device=device, dtype=torch.float64)[1:].reshape(b, n, n)
    gx = mask_sigmoid_grad(torch.tensor(g, device=device), m, 8.0, out=out)
    assert gx.data_ptr() == out.data_ptr()
    assert _rel(out.cpu(), O.grad_theta(g, ref_m, 8.0)) <= 1e-12
    raw, target = _fields(b, n, 12)
    r = resist_objective(torch.tensor(raw, device=device), torch.tensor(target, device=device),
                         0.9, 0.3, 50.0, mismatch=True)
    tot = 0.0
    for i in range(b):
        cost, d_raw, _, pe = O.relaxed_objective(raw[i], target[i], 0.9, 0.3, 50.0)
        assert float(r["cost"][i]) == pytest.approx(cost, rel=1e-12)
        assert _rel(r["d_intensity"][i].cpu(), d_raw) <= 1e-12
        assert int(r["mismatch"][i]) == pe
        tot += cost
    assert float(r["total"]) == pytest.approx(tot, rel=1e-12)


@pytest.mark.parametrize("precision", [64, 32])
def test_mask_penalties_match_oracle(device, precision):
    """TV + discreteness cost and gradient (periodic neighbours, odd sizes, several images)
    against the oracle's restatement; each term alone and both together."""
    import torch
    from lithobat_amd.objective import mask_penalties
    dt = torch.float64 if precision == 64 else torch.float32
    tol = 1e-12 if precision == 64 else 1e-5
    rng = np.random.default_rng(6)
    m = rng.random((3, 37, 52))
    for wtv, wd in [(0.3, 0.0), (0.0, 0.2), (0.3, 0.2)]:
        r = mask_penalties(torch.tensor(m, device=device, dtype=dt), wtv, wd)
        tot = 0.0
        for b in range(3):
            c, g = O.mask_penalties(m[b], wtv, wd)
            assert float(r["cost"][b]) == pytest.approx(c, rel=tol)
            assert _rel(r["grad"][b].cpu(), g) <= tol
            tot += c
        assert float(r["total"]) == pytest.approx(tot, rel=tol)


def test_mask_penalties_autograd(device):
    import torch
    from lithobat_amd.ilt import mask_penalties
    rng = np.random.default_rng(7)
    m = rng.random((64, 64))
    x = torch.tensor(m, device=device, requires_grad=True)
    loss = mask_penalties(x, 0.25, 0.1) * 3.0
    loss.backward()
    c, g = O.mask_penalties(m, 0.25, 0.1)
    assert float(loss) == pytest.approx(3.0 * c, rel=1e-12)
    assert _rel(x.grad.cpu(), 3.0 * g) <= 1e-12


@pytest.mark.parametrize("precision", [64, 32])
def test_adam_step_matches_oracle(device, precision):
    """Three in-place steps from zero moments with a changing gradient, against the oracle."""
    import torch
    from lithobat_amd.objective import adam_step
    dt = torch.float64 if precision == 64 else torch.float32
    tol = 1e-12 if precision == 64 else 1e-5
    rng = np.random.default_rng(9)
    th = rng.standard_normal(1001)
    t = torch.tensor(th, device=device, dtype=dt)
    m1, m2 = torch.zeros_like(t), torch.zeros_like(t)
    a, b = np.zeros(1001), np.zeros(1001)
    for it in range(1, 4):
        g = rng.standard_normal(1001)
        adam_step(t, m1, m2, torch.tensor(g, device=device, dtype=dt), it, 0.4)
        th, a, b = O.adam_step(th, a, b, g, it, 0.4)
        assert _rel(t.cpu(), th) <= tol and _rel(m1.cpu(), a) <= tol and _rel(m2.cpu(), b) <= tol
