"""Uneven-clip batches: one persistent coherent-system launch over clips of several sizes."""
import numpy as np
import pytest

from oracle import litho_oracle as O

pytestmark = pytest.mark.gpu


def _rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.abs(a - b).max() / np.abs(b).max())


def _clips(precision_sizes=(64, 96, 128), px=4e-9):
    from lithobat_amd.config import GridConfig, OpticsConfig
    rng = np.random.default_rng(12)
    zern = {j: float(rng.uniform(-0.02, 0.02)) for j in range(4, 12)}
    clips = []
    for k, n in enumerate(precision_sizes):
        o = OpticsConfig(NA=1.35, n_immersion=1.44, imaging_model="vector", polarisation="te",
                         source_type="quadrupole", sigma_outer=0.85, sigma_inner=0.55,
                         source_grid=9, zernike_coeffs=zern, defocus=(-40e-9, 10e-9, 35e-9)[k],
                         normalisation="clear")
        clips.append((o, GridConfig(n_pixels=n, pixel_size=px)))
    return clips


@pytest.mark.parametrize("precision", [32, 64])
def test_batch_forward_and_backward_equal_per_clip_plans(device, precision):
    import torch
    from lithobat_amd.batch import ImagingBatch
    from lithobat_amd.plan import ImagingPlan
    clips = _clips()
    batch = ImagingBatch(clips, precision=precision)
    dt = torch.float64 if precision == 64 else torch.float32
    rng = np.random.default_rng(1)
    masks = [torch.tensor(rng.random((g.n_pixels, g.n_pixels)), device=device, dtype=dt)
             for _, g in clips]
    sens = [torch.tensor(rng.standard_normal((1, g.n_pixels, g.n_pixels)), device=device, dtype=dt)
            for _, g in clips]
    imgs = batch.forward(masks)
    grads = batch.backward(masks, sens)
    tol_img = 1e-12 if precision == 64 else 1e-5
    tol_grad = 1e-10 if precision == 64 else 2e-5
    for (o, g), m, s, im, gr in zip(clips, masks, sens, imgs, grads):
        own = ImagingPlan(o, g, precision=precision)
        ref_im = own.forward(m)[0][0]
        ref_gr = own.backward(m[None], s[None])[0]
        assert _rel(im[0].cpu().numpy(), ref_im.cpu().numpy()) <= tol_img
        assert _rel(gr.cpu().numpy(), ref_gr.cpu().numpy()) <= tol_grad
        oracle = O.compute_aerial_image(m.double().cpu().numpy(), o, g, return_raw=True)[0]
        assert _rel(im[0].double().cpu().numpy(), oracle) <= tol_img
        og = O.abbe_gradient(m.double().cpu().numpy(), o, g, s[0].double().cpu().numpy())
        assert _rel(gr.double().cpu().numpy(), og) <= tol_grad


def test_graph_captured_ilt_step_replays_match_eager(device):
    import torch
    from lithobat_amd.batch import ILTStep, ImagingBatch
    from lithobat_amd.masks import manhattan
    clips = _clips()
    batch = ImagingBatch(clips, precision=32)
    rng = np.random.default_rng(3000)
    thetas = [torch.tensor(rng.standard_normal((g.n_pixels, g.n_pixels)), device=device,
                           dtype=torch.float32) for _, g in clips]
    targets = [torch.tensor(manhattan(g.n_pixels, g.pixel_size, seed=3001 + k, n_rect=(2, 6)),
                            device=device, dtype=torch.float32) for k, (_, g) in enumerate(clips)]
    step = ILTStep(batch, thetas, targets)
    step.run()
    torch.cuda.synchronize()
    loss0 = float(step.loss)
    grads0 = [g.clone() for g in step.grad]
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        step.run()
    torch.cuda.current_stream().wait_stream(s)
    torch.cuda.synchronize()
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        step.run()
    for rep in range(3):
        graph.replay()
        torch.cuda.synchronize()
        assert float(step.loss) == pytest.approx(loss0, rel=1e-6), rep
        for g, g0 in zip(step.grad, grads0):
            assert _rel(g.cpu().numpy(), g0.cpu().numpy()) <= 1e-6, rep
        # eager work on the same buffers between replays
        x = torch.randn(1000, device=device).sum()
        torch.cuda.synchronize()
        assert torch.isfinite(x)


def test_batch_autograd_vs_central_differences(device):
    import torch
    from lithobat_amd.batch import BatchIntensity, ImagingBatch
    clips = _clips(precision_sizes=(48, 64, 96), px=8e-9)
    batch = ImagingBatch(clips, precision=64)
    rng = np.random.default_rng(9)
    thetas = [torch.tensor(rng.standard_normal((g.n_pixels, g.n_pixels)), device=device,
                           requires_grad=True) for _, g in clips]
    weights = [torch.tensor(rng.standard_normal((1, g.n_pixels, g.n_pixels)), device=device)
               for _, g in clips]

    def loss_of(ths):
        raws = BatchIntensity.apply(batch, *[torch.sigmoid(8 * t) for t in ths])
        return sum((r * w).sum() for r, w in zip(raws, weights))

    loss_of(thetas).backward()
    h = 1e-6
    for c, t in enumerate(thetas):
        idx = rng.choice(t.numel(), 4, replace=False)
        for i in idx:
            with torch.no_grad():
                tp = [x.detach().clone() for x in thetas]
                tm = [x.detach().clone() for x in thetas]
                tp[c].view(-1)[i] += h
                tm[c].view(-1)[i] -= h
                fd = (float(loss_of(tp)) - float(loss_of(tm))) / (2 * h)
            g = float(t.grad.view(-1)[i])
            assert abs(g - fd) <= 1e-5 * max(1.0, abs(fd)), (c, i, g, fd)


def test_batch_refuses_mixed_precision(device):
    from lithobat_amd import _lib
    from lithobat_amd.plan import ImagingPlan
    import ctypes as C
    clips = _clips()
    a = ImagingPlan(clips[0][0], clips[0][1], precision=32, coarse_m=32)
    b = ImagingPlan(clips[1][0], clips[1][1], precision=64, coarse_m=32)
    lib = _lib.load()
    hs = (C.c_void_p * 2)(a._h.value, b._h.value)
    h = C.c_void_p()
    with pytest.raises(ValueError, match="share precision"):
        _lib.check(lib.lb_batch_create(hs, 2, C.byref(h)))


@pytest.mark.parametrize("precision", [64, 32])
def test_grouped_clips_in_place_and_staged(device, precision):
    """Clips of repeated sizes form shape groups whose band passes run as one pass: on
    buffers consecutive in memory (in place) and on separately allocated ones (staged
    through the workspace), forward and adjoint equal per-clip plans."""
    import torch
    from lithobat_amd.batch import ImagingBatch
    from lithobat_amd.config import GridConfig, OpticsConfig
    from lithobat_amd.plan import ImagingPlan
    sizes = (64, 96, 64, 96, 64)
    clips = []
    for k, n in enumerate(sizes):
        o = OpticsConfig(NA=1.35, n_immersion=1.44, source_type="annular", sigma_outer=0.9,
                         sigma_inner=0.6, source_grid=9, defocus=(-30e-9 + 15e-9 * k),
                         normalisation="clear")
        clips.append((o, GridConfig(n_pixels=n, pixel_size=4e-9)))
    batch = ImagingBatch(clips, precision=precision)
    assert sorted(len(g) for g in batch.groups) == [2, 3]
    dt = torch.float64 if precision == 64 else torch.float32
    rng = np.random.default_rng(21)
    loose = [torch.tensor(rng.random((n, n)), device=device, dtype=dt) for n in sizes]
    sens = [torch.tensor(rng.standard_normal((1, n, n)), device=device, dtype=dt) for n in sizes]
    packed = batch.grouped_empty(lambda p: (p.n, p.n), device)
    for a, b in zip(packed, loose):
        a.copy_(b)
    tol_img = 1e-12 if precision == 64 else 1e-5
    tol_grad = 1e-10 if precision == 64 else 2e-5
    refs = []
    for (o, g), m, s in zip(clips, loose, sens):
        own = ImagingPlan(o, g, precision=precision)
        refs.append((own.forward(m)[0][0].cpu().numpy(),
                     own.backward(m[None], s[None])[0].cpu().numpy()))
    for masks in (packed, loose):
        imgs = batch.forward(masks)
        grads = batch.backward(masks, sens)
        for (ri, rg), im, gr in zip(refs, imgs, grads):
            assert _rel(im[0].cpu().numpy(), ri) <= tol_img
            assert _rel(gr.cpu().numpy(), rg) <= tol_grad
        # separately allocated outputs (scattered back from the workspace)
        outs = [torch.empty((1, n, n), device=device, dtype=dt) for n in sizes]
        batch.forward_into(masks, outs)
        douts = [torch.empty((n, n), device=device, dtype=dt) for n in sizes]
        batch.backward_into(masks, sens, douts)
        for (ri, rg), im, gr in zip(refs, outs, douts):
            assert _rel(im[0].cpu().numpy(), ri) <= tol_img
            assert _rel(gr.cpu().numpy(), rg) <= tol_grad


def test_batch_backward_reuses_forward_spectra(device):
    """The adjoint after a forward of the same masks takes their band spectra from the
    forward's workspace: same gradients as the adjoint that analyses the masks."""
    import torch
    from lithobat_amd.batch import ImagingBatch
    clips = _clips()
    batch = ImagingBatch(clips, precision=64)
    rng = np.random.default_rng(31)
    masks = batch.grouped_empty(lambda p: (p.n, p.n), device)
    for m in masks:
        m.copy_(torch.tensor(rng.random(tuple(m.shape)), device=device))
    sens = [torch.tensor(rng.standard_normal((1,) + tuple(m.shape)), device=device)
            for m in masks]
    ref = batch.backward(masks, sens)
    imgs = batch.grouped_empty(lambda p: (len(p.planes), p.n, p.n), device)
    batch.forward_into(masks, imgs)
    got = batch.grouped_empty(lambda p: (p.n, p.n), device)
    batch.backward_into(masks, sens, got, after_forward=True)
    for a, b in zip(got, ref):
        assert torch.equal(a, b)
