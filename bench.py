"""Benchmark: batched FP32 Abbe aerial images at 1024^2 on one MI355X.

Workload (the headline configuration): 64 seeded Manhattan masks x 5 defocus
planes (-100, -50, 0, 50, 100 nm) at 1024 x 1024, 4 nm pixels, lambda 193 nm,
NA 0.93, disc quadrupole (centre 0.7, radius 0.15, 45 deg; source grid 47 ->
148 points), complex64 fields with FP32 accumulation.  One step images the
whole 320-image batch from masks already resident in HBM; the step is a
captured HIP graph replay.

Also reported on the same line:
* the plain single-image form (the same library called once per
  (mask, plane) with a host synchronisation after each image) and the
  batched / plain speed-up;
* the dense per-source-point form (the reference's dataflow on the device:
  one dense 1024^2 inverse transform per source point) and batched / dense;
* ILT forward + backward steps/s at 1024^2 (theta -> sigmoid(8 theta) -> Abbe
  at NA 1.35 quadrupole -> relaxed resist L2 vs a Manhattan target);
* the 16-clip vector TE ILT step (clips of 512/768/1024^2 in one persistent
  launch per pass, graph-captured) against one clip per call;
* Hopkins / SOCS: the FP64 device decomposition (K = 24) and 256 masks at
  512^2 imaged through it, with the truncation error against full Abbe;
* the depth-resolved image through a resist film and the full ILT solve
  (run_ilt) on the device;
* the FP32 image error against the float64 oracle on a sample image;
* the plain FP64 image of the CPU-runnable configuration (256^2 lines and
  spaces, NA 0.85 annular, 488 source points) on the device;
* roofline of the coherent-system kernel (HIP events on its stream; HBM
  traffic from the committed counter record when it matches the shape), and
  of the ILT adjoint kernel of the 16-mask step;
* the oracle (the reference algorithm in float64 NumPy/SciPy) timed on the
  host cores on a bounded sample, as the CPU baseline (the headline batch),
  and for the plain FP64, SOCS, ILT and clip configurations.

    python bench.py [--gpus 1] [--steps 10] [--warmup 3] [--legs batched,plain,...]
"""
from __future__ import annotations

import argparse
import json
import math
import os
import re
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

N = 1024
PIXEL = 4e-9
N_MASKS = 64
MASK_SEEDS = range(1000, 1000 + N_MASKS)
DEFOCUS = [-100e-9, -50e-9, 0.0, 50e-9, 100e-9]
PEAK_FP32_TFLOPS = 157.3
# the FP32 vector peak counts packed (two-wide) FP32 instructions; the kernels here are
# built without them, so one FP32 FMA per lane per cycle is their ceiling
VALU_UNPACKED_TFLOPS = 78.6
PEAK_HBM_GBS = 8000.0


def optics_cfg2():
    from lithobat_amd.config import OpticsConfig
    return OpticsConfig(wavelength=193e-9, NA=0.93, source_type="quadrupole", sigma_outer=0.85,
                        sigma_inner=0.55, source_grid=47, normalisation="none")


def optics_cfg4():
    from lithobat_amd.config import OpticsConfig
    return OpticsConfig(wavelength=193e-9, NA=1.35, n_immersion=1.44, source_type="quadrupole",
                        sigma_outer=0.85, sigma_inner=0.55, source_grid=47,
                        normalisation="clear")


def optics_cfg1():
    """The CPU-runnable plain configuration: NA 0.85 annular 0.5-0.8 on a 41-point source
    grid (488 points, 0.05 sigma spacing), zero defocus, peak normalisation."""
    from lithobat_amd.config import OpticsConfig
    return OpticsConfig(wavelength=193e-9, NA=0.85, source_type="annular", sigma_outer=0.8,
                        sigma_inner=0.5, source_grid=41, defocus=0.0)


def clips_cfg5():
    """16 clips, sizes from {512, 768, 1024} (seed 11), Zernike 4-11 ~ U(+-0.02) waves
    (seed 12), defocus ~ U(+-50 nm) per clip (seed 13); NA 1.35 TE vector, 4 nm."""
    from lithobat_amd.config import GridConfig, OpticsConfig
    sizes = np.random.default_rng(11).choice([512, 768, 1024], 16)
    zr = np.random.default_rng(12)
    zern = {j: float(zr.uniform(-0.02, 0.02)) for j in range(4, 12)}
    dz = np.random.default_rng(13).uniform(-50e-9, 50e-9, 16)
    clips = []
    for n, d in zip(sizes, dz):
        o = OpticsConfig(wavelength=193e-9, NA=1.35, n_immersion=1.44, imaging_model="vector",
                         polarisation="te", zernike_coeffs=zern, defocus=float(d),
                         source_type="quadrupole", sigma_outer=0.85, sigma_inner=0.55,
                         source_grid=47, normalisation="clear")
        clips.append((o, GridConfig(n_pixels=int(n), pixel_size=PIXEL)))
    return clips


def field_flops(M: int, W: int) -> float:
    """Algorithmic flops of one coherent system on the coarse grid.

    (W + M) length-M radix FFTs at 5 M log2 M, the spectrum x transfer product
    (6 W^2) and the weighted |E|^2 accumulation (4 M^2).
    """
    return 5.0 * M * math.log2(M) * (W + M) + 6.0 * W * W + 4.0 * M * M


COUNTERS = os.path.join(ROOT, "profiles", "k_field_fwd_abbe_1024_64x5_counters.md")
COUNTERS_ADJ = os.path.join(ROOT, "profiles", "k_field_adj_ilt_1024_16_counters.md")


def kernel_shape(plan, n_masks, adjoint=False):
    """The coherent-system kernel shape a call of ``n_masks`` masks runs."""
    i = plan.info
    pre = "adj_" if adjoint else ""
    chunk = plan.work_items(n_masks, adjoint)[0]
    return (f"M={i['m']} W={i['box']} threads={i[pre + 'block_threads']} "
            f"splits={i[pre + 'column_splits']} chunk={chunk}")


def recorded_traffic(plan, path=COUNTERS, n_masks=N_MASKS, adjoint=False):
    """HBM bytes per launch of the coherent-system kernel, as the committed counter summary
    (FETCH_SIZE and WRITE_SIZE passes, scaled by a 1 GiB copy) recorded them, when that
    record is for the kernel shape this run uses; else None."""
    shape = kernel_shape(plan, n_masks, adjoint)
    try:
        with open(path) as f:
            text = f.read()
    except OSError:
        return None, shape
    rec = dict(re.findall(r"^(traffic_bytes_per_launch|shape): (.+)$", text, re.M))
    if rec.get("shape", "").strip() != shape or "traffic_bytes_per_launch" not in rec:
        return None, shape
    return int(rec["traffic_bytes_per_launch"]), shape


def ratio(a, b):
    return None if a is None or b is None else a / b


def timed(fn, steps, warmup, sync):
    for _ in range(warmup):
        fn()
    sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        fn()
    sync()
    return (time.perf_counter() - t0) / steps


def cpu_baseline(mask, optics, grid, plane):
    """The oracle (reference algorithm, float64, scipy workers=-1) on one image."""
    from oracle import litho_oracle as O
    t0 = time.perf_counter()
    raw, _, _ = O.abbe_sum(mask, optics, grid, [plane])
    dt = time.perf_counter() - t0
    return raw[0], dt


ILT_BATCH = 16


def ilt_leg(args, ilt, dev, sync):
    """ILT forward + backward at 1024^2 (autograd, eager): seconds per step for one mask,
    per step of a batch of ILT_BATCH masks (the step count is ILT_BATCH), and per step of
    the dense per-source form (one dense fft2 / ifft2 per source point, host sync per
    step)."""
    import torch

    from lithobat_amd.ilt import abbe_intensity, ilt_objective, mask_of, relaxed_resist_loss
    from lithobat_amd.masks import manhattan
    scale = 1.0 / float(ilt.clear[0])
    rng = np.random.default_rng(3000)
    theta = torch.tensor(rng.standard_normal((1, N, N)), device=dev, dtype=torch.float32,
                         requires_grad=True)
    target = torch.tensor(manhattan(N, PIXEL, seed=3001)[None], device=dev, dtype=torch.float32)

    def ilt_step():
        theta.grad = None
        raw = abbe_intensity(mask_of(theta, 8.0), ilt)  # [1, 1, N, N]: one plane
        loss = relaxed_resist_loss(raw, target[:, None], scale, 0.3, 50.0)
        loss.backward()

    one = timed(ilt_step, args.ilt_steps, 2, sync)
    thetas = torch.tensor(np.random.default_rng(3100).standard_normal((ILT_BATCH, N, N)),
                          device=dev, dtype=torch.float32, requires_grad=True)
    targets = torch.tensor(np.stack([manhattan(N, PIXEL, seed=3200 + b)
                                     for b in range(ILT_BATCH)]), device=dev,
                           dtype=torch.float32)

    def batch_step():
        thetas.grad = None
        raw = abbe_intensity(mask_of(thetas, 8.0), ilt)
        loss = relaxed_resist_loss(raw, targets[:, None], scale, 0.3, 50.0)
        loss.backward()

    batched = timed(batch_step, max(2, args.ilt_steps // 2), 1, sync) / ILT_BATCH

    # the same objective as one autograd function (ILTObjective): the sigmoid inside the
    # band analysis, the loss gradient applied with the sigmoid derivative on the device
    def fused_step():
        theta.grad = None
        ilt_objective(theta, ilt, target, scale, 0.3, 50.0, beta=8.0).backward()

    def fused_batch_step():
        thetas.grad = None
        ilt_objective(thetas, ilt, targets, scale, 0.3, 50.0, beta=8.0).backward()

    fused_batched = timed(fused_batch_step, max(2, args.ilt_steps // 2), 1, sync) / ILT_BATCH
    # the adjoint coherent-system kernel of the batched step, from HIP events on its stream
    ilt.profile(True)
    for _ in range(max(3, args.ilt_steps // 2)):
        batch_step()
    sync()
    n_adj, adj_ms = ilt.profile_read(adjoint=True)
    n_fwd, fwd_ms = ilt.profile_read()
    ilt.profile(False)
    info = ilt.info
    M, W = info["m"], info["box"]
    flops = 2.0 * ILT_BATCH * info["n_systems"] * field_flops(M, W)
    adj_s = adj_ms / 1e3 / max(n_adj, 1)
    achieved = flops / adj_s / 1e12
    traffic, shape = recorded_traffic(ilt, COUNTERS_ADJ, ILT_BATCH, adjoint=True)
    roof = {"bound": "valu", "achieved": achieved, "peak": PEAK_FP32_TFLOPS,
            "unit": "TFLOP/s", "frac": achieved / PEAK_FP32_TFLOPS,
            "valu_unpacked_ceiling": VALU_UNPACKED_TFLOPS,
            "frac_of_unpacked": achieved / VALU_UNPACKED_TFLOPS, "traffic": traffic,
            "traffic_source": (os.path.relpath(COUNTERS_ADJ, ROOT) if traffic is not None
                               else None),
            "kernel_shape": shape,
            "kernel": "k_field_adj (ILT adjoint of the 16-mask step: fields recomputed, "
                      "transposed radix stages, FP32 VALU)",
            "kernel_ms": adj_s * 1e3, "flops_per_launch": flops,
            "forward_kernel_ms": fwd_ms / max(n_fwd, 1)}
    th = theta.detach()[0]

    def dense_step():
        m = torch.sigmoid(torch.clamp(8.0 * th, -60.0, 60.0))
        raw = ilt.forward_dense(m, batch=8).requires_grad_(True)
        loss = relaxed_resist_loss(raw[None], target, scale, 0.3, 50.0)
        loss.backward()
        gm = ilt.backward_dense(m, raw.grad, batch=8)
        (gm * 8.0 * m * (1.0 - m)).sum().item()  # host sync, as a per-step loop has

    dense = timed(dense_step, max(1, args.dense_images), 1, sync)

    def blur_step():  # the reference's objective with its blurs on: PEB 2 px, mask smoothing 1 px
        theta.grad = None
        raw = abbe_intensity(mask_of(theta, 8.0, 1.0), ilt)
        loss = relaxed_resist_loss(raw, target[:, None], scale, 0.3, 50.0, sigma_px=2.0)
        loss.backward()

    blurred = timed(blur_step, args.ilt_steps, 2, sync)

    # the same autograd steps (forward, loss, backward) captured once into a graph and
    # replayed: the single-mask step is bound by host launches when run eagerly
    g_one = graph_step_seconds(ilt_step, lambda: theta.grad, args, sync)
    g_batch = graph_step_seconds(batch_step, lambda: thetas.grad, args, sync)
    gf_one = graph_step_seconds(fused_step, lambda: theta.grad, args, sync)
    gf_batch = graph_step_seconds(fused_batch_step, lambda: thetas.grad, args, sync)
    fused = {"batched_steps_per_s": ratio(1.0, fused_batched),
             "graph_steps_per_s": ratio(1.0, gf_one),
             "graph_batched_steps_per_s": (None if gf_batch is None
                                           else ILT_BATCH / gf_batch),
             "definition": "ILTObjective: theta -> relaxed-resist cost of the Abbe image of "
                           "sigmoid(8 theta) in one autograd function (sigmoid inside the band "
                           "analysis, objective inside the image synthesis, adjoint from the "
                           "intensity derivative's band spectrum, sigmoid derivative inside the "
                           "mask-gradient synthesis); same value and gradient as the composed "
                           "path"}
    return (one, batched, dense, blurred, g_one,
            None if g_batch is None else g_batch / ILT_BATCH, roof, fused)


def graph_step_seconds(fn, result, args, sync):
    """Seconds per call of ``fn`` captured into one graph and replayed, or None when graphs
    are off or a replay's ``result()`` differs from an eager call's (relative 1e-5)."""
    import torch
    if args.no_graph:
        return None
    fn()
    sync()
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        fn()
    torch.cuda.current_stream().wait_stream(side)
    sync()
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        fn()
    sync()
    sec = timed(graph.replay, max(3, args.ilt_steps), 2, sync)
    graph.replay()
    sync()
    replayed = result().clone()
    fn()
    sync()
    eager = result()
    err = float((replayed - eager).abs().max() / eager.abs().max().clamp_min(1e-30))
    if not err <= 1e-5:
        print(f"graph replay differs from eager (rel {err:.3g}); not reported", file=sys.stderr)
        return None
    return sec


def clips_leg(args, dev, sync):
    """Seconds per batched vector ILT step over 16 uneven clips (graph replay), and per
    plain step (one clip per call, eager, synchronised after every clip)."""
    import torch

    from lithobat_amd.batch import ILTStep, ImagingBatch
    from lithobat_amd.masks import manhattan
    clips = clips_cfg5()
    batch = ImagingBatch(clips, precision=32)
    thetas = [torch.tensor(np.random.default_rng(5000 + c).standard_normal((g.n_pixels,) * 2),
                           device=dev, dtype=torch.float32) for c, (_, g) in enumerate(clips)]
    targets = [torch.tensor(manhattan(g.n_pixels, PIXEL, seed=6000 + c), device=dev,
                            dtype=torch.float32) for c, (_, g) in enumerate(clips)]
    bstep = ILTStep(batch, thetas, targets)
    bstep.run()
    sync()
    bgraph = None
    if not args.no_graph:
        s2 = torch.cuda.Stream()
        s2.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(s2):
            bstep.run()
        torch.cuda.current_stream().wait_stream(s2)
        sync()
        bgraph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(bgraph):
            bstep.run()
        sync()
    clip_ilt_s = timed(bgraph.replay if bgraph is not None else bstep.run,
                       max(3, args.ilt_steps // 2), 1, sync)
    singles = [ILTStep(ImagingBatch([c], precision=32), [t], [tg])
               for c, t, tg in zip(clips, thetas, targets)]

    def plain_clips():
        for one in singles:
            one.run()
            torch.cuda.synchronize()

    return clip_ilt_s, timed(plain_clips, 2, 1, sync)


def socs_cfg3():
    """Hopkins / SOCS configuration: 512^2 @ 8 nm, NA 1.35 (water), annular 0.6-0.9 on a
    21-point grid (142 source points), raw intensities."""
    from lithobat_amd.config import GridConfig, OpticsConfig
    grid = GridConfig(n_pixels=512, pixel_size=8e-9)
    optics = OpticsConfig(wavelength=193e-9, NA=1.35, n_immersion=1.44, source_type="annular",
                          sigma_outer=0.9, sigma_inner=0.6, source_grid=21,
                          normalisation="none")
    return optics, grid


def socs_masks(dev):
    import torch

    from lithobat_amd.masks import manhattan
    return torch.tensor(np.stack([manhattan(512, 8e-9, seed=s) for s in range(2000, 2256)]),
                        device=dev, dtype=torch.float32)


def socs_leg(args, dev, sync):
    """Hopkins / SOCS at 512^2, 8 nm: FP64 TCC decomposition on the device (Gram route,
    Jacobi) truncated to 24 kernels, then 256 masks (seeds 2000-2255) imaged through the
    kernels in complex64; error of the K = 24 image against the full Abbe sum (FP32
    device path, itself within 1e-5 of the float64 oracle) on mask 2000."""
    from lithobat_amd.hopkins import SOCSKernels
    from lithobat_amd.plan import ImagingPlan
    optics, grid = socs_cfg3()
    sync()
    t0 = time.perf_counter()
    k24 = SOCSKernels.build(optics, grid, tol=0.0, max_kernels=24, precision=32)
    sync()
    build_s = time.perf_counter() - t0
    masks = socs_masks(dev)
    plan = k24.plan

    def step():
        plan.forward(masks, with_peak=False)

    sec = timed(step, max(2, args.steps // 2), 1, sync)
    # the staged coherent-system passes of a call (row tables, then column
    # transforms, all passes), timed with HIP events on the call's stream
    plan.profile(True)
    for _ in range(3):
        step()
    sync()
    n_prof, prof_ms = plan.profile_read()
    plan.profile(False)
    kern_ms = prof_ms / max(n_prof, 1)
    info = plan.info
    flops = field_flops(info["m"], info["box"]) * 256 * info["n_systems"] * info["n_planes"]
    achieved = flops / (kern_ms * 1e-3) / 1e12
    roof = {"bound": "valu", "achieved": achieved, "peak": 157.3, "unit": "TFLOP/s",
            "frac": achieved / 157.3, "valu_unpacked_ceiling": 78.6,
            "frac_of_unpacked": achieved / 78.6, "traffic": None,
            "kernel": "k_rows_staged + k_cols_staged (each pass's row tables through the "
                      "last-level cache, every pass of the 256-mask call)",
            "kernel_ms": kern_ms, "flops_per_launch": flops,
            "kernel_shape": kernel_shape(plan, 256)}
    abbe = ImagingPlan(optics, grid, precision=32)
    ref = abbe.forward(masks[:1])[0][0, 0]
    img = plan.forward(masks[:1])[0][0, 0]
    err = float((img - ref).abs().max() / ref.abs().max())
    return k24, {"images_per_s": 256 / sec, "build_s": build_s, "kernels": len(k24),
            "n_source": k24.n_source, "n_band": k24.n_band,
            "energy_kept": float(k24.energy_kept), "rel_err_vs_abbe": err, "roofline": roof,
            "config": "256 masks 512^2 @ 8 nm, NA 1.35 annular 0.6-0.9 (142 points), "
                      "K = 24 of the FP64 Gram / Householder + divide-and-conquer QL decomposition, complex64 imaging"}


def volume_leg(args, masks, sync):
    """Seconds per depth-resolved image through a 100 nm resist (21 optical planes in one
    launch, linear interpolation onto 50 voxels of 2 nm), one 1024^2 mask."""
    from lithobat_amd.config import GridConfig, ResistConfig
    from lithobat_amd.resist3d import exposure_volume
    grid = GridConfig(n_pixels=N, pixel_size=PIXEL, dz=2e-9, n_z_slices=21)
    optics = optics_cfg2()
    resist = ResistConfig(thickness=100e-9, n_resist=1.70, focus_reference="mid")
    m = masks[0]
    return timed(lambda: exposure_volume(m, optics, grid, resist), max(2, args.steps // 2), 1,
                 sync)


def euv_optics(mask_model):
    from lithobat_amd.config import OpticsConfig
    return OpticsConfig(wavelength=13.5e-9, NA=0.33, source_type="annular", sigma_outer=0.8,
                        sigma_inner=0.4, source_grid=11, reduction=4.0, chief_ray_deg=6.0,
                        mask_model=mask_model, normalisation="clear")


def euv_leg(args, dev, sync):
    """Reflective EUV masks: 32 masks 1024^2 @ 2 nm, NA 0.33 annular 0.4-0.8, the default
    EUV stack (60 nm TaBN on 40 Mo/Si bilayers, Ru cap), 6 deg chief ray, FP32.  Images/s
    with every order reflected by the blank at its own angle, the same batch through the
    thin mask, and the plan set-up time (blank solve on the host + device tables)."""
    import torch
    from lithobat_amd.config import GridConfig
    from lithobat_amd.masks import manhattan
    from lithobat_amd.plan import ImagingPlan
    grid = GridConfig(n_pixels=1024, pixel_size=2e-9)
    masks = torch.tensor(np.stack([manhattan(1024, 2e-9, seed=s) for s in range(3000, 3032)]),
                         device=dev, dtype=torch.float32)
    res = {}
    for model in ("multilayer", "thin"):
        t0 = time.perf_counter()
        plan = ImagingPlan(euv_optics(model), grid, precision=32)
        sync()
        build = time.perf_counter() - t0
        sec = timed(lambda: plan.forward(masks, with_peak=False), max(3, args.steps // 2), 1,
                    sync)
        res[model] = {"images_per_s": 32 / sec, "plan_s": build,
                      "systems": plan.info["n_systems"]}
    return {"images_per_s": res["multilayer"]["images_per_s"],
            "thin_images_per_s": res["thin"]["images_per_s"],
            "plan_s": res["multilayer"]["plan_s"], "thin_plan_s": res["thin"]["plan_s"],
            "source_points": res["multilayer"]["systems"],
            "config": "32 masks 1024^2 @ 2 nm, NA 0.33 annular 0.4-0.8 (grid 11), 60 nm TaBN "
                      "on 40 Mo/Si bilayers, chief ray 6 deg on x, FP32, eager"}


def run_ilt_leg(args, sync):
    """The full inverse-lithography solve on the device: 512^2 @ 8 nm, NA 1.35 annular
    0.6-0.9 SOCS kernels (FP32), threshold resist with 10 nm PEB, 4 nm mask smoothing,
    TV and discreteness penalties, 120 Adam iterations (the reference's default; one
    history read per iteration, as its convergence test needs; iterations after the
    first replay one captured graph, capture included in the time).  Seconds per
    iteration."""
    from lithobat_amd.config import GridConfig, OpticsConfig, ResistConfig
    from lithobat_amd.masks import manhattan
    from lithobat_amd.opc import PrintModel, run_ilt
    grid = GridConfig(n_pixels=512, pixel_size=8e-9)
    optics = OpticsConfig(wavelength=193e-9, NA=1.35, n_immersion=1.44, source_type="annular",
                          sigma_outer=0.9, sigma_inner=0.6, source_grid=21,
                          normalisation="clear")
    model = PrintModel(optics, ResistConfig(threshold=0.3, diffusion_sigma=10e-9), grid,
                       precision=32)
    target = manhattan(512, 8e-9, seed=4000)
    kw = dict(smooth=4e-9, weight_tv=1e-3, weight_discrete=1e-3, tol=0.0)
    run_ilt(target, model, max_iter=2, **kw)  # kernels built, shapes warmed
    sync()
    t0 = time.perf_counter()
    res = run_ilt(target, model, max_iter=120, **kw)
    sync()
    return (time.perf_counter() - t0) / res.iterations, res.iterations


def cfg1_leg(args, dev, sync, with_cpu):
    """The plain FP64 configuration on the device: one 256^2 image per call (5 nm pixels,
    vertical lines / spaces pitch 200 nm CD 100 nm, 488 source points), images/s with the
    mask resident; its error against the float64 oracle; and the oracle's own image timed
    in full on the host (median of 5)."""
    import torch

    from lithobat_amd.config import GridConfig
    from lithobat_amd.masks import lines_and_spaces
    from lithobat_amd.plan import ImagingPlan
    grid = GridConfig(n_pixels=256, pixel_size=5e-9)
    optics = optics_cfg1()
    mask = lines_and_spaces(256, 5e-9, 200e-9, 100e-9)
    plan = ImagingPlan(optics, grid, precision=64)
    m = torch.tensor(mask, device=dev, dtype=torch.float64)
    sec = timed(lambda: plan.forward(m), max(10, args.steps), 3, sync)
    raw = plan.forward(m)[0][0, 0].cpu().numpy()
    out = {"images_per_s": 1.0 / sec, "source_points": len(plan.src),
           "coarse_grid": plan.info["m"], "pupil_box": plan.info["box"],
           "config": "1 image 256^2 @ 5 nm, lines/spaces pitch 200 nm CD 100 nm, NA 0.85 "
                     "annular 0.5-0.8 (488 points), FP64 (complex128), one image per call"}
    if with_cpu:
        from oracle import litho_oracle as O
        times = []
        for _ in range(5):
            t0 = time.perf_counter()
            ref = O.compute_aerial_image(mask, optics, grid, return_raw=True)[0]
            times.append(time.perf_counter() - t0)
        cpu_s = float(np.median(times))
        out["fp64_rel_err_vs_oracle"] = float(np.abs(raw - ref).max() / np.abs(ref).max())
        out["cpu_baseline"] = {"value": 1.0 / cpu_s, "unit": "images/s", "cores": _cores(),
                               "kind": "port",
                               "sample": f"the whole image, median of 5 ({cpu_s:.3f} s)"}
        out["vs_cpu"] = cpu_s / sec
    return out


def _cores():
    return len(os.sched_getaffinity(0))


def cpu_socs_leg(k24, sync):
    """CPU baseline of the SOCS configuration: the reference's convolution (fft2 of the
    mask, K inverse transforms through the kernels, sum of lambda |E|^2) on one 512^2 mask,
    three times, with the device's K = 24 kernels; and the reference's TCC build and
    eigendecomposition (the full n_band x n_band TCC, eigh) timed on the same optics at half
    the field (n_band / 4) and scaled by n_band^3 to the full band."""
    from lithobat_amd.config import GridConfig
    from lithobat_amd.masks import manhattan
    from oracle import litho_oracle as O
    optics, grid = socs_cfg3()
    kern = k24.plan.kernels().cpu().numpy()
    k = O.SOCS(kern, np.asarray(k24.weights, dtype=np.float64), float(k24.clear),
               float(k24.energy_kept), int(k24.n_band), int(k24.n_source), grid.n_pixels)
    mask = manhattan(512, 8e-9, seed=2000)
    times = []
    for _ in range(3):
        t0 = time.perf_counter()
        O.socs_raw(k, mask)
        times.append(time.perf_counter() - t0)
    conv_s = float(np.median(times))
    half = GridConfig(n_pixels=256, pixel_size=8e-9)
    t0 = time.perf_counter()
    kh = O.socs_build(optics, half, tol=0.0, max_kernels=24)
    tcc_half = time.perf_counter() - t0
    scale = (k24.n_band / kh.n_band) ** 3
    return {"value": 1.0 / conv_s, "unit": "images/s", "cores": _cores(), "kind": "port",
            "sample": f"1 mask (seed 2000) x 3, median {conv_s:.3f} s, K = 24 kernels",
            "tcc_build_s_sample": tcc_half, "tcc_n_band_sample": int(kh.n_band),
            "tcc_build_s_extrapolated": tcc_half * scale,
            "tcc_sample": f"TCC + eigh at 256^2 @ 8 nm (n_band {kh.n_band}) in "
                          f"{tcc_half:.2f} s, scaled by (n_band ratio)^3 = {scale:.1f} to "
                          f"n_band {k24.n_band}"}


def _oracle_ilt_step(mask, target, optics, grid, every):
    """Seconds of one oracle ILT step (forward image, relaxed resist, Abbe adjoint,
    sigmoid chain) over every ``every``-th source point, and the full source size."""
    from oracle import litho_oracle as O
    full = O.source_points(optics)
    sub = O.SourcePts(full.weight[::every], full.fx[::every], full.fy[::every],
                      full.phi[::every], full.grid)
    planes = [(optics.defocus, optics.n_immersion if optics.n_image is None
               else optics.n_image)]
    t0 = time.perf_counter()
    raw, _, clear = O.abbe_sum(mask, optics, grid, planes, src=sub)
    _, d_raw, _, _ = O.relaxed_objective(raw[0], target, 1.0 / max(clear[0], 1e-30), 0.3, 50.0)
    O.grad_theta(O.abbe_gradient(mask, optics, grid, d_raw, src=sub), mask, 8.0)
    return time.perf_counter() - t0, len(full), len(sub)


def cpu_ilt_leg():
    """CPU baseline of the ILT step at 1024^2, NA 1.35 (148 points): the oracle's forward
    image, relaxed-resist cost and derivative, and Abbe adjoint (opc/ilt.py:311-337 and
    235-266 restated for the Abbe systems) for one mask over every 4th source point, three
    times (median), scaled by the source count (the step is one transform pair per source
    point, plus the mask's own transform)."""
    from lithobat_amd.config import GridConfig
    from lithobat_amd.masks import manhattan
    from oracle import litho_oracle as O
    optics = optics_cfg4()
    grid = GridConfig(n_pixels=N, pixel_size=PIXEL)
    mask = O.mask_of(np.random.default_rng(3000).standard_normal((N, N)), 8.0)
    target = manhattan(N, PIXEL, seed=3001)
    runs = [_oracle_ilt_step(mask, target, optics, grid, 4) for _ in range(3)]
    t, s_full, s_sub = sorted(runs)[1]
    step_s = t * s_full / s_sub
    return {"value": 1.0 / step_s, "unit": "steps/s", "cores": _cores(), "kind": "port",
            "sample": f"1 step (theta seed 3000) over {s_sub} of the {s_full} source points, "
                      f"x 3, median {t:.2f} s, scaled by {s_full}/{s_sub}: {step_s:.1f} s"}


def cpu_clips_leg():
    """CPU baseline of the 16-clip vector ILT step: one clip of each size (512, 768, 1024)
    through the oracle's vector forward, relaxed resist and Abbe adjoint over every 8th
    source point, scaled by the source count; the step time is the sum over the 16 clips
    of the time of a clip of the same size."""
    from lithobat_amd.masks import manhattan
    from oracle import litho_oracle as O
    clips = clips_cfg5()
    per = {}
    for c, (o, g) in enumerate(clips):
        n = int(g.n_pixels)
        if n in per:
            continue
        mask = O.mask_of(np.random.default_rng(5000 + c).standard_normal((n, n)), 8.0)
        t, s_full, s_sub = _oracle_ilt_step(mask, manhattan(n, PIXEL, seed=6000 + c), o, g, 8)
        per[n] = t * s_full / s_sub
    counts = {n: sum(1 for _, g in clips if int(g.n_pixels) == n) for n in per}
    step_s = sum(per[n] * counts[n] for n in per)
    return {"value": 1.0 / step_s, "unit": "steps/s", "cores": _cores(), "kind": "port",
            "sample": f"one clip of each size over {s_sub} of the {s_full} source points, "
                      "scaled by the source count (" +
                      ", ".join(f"{n}^2: {per[n]:.2f} s x {counts[n]}" for n in sorted(per)) +
                      f"), summed over the 16 clips: {step_s:.1f} s per step"}


def dense_leg(args, plan, masks, sync):
    """Images/s of the reference's dataflow on the device: one dense N x N inverse
    transform per source point (8 per launch), host synchronisation per image."""
    import torch
    kd = max(1, args.dense_images)
    plan.forward_dense(masks[0], batch=8)
    sync()
    t0 = time.perf_counter()
    for i in range(kd):
        plan.forward_dense(masks[i % N_MASKS], batch=8)
        torch.cuda.synchronize()
    return kd / (time.perf_counter() - t0)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=40)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-graph", action="store_true")
    ap.add_argument("--plain-images", type=int, default=20,
                    help="images timed for the plain single-image form")
    ap.add_argument("--ilt-steps", type=int, default=10)
    ap.add_argument("--dense-images", type=int, default=3,
                    help="images timed for the dense per-source-point form")
    ap.add_argument("--legs",
                    default="batched,plain,ilt,clips,dense,socs,volume,solve,euv,cfg1,cpu",
                    help="comma list of measurements to run (batched always runs; cpu = the "
                         "CPU baselines of every configuration)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="after the legs, copy a MIB-sized buffer once (a plain copy kernel of "
                         "known size, to scale hardware byte counters)")
    ap.add_argument("--tune", metavar="TABLE", default=None,
                    help="autotune the kernel shapes of this workload and write the table")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU: this build has no multi-GPU path (--gpus {args.gpus}"
              " requested). Shard masks x planes across processes to use more GPUs.",
              file=sys.stderr)
        sys.exit(2)

    legs = set(args.legs.split(","))
    if args.no_cpu:
        legs.discard("cpu")

    import torch

    from lithobat_amd.config import GridConfig
    from lithobat_amd.masks import manhattan
    from lithobat_amd.plan import ImagingPlan

    dev = torch.device("cuda", 0)
    sync = torch.cuda.synchronize
    grid = GridConfig(n_pixels=N, pixel_size=PIXEL)
    optics = optics_cfg2()
    planes = [(dz, 1.0) for dz in DEFOCUS]
    masks_np = np.stack([manhattan(N, PIXEL, seed=s) for s in MASK_SEEDS])
    masks = torch.tensor(masks_np, device=dev, dtype=torch.float32)

    # ---------------- batched: 64 masks x 5 planes per step --------------------
    plan = ImagingPlan(optics, grid, planes, precision=32)
    ilt = ImagingPlan(optics_cfg4(), grid, precision=32)
    if args.tune:
        from lithobat_amd import autotune
        # the coarse grid (80, 96 or 128 points here) and its kernel shapes
        res = autotune.tune_grid(
            lambda m: ImagingPlan(optics, grid, planes, precision=32, coarse_m=m), masks,
            adjoint=False, save=False)
        th = torch.randn((1, N, N), device=dev)
        res.update(autotune.tune(ilt, torch.sigmoid(8 * th), adjoint=True, save=False))
        thb = torch.randn((ILT_BATCH, N, N), device=dev)
        res.update(autotune.tune(ilt, torch.sigmoid(8 * thb), adjoint=True, save=False))
        from lithobat_amd.hopkins import SOCSKernels
        so, sg = socs_cfg3()
        k24 = SOCSKernels.build(so, sg, tol=0.0, max_kernels=24, precision=32)
        sm = socs_masks(dev)
        res.update(autotune.tune(k24.plan, sm, adjoint=False, save=False))
        res.update(autotune.tune(k24.plan, sm[:1], adjoint=True, save=False))
        one = ImagingPlan(optics, grid, planes[:1], precision=32)
        res.update(autotune.tune(one, masks[:1], adjoint=False, save=False))
        table = autotune.load_table()
        table.update(res)
        autotune.save_table(table, args.tune)
        print(json.dumps(res), file=sys.stderr)
        plan = ImagingPlan(optics, grid, planes, precision=32)  # on the chosen grid
    plan._use_tuning(N_MASKS)
    n_img = N_MASKS * len(planes)
    out = torch.empty((N_MASKS, len(planes), N, N), device=dev, dtype=torch.float32)
    peak = torch.empty((N_MASKS, len(planes)), device=dev, dtype=torch.float32)
    from lithobat_amd import _lib
    lib = _lib.load()
    ws = torch.empty(plan.workspace_bytes(N_MASKS), dtype=torch.uint8, device=dev)

    def step_on(stream):
        _lib.check(lib.lb_forward(plan._h, masks.data_ptr(), 0, N_MASKS, out.data_ptr(),
                                  peak.data_ptr(), ws.data_ptr(), ws.numel(), stream))

    step_on(torch.cuda.current_stream().cuda_stream)
    sync()
    graph = None
    if not args.no_graph:
        try:
            s = torch.cuda.Stream()
            s.wait_stream(torch.cuda.current_stream())
            with torch.cuda.stream(s):
                step_on(s.cuda_stream)
            torch.cuda.current_stream().wait_stream(s)
            sync()
            graph = torch.cuda.CUDAGraph()
            with torch.cuda.graph(graph):
                step_on(torch.cuda.current_stream().cuda_stream)
            sync()
        except Exception as e:  # noqa: BLE001 - report and time eagerly
            print(f"graph capture unavailable ({e}); timing eager launches", file=sys.stderr)
            graph = None

    def step():
        if graph is not None:
            graph.replay()
        else:
            step_on(torch.cuda.current_stream().cuda_stream)

    sec = timed(step, args.steps, args.warmup, sync)
    images_per_s = n_img / sec

    # coherent-system kernel time from HIP events on its stream (eager launches)
    plan.profile(True)
    for _ in range(max(3, args.steps // 2)):
        step_on(torch.cuda.current_stream().cuda_stream)
    sync()
    launches, kern_ms = plan.profile_read()
    plan.profile(False)
    kern_s = kern_ms / 1e3 / launches
    M, W = plan.info["m"], plan.info["box"]
    n_sys = plan.info["n_systems"] * n_img
    flops = n_sys * field_flops(M, W)
    achieved_tflops = flops / kern_s / 1e12
    traffic, shape = recorded_traffic(plan)

    # ---------------- plain: one (mask, plane) per call, sync per image ---------
    plain_plans = [ImagingPlan(optics, grid, [pl], precision=32) for pl in planes]
    plain_images_per_s = None
    if "plain" in legs:
        k = max(1, args.plain_images)

        def plain_image(i):
            p = plain_plans[i % len(planes)]
            p.forward(masks[(i // len(planes)) % N_MASKS])
            torch.cuda.synchronize()

        for i in range(3):
            plain_image(i)
        t0 = time.perf_counter()
        for i in range(k):
            plain_image(i)
        plain_images_per_s = k / (time.perf_counter() - t0)

    # ---------------- ILT forward + backward at 1024^2 ---------------------------
    ilt_s = ilt_batched_s = ilt_dense_s = ilt_blur_s = vol_s = None
    ilt_graph_s = ilt_graph_batched_s = None
    clip_ilt_s = clip_plain_s = dense_images_per_s = None
    ilt_roof = ilt_fused = None
    if "ilt" in legs:
        (ilt_s, ilt_batched_s, ilt_dense_s, ilt_blur_s, ilt_graph_s,
         ilt_graph_batched_s, ilt_roof, ilt_fused) = ilt_leg(args, ilt, dev, sync)
    if "volume" in legs:
        vol_s = volume_leg(args, masks, sync)
    solve = run_ilt_leg(args, sync) if "solve" in legs else None
    if "clips" in legs:
        clip_ilt_s, clip_plain_s = clips_leg(args, dev, sync)
    if "dense" in legs:
        dense_images_per_s = dense_leg(args, plain_plans[0], masks, sync)
    socs = k24 = None
    if "socs" in legs:
        k24, socs = socs_leg(args, dev, sync)
    cfg1 = cfg1_leg(args, dev, sync, "cpu" in legs) if "cfg1" in legs else None
    euv = euv_leg(args, dev, sync) if "euv" in legs else None
    if args.copy_calibration > 0:
        nb = args.copy_calibration << 20
        a = torch.empty(nb, dtype=torch.uint8, device=dev).fill_(1)
        b = torch.empty_like(a)
        sync()
        b.copy_(a)
        sync()
        del a, b

    # ---------------- error vs float64 and the CPU baseline ---------------------
    err = None
    cpu = None
    if "cpu" in legs:
        ref, cpu_s = cpu_baseline(masks_np[0], optics, grid, planes[0])
        gpu0 = out[0, 0].double().cpu().numpy()
        err = float(np.abs(gpu0 - ref).max() / np.abs(ref).max())
        cores = len(os.sched_getaffinity(0))
        cpu = {"value": 1.0 / cpu_s, "unit": "images/s", "cores": cores, "kind": "port",
               "sample": "1 image (mask seed 1000, defocus -100 nm) of the 64 x 5 batch: the "
                         "oracle's float64 Abbe sum with scipy.fft workers=-1 and 64 MiB "
                         f"transform batches; {cpu_s:.2f} s on {cores} host threads"}
        if socs is not None:
            socs["cpu_baseline"] = cpu_socs_leg(k24, sync)
            socs["vs_cpu"] = socs["images_per_s"] / socs["cpu_baseline"]["value"]
        cpu_ilt = cpu_ilt_leg() if "ilt" in legs else None
        cpu_clips = cpu_clips_leg() if "clips" in legs else None
    else:
        cpu_ilt = cpu_clips = None

    line = {
        "metric": "aerial images/s and ILT grad steps/s at 1024² (batched vs plain, ≤1e-5 rel err)",
        "value": images_per_s,
        "unit": "images/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": sec * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32 (complex64 fields, FP32 accumulation)",
        "data": "synthetic: 64 seeded Manhattan masks (seeds 1000-1063), 5 defocus planes",
        "config": {"workload": "batched scalar Abbe, 64 masks x 5 defocus, 1024^2 @ 4 nm, "
                               "NA 0.93 quadrupole (148 source points), FP32",
                   "images_per_step": n_img, "n": N, "coarse_grid": M, "pupil_box": W,
                   "graph": graph is not None, "parallelism": "single GPU"},
        "plain_images_per_s": plain_images_per_s,
        "batched_vs_plain": ratio(images_per_s, plain_images_per_s),
        "plain_dense_images_per_s": dense_images_per_s,
        "batched_vs_plain_dense": ratio(images_per_s, dense_images_per_s),
        "plain_definitions": {
            "plain": "this library, one (mask, plane) per call, host sync per image, no graph",
            "plain_dense": "the reference's dataflow on the device: one dense 1024^2 inverse "
                           "transform per source point (8 per launch), |E|^2 accumulated, host "
                           "sync per image"},
        "clip_ilt_steps_per_s": ratio(1.0, clip_ilt_s),
        "clip_ilt_plain_steps_per_s": ratio(1.0, clip_plain_s),
        "clip_ilt_batched_vs_plain": ratio(clip_plain_s, clip_ilt_s),
        "clip_ilt_config": "16 clips {512,768,1024}^2 @ 4 nm, NA 1.35 TE vector, Zernike 4-11, "
                           "defocus +-50 nm, forward + adjoint, one persistent coherent-system "
                           "launch per pass, graph-captured step; plain = one clip per call, "
                           "eager, sync per clip",
        "clip_ilt_cpu_baseline": cpu_clips,
        "clip_ilt_vs_cpu": ratio(None if cpu_clips is None else 1.0 / cpu_clips["value"],
                                 clip_ilt_s),
        "ilt_steps_per_s": ratio(1.0, ilt_s),
        "ilt_batched_steps_per_s": ratio(1.0, ilt_batched_s),
        "ilt_blur_steps_per_s": ratio(1.0, ilt_blur_s),
        "ilt_graph_steps_per_s": ratio(1.0, ilt_graph_s),
        "ilt_graph_batched_steps_per_s": ratio(1.0, ilt_graph_batched_s),
        "ilt_fused": ilt_fused,
        "exposure_volumes_per_s": ratio(1.0, vol_s),
        "run_ilt_iterations_per_s": None if solve is None else 1.0 / solve[0],
        "run_ilt_config": "512^2 @ 8 nm, NA 1.35 annular (142 pts) SOCS kernels FP32, "
                          "threshold resist, PEB 10 nm, smoothing 4 nm, TV + discreteness, "
                          "120 Adam iterations (graph-replayed after the first)",
        "exposure_volume_config": "1 mask 1024^2 @ 4 nm, NA 0.93 quadrupole (148 pts), resist "
                                  "100 nm / n 1.70, focus mid, 21 optical planes -> 50 voxels "
                                  "of 2 nm, FP32",
        "ilt_plain_dense_steps_per_s": ratio(1.0, ilt_dense_s),
        "ilt_batched_vs_plain_dense": ratio(ilt_dense_s, ilt_batched_s),
        "ilt_single_vs_plain_dense": ratio(ilt_dense_s, ilt_s),
        "ilt_config": "1024^2 @ 4 nm, NA 1.35 quadrupole (148 pts), sigmoid(8 theta), "
                      "threshold 0.3 steepness 50, FP32 forward + adjoint; batched = "
                      f"{ILT_BATCH} masks per call (steps/s counts masks); plain dense = "
                      "one dense fft2/ifft2 pair per source point, host sync per step; "
                      "blur = PEB sigma 2 px and mask smoothing 1 px on; graph = the "
                      "same autograd step captured once into a graph and replayed, 1 mask "
                      f"and {ILT_BATCH} masks per step",
        "ilt_cpu_baseline": cpu_ilt,
        "ilt_batched_vs_cpu": ratio(None if cpu_ilt is None else 1.0 / cpu_ilt["value"],
                                    ilt_batched_s),
        "ilt_roofline": ilt_roof,
        "socs": socs,
        "plain_fp64": cfg1,
        "euv_multilayer": euv,
        "fp32_rel_err_vs_fp64": err,
        "roofline": {"bound": "valu", "achieved": achieved_tflops, "peak": PEAK_FP32_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved_tflops / PEAK_FP32_TFLOPS,
                     "valu_unpacked_ceiling": VALU_UNPACKED_TFLOPS,
                     "frac_of_unpacked": achieved_tflops / VALU_UNPACKED_TFLOPS,
                     "traffic": traffic,
                     "traffic_source": ("profiles/k_field_fwd_abbe_1024_64x5_counters.md"
                                        if traffic is not None else None),
                     "kernel_shape": shape,
                     "kernel": "k_field_fwd (coherent systems, FP32 VALU radix FFTs)",
                     "kernel_ms": kern_s * 1e3, "flops_per_launch": flops},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
