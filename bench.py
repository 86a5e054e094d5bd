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
* ILT forward + backward steps/s at 1024^2 (theta -> sigmoid(8 theta) -> Abbe
  at NA 1.35 quadrupole -> relaxed resist L2 vs a Manhattan target);
* the FP32 image error against the float64 oracle on a sample image;
* roofline of the coherent-system kernel (HIP events on its stream);
* the oracle (the reference algorithm in float64 NumPy/SciPy) timed on the
  host cores on a bounded sample, as the CPU baseline.

    python bench.py [--gpus 1] [--steps 10] [--warmup 3]
"""
from __future__ import annotations

import argparse
import json
import math
import os
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


def field_flops(M: int, W: int) -> float:
    """Algorithmic flops of one coherent system on the coarse grid.

    (W + M) length-M radix FFTs at 5 M log2 M, the spectrum x transfer product
    (6 W^2) and the weighted |E|^2 accumulation (4 M^2).
    """
    return 5.0 * M * math.log2(M) * (W + M) + 6.0 * W * W + 4.0 * M * M


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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-graph", action="store_true")
    ap.add_argument("--plain-images", type=int, default=20,
                    help="images timed for the plain single-image form")
    ap.add_argument("--ilt-steps", type=int, default=10)
    ap.add_argument("--tune", metavar="TABLE", default=None,
                    help="autotune the kernel shapes of this workload and write the table")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU: this build has no multi-GPU path (--gpus {args.gpus}"
              " requested). Shard masks x planes across processes to use more GPUs.",
              file=sys.stderr)
        sys.exit(2)

    import torch

    from lithobat_amd.config import GridConfig
    from lithobat_amd.ilt import abbe_intensity, mask_of, relaxed_resist_loss
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
        res = autotune.tune(plan, masks, adjoint=False, save=False)
        th = torch.randn((1, N, N), device=dev)
        res.update(autotune.tune(ilt, torch.sigmoid(8 * th), adjoint=True, save=False))
        one = ImagingPlan(optics, grid, planes[:1], precision=32)
        res.update(autotune.tune(one, masks[:1], adjoint=False, save=False))
        table = autotune.load_table()
        table.update(res)
        autotune.save_table(table, args.tune)
        print(json.dumps(res), file=sys.stderr)
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

    # ---------------- plain: one (mask, plane) per call, sync per image ---------
    plain_plans = [ImagingPlan(optics, grid, [pl], precision=32) for pl in planes]
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
    plain_s = (time.perf_counter() - t0) / k
    plain_images_per_s = 1.0 / plain_s

    # ---------------- ILT forward + backward at 1024^2 ---------------------------
    scale = 1.0 / float(ilt.clear[0])
    rng = np.random.default_rng(3000)
    theta = torch.tensor(rng.standard_normal((1, N, N)), device=dev, dtype=torch.float32,
                         requires_grad=True)
    target = torch.tensor(manhattan(N, PIXEL, seed=3001)[None], device=dev, dtype=torch.float32)

    def ilt_step():
        theta.grad = None
        raw = abbe_intensity(mask_of(theta, 8.0), ilt)[:, 0]
        loss = relaxed_resist_loss(raw, target, scale, 0.3, 50.0)
        loss.backward()

    ilt_s = timed(ilt_step, args.ilt_steps, 2, sync)

    # ---------------- error vs float64 and the CPU baseline ---------------------
    err = None
    cpu = None
    if not args.no_cpu:
        ref, cpu_s = cpu_baseline(masks_np[0], optics, grid, planes[0])
        gpu0 = out[0, 0].double().cpu().numpy()
        err = float(np.abs(gpu0 - ref).max() / np.abs(ref).max())
        cores = len(os.sched_getaffinity(0))
        cpu = {"value": 1.0 / cpu_s, "unit": "images/s", "cores": cores, "kind": "port",
               "sample": "1 image (mask seed 1000, defocus -100 nm) of the 64 x 5 batch: the "
                         "oracle's float64 Abbe sum with scipy.fft workers=-1 and 64 MiB "
                         f"transform batches; {cpu_s:.2f} s on {cores} host threads"}

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
        "batched_vs_plain": images_per_s / plain_images_per_s,
        "ilt_steps_per_s": 1.0 / ilt_s,
        "ilt_config": "1024^2 @ 4 nm, NA 1.35 quadrupole (148 pts), sigmoid(8 theta), "
                      "threshold 0.3 steepness 50, FP32 forward + adjoint",
        "fp32_rel_err_vs_fp64": err,
        "roofline": {"bound": "mfma", "achieved": achieved_tflops, "peak": PEAK_FP32_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved_tflops / PEAK_FP32_TFLOPS,
                     "traffic": None,
                     "kernel": "k_field_fwd (coherent systems, FP32 VALU radix FFTs)",
                     "kernel_ms": kern_s * 1e3, "flops_per_launch": flops},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
