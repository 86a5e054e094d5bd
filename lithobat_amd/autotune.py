"""Autotuner for the coherent-system kernel schedule.

For a plan and a batch size it times every compiled kernel shape that fits
(threads per block x column splits, i.e. the radix-stage tiling of the
coarse-grid transforms) and a few work-item sizes (systems per block), with
HIP events on the launch stream, and keeps the fastest.  Winners are stored in
a JSON table keyed on (precision, coarse grid M, pupil box W, images per call
and systems per image, both bucketed to powers of two), so one tuning run
serves every later plan of the same shape.  The table lives at
``$LITHOBAT_TUNING`` or ``lithobat_amd/tuning.json``.

``tune_grid`` also chooses the coarse grid itself (and with it the radix
factorisation of its transforms: 80 = 10 x 8, 96 = 12 x 8, 128 = 16 x 8, ...):
it builds the plan on every admissible grid (side >= 2W - 1, coarser than the
image), tunes each, and records the fastest under (precision, N, W); a plan
created without an explicit ``coarse_m`` is built on the recorded grid.
"""
from __future__ import annotations

import ctypes as C
import json
import os
import threading

import torch

from . import _lib

TABLE_PATH = os.environ.get(
    "LITHOBAT_TUNING", os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuning.json"))
_lock = threading.Lock()
SPREAD = -2  # chunk value: the call's systems in equal runs, one per resident block
_table: dict | None = None


def _bucket(x: int) -> int:
    b = 1
    while b < x:
        b *= 2
    return b


def key_of(plan, n_masks: int, adjoint: bool) -> str:
    i = plan.info
    n_img = n_masks * i["n_planes"]
    return (f"p{plan.precision}-M{i['m']}-W{i['box']}-img{_bucket(n_img)}"
            f"-sys{_bucket(i['n_systems'])}-{'adj' if adjoint else 'fwd'}")


def load_table() -> dict:
    global _table
    with _lock:
        if _table is None:
            try:
                with open(TABLE_PATH) as f:
                    _table = json.load(f)
            except (OSError, ValueError):
                _table = {}
        return _table


def save_table(table: dict, path: str | None = None) -> None:
    path = path or TABLE_PATH
    tmp = path + ".tmp"
    with open(tmp, "w") as f:
        json.dump(table, f, indent=1, sort_keys=True)
    os.replace(tmp, path)


def shapes(plan, adjoint: bool) -> list[tuple[int, int]]:
    lib = _lib.load()
    buf = (C.c_int32 * 64)()
    n = C.c_int32()
    _lib.check(lib.lb_plan_shapes(plan._h, 1 if adjoint else 0, buf, 32, C.byref(n)))
    return [(buf[2 * k], buf[2 * k + 1]) for k in range(min(n.value, 32))]


def set_shape(plan, adjoint: bool, nt: int, cs: int, chunk: int, pin: bool = True) -> None:
    """Select a kernel shape; ``pin`` keeps the plan's automatic per-batch-size selection
    from replacing it (``unpin`` hands the choice back)."""
    _set(plan, adjoint, nt, cs, chunk)
    if pin:
        plan._pinned = True


def unpin(plan) -> None:
    plan._pinned = False
    plan._shape_for = None


def _set(plan, adjoint: bool, nt: int, cs: int, chunk: int) -> None:
    _lib.check(_lib.load().lb_plan_set_shape(plan._h, 1 if adjoint else 0, nt, cs, chunk))


def apply(plan, n_masks: int) -> bool:
    """Select the tuned shapes for this batch size, or the plan's defaults (own kernel
    shape, automatic work items) where the table has none."""
    table = load_table()
    hit = False
    for adjoint in (False, True):
        e = table.get(key_of(plan, n_masks, adjoint))
        done = False
        if e:
            try:
                _set(plan, adjoint, e["threads"], e["splits"], e["chunk"])
                hit = done = True
            except (ValueError, NotImplementedError):
                pass
        if not done:
            _set(plan, adjoint, 0, 0, 0)
    return hit


# coarse grid sides with compiled coherent-system kernels (radix plans)
FIELD_SIZES = (16, 32, 48, 64, 80, 96, 128, 256)


def grid_key(precision: int, n: int, box: int) -> str:
    return f"p{precision}-N{n}-W{box}-grid"


def grid_candidates(n: int, box: int, limit: int = 3) -> list[int]:
    """Admissible coarse grids for an n-pixel image with a pupil box of ``box`` bins: the
    smallest ``limit`` compiled sides >= 2 box - 1 that are coarser than the image."""
    return [m for m in FIELD_SIZES if m >= 2 * box - 1 and m < n][:limit]


def recorded_grid(precision: int, n: int, box: int) -> int:
    """The coarse grid recorded by ``tune_grid`` for this geometry, or 0."""
    e = load_table().get(grid_key(precision, n, box))
    try:
        return int(e["coarse_m"]) if e else 0
    except (KeyError, TypeError, ValueError):
        return 0


def tune_grid(make_plan, masks: torch.Tensor, adjoint: bool = True, reps: int = 5,
              save: bool = True, candidates=None) -> dict:
    """Choose the coarse grid: for each admissible side M, ``make_plan(M)`` builds the
    plan on it, its kernel shapes are tuned for this batch (``tune``), and the forward
    (plus adjoint) time with the winning shapes decides.  Returns the shape winners of
    every grid and, under ``grid_key``, the chosen side with every candidate's time."""
    probe = make_plan(0)
    i = probe.info
    prec, n, box = probe.precision, probe.n, i["box"]
    cands = list(candidates) if candidates else grid_candidates(n, box)
    if i["m"] not in cands and i["m"] < n:
        cands.append(i["m"])
    del probe
    results, times = {}, {}
    for m in sorted(set(cands)):
        plan = make_plan(m)
        if plan.info["m"] != m:
            continue
        res = tune(plan, masks, adjoint=adjoint, reps=reps, save=False)
        results.update(res)
        times[m] = sum(v["ms"] for v in res.values())
    if not times:
        raise ValueError("no admissible coarse grid")
    best = min(times, key=times.get)
    results[grid_key(prec, n, box)] = {
        "coarse_m": best, "ms": round(times[best], 4),
        "candidates": {str(m): round(t, 4) for m, t in sorted(times.items())}}
    if save:
        table = load_table()
        with _lock:
            table.update(results)
        save_table(table)
    return results


def _time(fn, reps: int) -> float:
    fn()
    torch.cuda.synchronize()
    start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    start.record()
    for _ in range(reps):
        fn()
    stop.record()
    stop.synchronize()
    return start.elapsed_time(stop) / reps


def tune(plan, masks: torch.Tensor, adjoint: bool = True, reps: int = 5,
         chunks=(0, 1, 2, 4, 8, 12, 16, 24, 32), save: bool = True,
         candidates: bool = False) -> dict:
    """Time every candidate for ``masks`` (the real batch) and record the winners.

    ``candidates``: also return every candidate's time under ``"candidates"``
    (the saved table keeps the winners only).
    """
    n_masks = masks.shape[0] if masks.dim() == 3 else 1
    was_pinned = plan._pinned
    plan._pinned = True  # keep the table from overriding the candidates
    P = len(plan.planes)
    results = {}
    spp = plan.info["n_systems"]
    chunk_cands = set(chunks)
    # work-item sizes that fill the CUs in k rounds of one block each: a block
    # of the large shapes holds a whole CU, and a round with a few idle CUs or
    # a short last round costs as much as a full one
    cus = torch.cuda.get_device_properties(masks.device).multi_processor_count
    n_img = n_masks * P
    for k in (1, 2, 3, 4):
        per_img = (k * cus) // max(n_img, 1)
        if per_img >= 1:
            chunk_cands.add(-(-spp // per_img))
    chunk_cands = sorted({c for c in chunk_cands if c <= max(spp, 1)})
    chunk_cands.append(SPREAD)  # one run of systems per resident block
    timings = {}
    for adj in ([False, True] if adjoint else [False]):
        best = None
        cands = timings.setdefault(key_of(plan, n_masks, adj), [])
        d_img = torch.randn((n_masks, P, plan.n, plan.n), device=masks.device,
                            dtype=plan.real_dtype) if adj else None
        for nt, cs in shapes(plan, adj):
            for ch in chunk_cands:
                if ch == SPREAD and cs == 0:
                    continue  # the staged kernels lay out their own work
                set_shape(plan, adj, nt, cs, ch)
                if adj:
                    ms = _time(lambda: plan.backward(masks, d_img), reps)
                else:
                    ms = _time(lambda: plan.forward(masks, with_peak=False), reps)
                cands.append({"threads": nt, "splits": cs, "chunk": ch, "ms": round(ms, 4)})
                if best is None or ms < best[0]:
                    best = (ms, nt, cs, ch)
        ms, nt, cs, ch = best
        set_shape(plan, adj, nt, cs, ch)
        results[key_of(plan, n_masks, adj)] = {"threads": nt, "splits": cs, "chunk": ch,
                                               "ms": round(ms, 4)}
    if save:
        table = load_table()
        with _lock:
            table.update(results)
        save_table(table)
    # the winners stay selected for this batch size; another size selects its own
    plan._pinned = was_pinned
    plan._shape_for = n_masks
    if candidates:
        for k, v in timings.items():
            results[k] = dict(results[k], candidates=v)
    return results
