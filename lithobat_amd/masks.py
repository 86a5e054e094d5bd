"""Synthetic mask inputs: line/space gratings and seeded Manhattan layouts."""
from __future__ import annotations

import numpy as np


def lines_and_spaces(n_pixels: int, pixel_size: float, pitch: float, cd: float,
                     orientation: str = "vertical") -> np.ndarray:
    """Binary L/S grating with a line centred at the origin (mask/patterns.py:22-66)."""
    coords = (np.arange(n_pixels) - n_pixels // 2) * pixel_size
    line = (np.mod(coords + pitch / 2, pitch) < cd).astype(np.float64)
    if orientation == "vertical":
        return np.tile(line[np.newaxis, :], (n_pixels, 1))
    if orientation == "horizontal":
        return np.tile(line[:, np.newaxis], (1, n_pixels))
    raise ValueError(f"orientation must be 'vertical' or 'horizontal', got '{orientation}'")


def manhattan(n: int, pixel_size: float, seed: int, n_rect=(20, 60), wmin: float = 40e-9,
              wmax: float = 200e-9) -> np.ndarray:
    """Seeded union of axis-aligned rectangles, transmission 0/1.

    Rectangle count uniform in ``n_rect`` (inclusive), widths and heights
    uniform in [wmin, wmax] snapped to whole pixels, positions uniform on the
    periodic grid.  Deterministic in ``seed`` (numpy ``default_rng``).
    """
    rng = np.random.default_rng(seed)
    m = np.zeros((n, n), dtype=np.float64)
    count = int(rng.integers(n_rect[0], n_rect[1] + 1))
    for _ in range(count):
        w = max(1, int(round(rng.uniform(wmin, wmax) / pixel_size)))
        h = max(1, int(round(rng.uniform(wmin, wmax) / pixel_size)))
        y0 = int(rng.integers(0, n))
        x0 = int(rng.integers(0, n))
        m[np.ix_(np.arange(y0, y0 + h) % n, np.arange(x0, x0 + w) % n)] = 1.0
    return m
