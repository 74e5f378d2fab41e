"""Seeded synthetic NV12 frames: the workload generator for tests and bench.

Y plane: 128 plus three sinusoidal gratings (orientation U[0, pi), period
U[16, 128] px, amplitude U[12, 28], phase U[0, 2 pi)), then 16 axis-aligned
rectangles (side U[W/32, W/4] x U[H/32, H/4], uniform position) painted at a
level U[16, 235] that keeps half of the underlying texture, then Gaussian noise
sigma = 4, rounded and clipped to [16, 235].  UV: (W/2) x (H/2) interleaved
Cb, Cr, each a linear gradient 128 +- 40 across the plane along its own random
direction.  Every draw comes from numpy's PCG64 generator seeded with `seed`.
"""

from __future__ import annotations

import numpy as np


def synthetic_nv12(width: int, height: int, seed: int) -> tuple[np.ndarray, np.ndarray]:
    """Returns (luma[h, w] uint8, chroma[h/2, w] uint8 interleaved Cb/Cr)."""
    if width % 2 or height % 2:
        raise ValueError("NV12 needs even dimensions")
    rng = np.random.default_rng(seed)
    yy, xx = np.mgrid[0:height, 0:width].astype(np.float64)
    y = np.full((height, width), 128.0)
    for _ in range(3):
        theta = rng.uniform(0.0, np.pi)
        period = rng.uniform(16.0, 128.0)
        amp = rng.uniform(12.0, 28.0)
        phase = rng.uniform(0.0, 2.0 * np.pi)
        y += amp * np.sin(2.0 * np.pi * (xx * np.cos(theta) + yy * np.sin(theta)) / period + phase)
    del xx, yy
    for _ in range(16):
        rw = int(rng.integers(max(1, width // 32), width // 4 + 1))
        rh = int(rng.integers(max(1, height // 32), height // 4 + 1))
        x0 = int(rng.integers(0, width - rw + 1))
        y0 = int(rng.integers(0, height - rh + 1))
        level = rng.uniform(16.0, 235.0)
        sub = y[y0 : y0 + rh, x0 : x0 + rw]
        y[y0 : y0 + rh, x0 : x0 + rw] = level + 0.5 * (sub - 128.0)
    y += rng.normal(0.0, 4.0, size=(height, width))
    luma = np.clip(np.rint(y), 16, 235).astype(np.uint8)
    del y

    ch, cw = height // 2, width // 2
    cyy, cxx = np.mgrid[0:ch, 0:cw].astype(np.float64)
    radius = 0.5 * np.hypot(cw, ch)
    chroma = np.empty((ch, width), np.uint8)
    for k in range(2):
        a = rng.uniform(0.0, 2.0 * np.pi)
        u = ((cxx - cw / 2.0) * np.cos(a) + (cyy - ch / 2.0) * np.sin(a)) / radius
        chroma[:, k::2] = np.clip(np.rint(128.0 + 40.0 * u), 88, 168).astype(np.uint8)
    return luma, chroma


def nv12_bytes(luma: np.ndarray, chroma: np.ndarray) -> np.ndarray:
    """Packed NV12 (luma rows then interleaved chroma rows), as pycujpegxl takes."""
    return np.concatenate([luma.reshape(-1), chroma.reshape(-1)])
