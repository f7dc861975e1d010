"""Seeded synthetic terrains for the benchmark and parity workloads.

Generators use numpy PCG64 (``default_rng(seed)``); every workload is named by
its sizes.  The reference generates piles with gradient noise
(slope_model.py:77-129); the multi-octave value noise here is this project's
own choice and only needs to be seeded and reproducible.
"""

from __future__ import annotations

import numpy as np


def gaussian_pile(ny: int, nx: int, dx: float, rng: np.random.Generator, bumps: int = 6,
                  height=(0.3, 2.0), sigma=(0.5, 1.5)) -> np.ndarray:
    y = np.arange(ny)[:, None] * dx
    x = np.arange(nx)[None, :] * dx
    b = np.zeros((ny, nx))
    for _ in range(bumps):
        cy, cx = rng.uniform(0, ny * dx), rng.uniform(0, nx * dx)
        a, s = rng.uniform(*height), rng.uniform(*sigma)
        b += a * np.exp(-((x - cx) ** 2 + (y - cy) ** 2) / (2.0 * s * s))
    return b


def mobile_patch(ny: int, nx: int, rng: np.random.Generator, size: int = 32, hmax: float = 0.2) -> np.ndarray:
    h = np.zeros((ny, nx))
    s = min(size, ny, nx)
    r0, c0 = int(rng.integers(0, ny - s + 1)), int(rng.integers(0, nx - s + 1))
    h[r0:r0 + s, c0:c0 + s] = rng.uniform(0.0, hmax, size=(s, s))
    return h


def pile_with_patch(n: int = 128, dx: float = 0.05, seed: int = 0):
    """128^2 FP64 CPU-reference workload: 6 Gaussian bumps + uniform(0, 0.2) 32^2 patch."""
    rng = np.random.default_rng(seed)
    b = gaussian_pile(n, n, dx, rng)
    h = mobile_patch(n, n, rng)
    return b, h, np.zeros((n, n)), np.zeros((n, n))


def value_noise(ny: int, nx: int, rng: np.random.Generator, octaves: int = 5, base_cells: int = 4,
                lacunarity: float = 2.0, persistence: float = 0.5) -> np.ndarray:
    out = np.zeros((ny, nx))
    amp, total, cells = 1.0, 0.0, float(base_cells)
    for _ in range(octaves):
        gy, gx = int(np.ceil(cells)) + 2, int(np.ceil(cells)) + 2
        lattice = rng.uniform(0.0, 1.0, size=(gy, gx))
        ys = np.linspace(0.0, cells, ny)
        xs = np.linspace(0.0, cells, nx)
        y0, x0 = np.floor(ys).astype(int), np.floor(xs).astype(int)
        fy, fx = ys - y0, xs - x0
        fy = fy * fy * (3 - 2 * fy)
        fx = fx * fx * (3 - 2 * fx)
        a = lattice[y0[:, None], x0[None, :]]
        b = lattice[y0[:, None], x0[None, :] + 1]
        c = lattice[y0[:, None] + 1, x0[None, :]]
        d = lattice[y0[:, None] + 1, x0[None, :] + 1]
        top = a + (b - a) * fx[None, :]
        bot = c + (d - c) * fx[None, :]
        out += amp * (top + (bot - top) * fy[:, None])
        total += amp
        amp *= persistence
        cells *= lacunarity
    return out / total


def cut_face_terrain(n: int = 701, dx: float = 0.05, seed: int = 0, amplitude: float = 1.5,
                     face_deg: float = 40.0):
    """701^2 full-size workload: value-noise heightfield with a planar 40 degree cut face.

    Returns (b_eff, h_mobile, face_mask).  The lower floor and upper plateau are
    noise; the face rises at ``face_deg`` between them; h_mobile is
    uniform(0.05, 0.3) m on the face cells.
    """
    rng = np.random.default_rng(seed)
    upper = 0.5 + amplitude * value_noise(n, n, rng)
    floor = 0.2 + 0.3 * value_noise(n, n, rng, octaves=3)
    x = np.arange(n)[None, :] * dx
    x0 = 0.35 * n * dx
    plane = floor + np.tan(np.deg2rad(face_deg)) * (x - x0)
    b = np.clip(plane, floor, upper)
    face = (plane > floor) & (plane < upper)
    h = np.where(face, rng.uniform(0.05, 0.3, size=(n, n)), 0.0)
    return b, h, face


BATCH_SIZES = (64, 96, 128, 192, 256, 384, 512)


def batched_piles(count: int = 256, dx: float = 0.05, seed: int = 0, sizes=BATCH_SIZES):
    """Uneven batch of square terrains, each with its own pile and mobile patch."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        n = int(rng.choice(sizes))
        b = gaussian_pile(n, n, dx, rng)
        h = mobile_patch(n, n, rng, size=min(32, n))
        out.append((b, h))
    return out


# bucket footprint of the digging workload: 1.5 m along the sweep x 0.8 m across
BUCKET_ALONG_M, BUCKET_ACROSS_M = 1.5, 0.8


def digging_sites(count: int = 64, n: int = 256, dx: float = 0.05, seed: int = 0, sweep_m: float = 3.6):
    """Digging workload: `count` terrains of n^2 (Gaussian piles on a 0.2 m
    floor, no mobile layer) and per terrain a seeded straight bucket line.
    Returns (terrains, lines); a line is (start_y_m, start_x_m, heading_rad),
    chosen so that the bucket stays inside the terrain over `sweep_m`."""
    rng = np.random.default_rng(seed)
    terrains, lines = [], []
    size = n * dx
    margin = sweep_m + 0.5 * BUCKET_ALONG_M + dx
    for _ in range(count):
        b = 0.2 + gaussian_pile(n, n, dx, rng)
        terrains.append((b, np.zeros((n, n))))
        heading = float(rng.uniform(0.0, 2.0 * np.pi))
        y0 = float(rng.uniform(margin * abs(np.sin(heading)) + 1.0, size - margin * abs(np.sin(heading)) - 1.0))
        x0 = float(rng.uniform(margin * abs(np.cos(heading)) + 1.0, size - margin * abs(np.cos(heading)) - 1.0))
        lines.append((y0, x0, heading))
    return terrains, lines


def bucket_cells(line, travelled_m: float, n: int, dx: float = 0.05) -> np.ndarray:
    """Flat indices of the cells under the bucket after `travelled_m` along
    `line`: an axis-aligned 1.5 m x 0.8 m rectangle (the long side on the axis
    nearer the heading) centred on the bucket position, clipped to the terrain."""
    y0, x0, heading = line
    cy, cx = y0 + travelled_m * np.sin(heading), x0 + travelled_m * np.cos(heading)
    along, across = int(round(BUCKET_ALONG_M / dx)), int(round(BUCKET_ACROSS_M / dx))
    rows, cols = (across, along) if abs(np.cos(heading)) >= abs(np.sin(heading)) else (along, across)
    r0 = int(np.floor(cy / dx)) - rows // 2
    c0 = int(np.floor(cx / dx)) - cols // 2
    rr = np.clip(np.arange(r0, r0 + rows), 0, n - 1)
    cc = np.clip(np.arange(c0, c0 + cols), 0, n - 1)
    return np.unique((rr[:, None] * n + cc[None, :]).ravel()).astype(np.int32)
