"""Seeded inputs of the Mobile -> Resting deposition step (shared by the CPU
and GPU tests) and an independent NumPy statement of its decision rule."""

from __future__ import annotations

import numpy as np

# material of the reference's device-state tests (tests/test_device_bulk_state.py:35-36)
DENSITY = 1370.0
COHESION = 800.0
STOP_DEG = 30.0
GRAVITY = 9.81
RATE = 0.30          # WarpDepositionOperator defaults (warp_deposition.py:151-152)
THRESHOLD = 0.18


def fields(ny, nx, seed, *, dx=0.05, dy=0.05, wet=0.7, forcing_fraction=0.05):
    """Rough surface with slopes of order one, a partial mobile layer with
    speeds on both sides of the threshold, a sparse forcing mask and a
    non-zero cumulative field."""
    rng = np.random.default_rng(seed)
    y = np.arange(ny)[:, None] * dy
    x = np.arange(nx)[None, :] * dx
    b = 0.5 + 0.4 * np.sin(3.0 * x + 1.7 * y) + 0.3 * np.cos(2.3 * y - 0.4 * x)
    b = b + rng.uniform(0.0, 0.08, (ny, nx))
    h = np.where(rng.random((ny, nx)) < wet, rng.uniform(1e-4, 0.06, (ny, nx)), 0.0)
    speed = rng.uniform(0.0, 0.3, (ny, nx))
    angle = rng.uniform(0.0, 2.0 * np.pi, (ny, nx))
    qx = h * speed * np.cos(angle)
    qy = h * speed * np.sin(angle)
    forcing = (rng.random((ny, nx)) < forcing_fraction).astype(np.int32)
    cumulative = rng.uniform(0.0, 1e-3, (ny, nx))
    return b, h, qx, qy, forcing, cumulative


def numpy_rule(b, h, qx, qy, forcing, *, dx, dy, dt, cohesion=COHESION, tan_stop=None,
               density=DENSITY, gravity=GRAVITY, rate=RATE, threshold=THRESHOLD):
    """Deposited amount per cell from the physical statement of the rule:
    free surface H = b + h, its gradient by np.gradient (central inside,
    one-sided at the edges), drive rho g h sin(theta) against cohesion +
    rho g h cos(theta) tan(stop).  Returns (amount, knife) where knife marks
    cells whose decision lies within rounding of the threshold."""
    if tan_stop is None:
        tan_stop = float(np.tan(np.deg2rad(STOP_DEG)))
    free = b + h
    gy, gx = np.gradient(free, dy, dx, edge_order=1)
    theta = np.arctan(np.hypot(gx, gy))
    drive = density * gravity * h * np.sin(theta)
    resist = cohesion + density * gravity * h * np.cos(theta) * tan_stop
    safe = np.where(h > 0.0, h, 1.0)
    speed = np.hypot(qx / safe, qy / safe)
    wet = (h > 0.0) & (forcing == 0)
    eligible = wet & (speed <= threshold) & (drive <= resist)
    knife = wet & ((np.abs(resist - drive) <= 1e-9 * (np.abs(resist) + 1.0))
                   | (np.abs(speed - threshold) <= 1e-12))
    amount = np.where(eligible, np.minimum(h, rate * dt), 0.0)
    return amount, knife
