"""Inputs of the reference's 41x41 Mobile V2 contract suite.

Restated from tools/run_mobile_v2_reference_contracts.py:78-134 (shapes,
fields, dt per case); the expected outputs live in
tests/golden/mobile_v2_cpu_contracts.json.
"""

from __future__ import annotations

import numpy as np

DX = 0.05


def _state(b, h, velocity=None):
    v = np.zeros(h.shape + (2,)) if velocity is None else np.broadcast_to(velocity, h.shape + (2,)).copy()
    q = h[..., None] * v
    return np.asarray(b, dtype=np.float64), np.asarray(h, dtype=np.float64), q


def cases():
    """Return {name: (b, h, q(ny,nx,2), dt)} for the seven dynamic/static cases."""
    shape = (41, 41)
    y, x = np.indices(shape)
    out = {}
    out["STATIC_FLAT"] = (*_state(np.ones(shape), np.full(shape, 0.1)), 0.05)
    free = 2.0 - np.tan(np.deg2rad(10.0)) * x * DX
    out["STATIC_SLOPE"] = (*_state(free - 0.08, np.full(shape, 0.08)), 0.05)
    H = np.full(shape, 0.8266754619)
    H[:, 19] = 1.82
    H[:, 20] = 0.5144415194
    H[:, 21] = 0.8266754619
    b = np.full(shape, 0.12)
    h = np.maximum(H - b, 0.0)
    velocity = np.zeros(shape + (2,))
    velocity[:, 20, 0] = 0.8
    out["VALLEY"] = (b, h, h[..., None] * velocity, 1.0 / 240.0)
    mound_h = 0.03 + 0.30 * np.exp(-((x - 20) ** 2 + (y - 20) ** 2) / 18.0)
    out["MOUND"] = (*_state(np.full(shape, 0.12), mound_h), 0.02)
    packet_h = 0.20 * np.exp(-((x - 13) ** 2 + (y - 20) ** 2) / 30.0)
    packet_v = np.zeros(shape + (2,))
    packet_v[..., 0] = 0.8
    out["SMOOTH_PACKET"] = (*_state(1.5 - 0.08 * x * DX, packet_h, packet_v), 0.04)
    opposing_h = 0.12 * (np.exp(-((x - 12) ** 2 + (y - 20) ** 2) / 20.0)
                         + np.exp(-((x - 28) ** 2 + (y - 20) ** 2) / 20.0))
    opposing_v = np.zeros(shape + (2,))
    opposing_v[..., 0] = np.where(x < 20, 0.8, -0.8)
    out["OPPOSING_PACKETS"] = (*_state(np.full(shape, 0.12), opposing_h, opposing_v), 0.04)
    wet_h = np.zeros(shape)
    wet_h[12:29, 12:29] = 0.12
    wet_v = np.zeros(shape + (2,))
    wet_v[..., 0] = 0.5
    out["WET_DRY"] = (*_state(np.full(shape, 0.12), wet_h, wet_v), 0.04)
    return out
