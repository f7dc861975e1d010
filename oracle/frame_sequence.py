"""FP64 oracle of the full-size frame (test infrastructure only).

TEST INFRASTRUCTURE ONLY: used by ``tests/`` and by the ``cpu_baseline`` leg
of ``bench.py`` as the checker of the FP32 device frame -- never by the
product package.

One frame, as the reference's GPU core runs its pieces
(pkg/runtime/v2_physics_core.py:708-1105, restricted to this workload):

1. the Mobile V2 step of ``dt`` (``MobileV2ReferenceSolver.step``,
   pkg/experimental/mobile_v2_reference.py:191-314, restated in
   ``mv2_oracle.c:orc_v2_step``; the production loop-limit semantics of
   pkg/bulk_interaction/warp_mobile_v2.py:181-216);
2. the plain angle-of-repose relaxation of b_eff
   (``relax_critical_slope``, slope_model.py:273-343, behind
   ``MinimumSlopeAdapter``'s closed boundary,
   pkg/solvers/minimum_slope_adapter.py:153-182);
3. the LargeAvalanche failure mask (``initialize``,
   pkg/runtime/gpu_large_avalanche.py:59-162) and its canonical 8-connected
   labels (root = minimum linear index, :165-198).

``field_errors`` gives the comparison numbers of a device frame against the
oracle's: L-inf and RMSE of h, qx, qy, b_eff and H_free.
"""

from __future__ import annotations

import numpy as np

import oracle

QUANTUM_M = 2.0 ** -24


class OracleFrameSequence:
    """The FP64 frame sequence from given initial fields (b_eff, h; q = 0)."""

    def __init__(self, b, h, *, dx: float, dt: float, v2: dict, relax_angle_deg: float,
                 relax_max_iterations: int, relax_tolerance: float, avalanche: dict,
                 connectivity: int = 8) -> None:
        self.b = np.array(b, dtype=np.float64)
        self.h = np.array(h, dtype=np.float64)
        self.qx = np.zeros_like(self.h)
        self.qy = np.zeros_like(self.h)
        self.dx, self.dt = float(dx), float(dt)
        self.v2 = dict(v2)
        self.crit = float(np.tan(np.deg2rad(relax_angle_deg)))
        self.relax_max_iterations = int(relax_max_iterations)
        self.relax_tolerance = float(relax_tolerance)
        self.avalanche = dict(avalanche)
        self.connectivity = int(connectivity)
        self.latch = np.zeros(self.h.shape, np.int32)
        self.mass0 = float(np.sum(self.h))
        self.total0 = float(np.sum(self.b) + np.sum(self.h))
        self.v2_seconds = 0.0

    def mask(self, b, h, qx, qy, latch=None):
        """Failure mask of given fields (z_base 0, no export history)."""
        z = np.zeros_like(b)
        w = oracle.vertex_weights_fast(b.shape[0], b.shape[1], self.dx, self.dx)
        lat = np.zeros(b.shape, np.int32) if latch is None else latch
        return oracle.avalanche_initialize(z, b, h, qx, qy, w, z, z, lat, dx=self.dx, dy=self.dx,
                                           **self.avalanche)

    def step(self) -> dict:
        import time

        t0 = time.perf_counter()
        self.h, self.qx, self.qy, rec = oracle.v2_step(self.b, self.h, self.qx, self.qy, self.dt,
                                                       dx=self.dx, dy=self.dx, **self.v2)
        self.v2_seconds += time.perf_counter() - t0
        if rec["status"] != 0:
            raise RuntimeError(f"oracle V2 step failed with status {rec['status']}")
        self.b, it, conv, last = oracle.relax_yx(self.b, self.dx, self.dx, self.crit,
                                                 self.relax_max_iterations, self.relax_tolerance)
        m = self.mask(self.b, self.h, self.qx, self.qy, self.latch)
        self.latch = m["latch"]
        labels = oracle.label(m["unstable"], self.connectivity)
        return {"substeps": rec["substeps"], "relax_iterations": it, "relax_converged": conv,
                "unstable": m["unstable"], "labels": labels}

    def mass_error(self) -> float:
        """|M_t - M_0| / M_0 of the mobile layer (uniform cell area)."""
        return abs(float(np.sum(self.h)) - self.mass0) / self.mass0

    def total_error(self) -> float:
        """|V_t - V_0| / V_0 of b_eff + h (what the relaxation must also keep)."""
        return abs(float(np.sum(self.b) + np.sum(self.h)) - self.total0) / self.total0


def field_errors(dev: dict, ref: dict) -> dict:
    """L-inf and RMSE of every named field (dicts of equal-shape arrays)."""
    out = {}
    for name in ref:
        e = np.asarray(dev[name], dtype=np.float64) - np.asarray(ref[name], dtype=np.float64)
        out[name] = {"linf": float(np.max(np.abs(e))), "rmse": float(np.sqrt(np.mean(e * e)))}
    return out
