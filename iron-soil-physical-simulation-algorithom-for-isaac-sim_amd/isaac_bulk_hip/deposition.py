"""Device-resident Mobile -> Resting deposition (drop-in for WarpDepositionOperator).

Mirrors pkg/bulk_interaction/warp_deposition.py:136-246: the same constructor
arguments and checks, ``bind_device_state`` over the shared DeviceBulkState
arrays (creating ``deposition_removed_momentum_x/y`` when absent), and
``step_resident(material, dt_s, settle_subcell_tail=...)`` returning a
``WarpDepositionStep``.  The step runs as two HIP kernels through
``mv2_deposition_step``: decisions from the pre-call fields, then the in-place
transfer; the three weighted sums come back after one stream sync, as in the
reference.
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

from . import _lib
from .config import MaterialScenario
from .runtime import HipRuntime

PARAMETER_STATUS = "LITERATURE_BASED_REDUCED_ORDER_UNCALIBRATED"
GRAVITY_M_S2 = 9.81


@dataclass(frozen=True)
class WarpDepositionStep:
    deposited_volume_m3: float
    parameter_status: str
    removed_mobile_momentum_terrain_kg_m_s: np.ndarray


class WarpDepositionOperator:
    """Shared-state incremental deposition below the cohesive stop yield."""

    backend_name = "GPU_HIP_INCREMENTAL_DEPOSITION"

    def __init__(self, *, settling_rate_m_s: float = 0.30, speed_threshold_m_s: float = 0.18,
                 device: str = "cuda:0", runtime: HipRuntime | None = None) -> None:
        if settling_rate_m_s <= 0.0 or speed_threshold_m_s < 0.0:
            raise ValueError("[WarpDeposition] invalid configuration")
        self.settling_rate_m_s = float(settling_rate_m_s)
        self.speed_threshold_m_s = float(speed_threshold_m_s)
        self.runtime = runtime or HipRuntime(device)
        self._state = None

    def bind_device_state(self, state) -> None:
        if getattr(state, "runtime", None) is not self.runtime:
            raise ValueError("[WarpDeposition] state/runtime ownership mismatch")
        required = {"b_eff", "mobile", "momentum_x", "momentum_y", "material_mask",
                    "deposition_work", "weights", "avalanche_m2r_cumulative"}
        missing = required - set(self.runtime.arrays)
        if missing:
            raise ValueError(f"[WarpDeposition] missing shared fields: {sorted(missing)}")
        if "deposition_removed_momentum_x" not in self.runtime.arrays:
            dtype = self.runtime.arrays["momentum_x"].dtype
            self.runtime.zeros("deposition_removed_momentum_x", state.size, dtype)
            self.runtime.zeros("deposition_removed_momentum_y", state.size, dtype)
        self._state = state

    def step_resident(self, material: MaterialScenario, dt_s: float, *,
                      settle_subcell_tail: bool = False) -> WarpDepositionStep:
        if self._state is None:
            raise RuntimeError("[WarpDeposition] bind DeviceBulkState first")
        dt = float(dt_s)
        if not np.isfinite(dt) or dt <= 0.0:
            raise ValueError("[WarpDeposition] dt_s must be finite/positive")
        rt = self.runtime
        arrays = rt.arrays
        cfg = _lib.DepositionConfig(
            dt, self.settling_rate_m_s, self.speed_threshold_m_s,
            float(material.assumed_bulk_density_kg_m3), GRAVITY_M_S2,
            float(material.cohesion_proxy_pa), float(np.tan(np.deg2rad(material.stop_angle_deg))),
            int(bool(settle_subcell_tail)), 1)
        out = np.zeros(3)
        rt.bind_stream()
        rc = rt.lib.mv2_deposition_step(
            rt.ctx, ctypes.byref(cfg), rt.ptr("material_mask"), rt.ptr("deposition_work"),
            rt.ptr("deposition_removed_momentum_x"), rt.ptr("deposition_removed_momentum_y"),
            rt.ptr("avalanche_m2r_cumulative"), out.ctypes.data)
        rt.check(rc)
        rt.telemetry.record_call(launches=2)
        rt.telemetry.record_d2h(out)
        return WarpDepositionStep(
            deposited_volume_m3=float(out[0]),
            parameter_status=PARAMETER_STATUS,
            removed_mobile_momentum_terrain_kg_m_s=np.asarray([out[1], out[2], 0.0]),
        )

    def diagnostics(self) -> dict[str, object]:
        result = self.runtime.diagnostics(self.backend_name)
        result.update({
            "resident_state": ["b_eff", "mobile", "momentum_x", "momentum_y"],
            "semantics": "incremental_mobile_to_resting_below_cohesive_stop_yield",
            "momentum_removal": "EXPLICITLY_REPORTED",
            "decision_snapshot": "PRE_CALL_FIELDS",
        })
        return result
