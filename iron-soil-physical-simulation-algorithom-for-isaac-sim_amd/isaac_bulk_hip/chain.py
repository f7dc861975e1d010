"""Shared-device execution chain of the hot-path operators.

Counterpart of GpuBulkOperatorChain (pkg/runtime/gpu_bulk_operator_chain.py:53-296)
restricted to the Mobile V2 hot path: the production V2 solver, deposition,
the compact-frontier sweep, FailureZone activation and the LargeAvalanche
detector, all bound to one DeviceBulkState.  Each call keeps the reference's
order -- surface capture, operator, dirty-tile collection -- and records its
wall time and transfer snapshot.  Track-soil, bucket intake and airborne
parcels are outside the hot path and not part of this chain.
"""

from __future__ import annotations

from dataclasses import dataclass
from time import perf_counter

from .bulk_state import DeviceBulkState
from .config import LargeAvalancheTransitionConfig, MaterialScenario, MobileV2Config, TerrainGrid
from .deposition import WarpDepositionOperator
from .failure_zone import DeviceFailureZoneBridge, DeviceFailureZoneResult
from .large_avalanche import DeviceLargeAvalancheBridge
from .mobile_v2 import WarpProductionMobileV2Solver


@dataclass(frozen=True)
class GpuBulkOperatorStep:
    module: str
    wall_time_ms: float
    transfer: dict
    dirty_tile_count: int


class GpuBulkOperatorChain:
    """One-device bridge for the hot-path terrain operators."""

    backend_identity = "GPU_RUNTIME_SHARED_DEVICE_OPERATOR_CHAIN_HIP"

    def __init__(self, state: DeviceBulkState, material: MaterialScenario, grid: TerrainGrid, integrator,
                 large_avalanche_config: LargeAvalancheTransitionConfig | None = None, *,
                 maximum_wedge_length_m: float = 1.0) -> None:
        if state.grid is not grid:
            raise ValueError("[GpuBulkChain] state/grid identity mismatch")
        self.state, self.material, self.grid, self.integrator = state, material, grid, integrator
        # the production binding of the V2 solver (:73-82)
        self.mobile = WarpProductionMobileV2Solver(
            runtime=state.runtime,
            config=MobileV2Config(dx_m=grid.dx, dy_m=grid.dy, start_angle_deg=material.start_angle_deg,
                                  stop_angle_deg=material.stop_angle_deg,
                                  basal_friction_coefficient=material.mobile_friction_coefficient))
        self.mobile.bind_device_state(state, material, grid, integrator)
        self.deposition = WarpDepositionOperator(runtime=state.runtime)
        self.deposition.bind_device_state(state)
        self.failure_zone = DeviceFailureZoneBridge(state, material, grid, integrator,
                                                    maximum_wedge_length_m=maximum_wedge_length_m)
        self.large_avalanche = DeviceLargeAvalancheBridge(state, material, grid, integrator,
                                                          large_avalanche_config or LargeAvalancheTransitionConfig())
        self._frontier = None
        self._records: list[GpuBulkOperatorStep] = []

    @property
    def frontier(self):
        """The compact-frontier operator bound to the state (FP64 states only)."""
        if self._frontier is None:
            from .relaxation import WarpCompactActiveEdgeOperator

            op = WarpCompactActiveEdgeOperator(self.grid.shape, runtime=self.state.runtime,
                                               tile_size=self.state.tile_size)
            op.bind_device_state(self.state)
            self._frontier = op
        return self._frontier

    @property
    def records(self) -> list[GpuBulkOperatorStep]:
        return list(self._records)

    def _record(self, module: str, start: float, dirty_tile_count: int) -> None:
        self._records.append(GpuBulkOperatorStep(module, (perf_counter() - start) * 1e3,
                                                 self.state.transfer_snapshot().to_dict(), int(dirty_tile_count)))

    # -- operators (:131-270) ------------------------------------------------------------
    def step_mobile(self, dt_s: float):
        start = perf_counter()
        self.state.capture_surface_for_dirty_tracking()
        result = self.mobile.step_resident(dt_s)
        self.state.advance_time(dt_s)
        dirty = self.state.collect_surface_dirty_tiles()
        self._record("Mobile", start, dirty.size)
        return result

    def apply_failure_zone(self, geometry, requested_bbox_yx) -> DeviceFailureZoneResult:
        """FailureZone activation with the host CAD geometry as a callable."""
        start = perf_counter()
        result = self.failure_zone.execute(geometry, requested_bbox_yx)
        self._record("FailureZoneActivation", start, 0)
        return result

    def activate_failure_zone(self, bbox_yx, activated_height_m, affected_mask) -> DeviceFailureZoneResult:
        """FailureZone activation of an already computed patch."""
        start = perf_counter()
        result = self.failure_zone.activate(bbox_yx, activated_height_m, affected_mask)
        self._record("FailureZoneActivation", start, 0)
        return result

    def clear_failure_tool_forcing(self, result: DeviceFailureZoneResult) -> None:
        self.failure_zone.clear_tool_forcing(result.forcing_flat_indices)

    def step_deposition(self, dt_s: float, *, settle_subcell_tail: bool = False):
        start = perf_counter()
        self.state.capture_surface_for_dirty_tracking()
        result = self.deposition.step_resident(self.material, dt_s, settle_subcell_tail=settle_subcell_tail)
        dirty = self.state.collect_surface_dirty_tiles()
        self._record("Deposition", start, dirty.size)
        return result

    def advance_large_avalanche(self, dt_s: float, *, release_settled_latches: bool = True):
        start = perf_counter()
        result = self.large_avalanche.observe_and_maybe_mobilize(dt_s, release_settled_latches=release_settled_latches)
        self._record("LargeAvalancheDeviceTransition", start, 0)
        return result

    def relax_frontier_phase(self, batch, *, critical_difference_m: float, tolerance_m: float):
        start = perf_counter()
        self.state.capture_surface_for_dirty_tracking()
        result = self.frontier.run_phase(batch, critical_difference_m=critical_difference_m, tolerance_m=tolerance_m)
        dirty = self.state.collect_surface_dirty_tiles()
        self._record("MiniSlopeCompactFrontier", start, dirty.size)
        return result

    def transfer_frontier_phase(self, batch, *, critical_difference_m: float):
        start = perf_counter()
        self.state.capture_surface_for_dirty_tracking()
        result = self.frontier.relax_phase(batch, critical_difference_m=critical_difference_m)
        dirty = self.state.collect_surface_dirty_tiles()
        self._record("MiniSlopeCompactFrontier", start, dirty.size)
        return result

    def scan_frontier_phase(self, batch, *, critical_difference_m: float, tolerance_m: float):
        start = perf_counter()
        result = self.frontier.scan_phase(batch, critical_difference_m=critical_difference_m, tolerance_m=tolerance_m)
        self._record("MiniSlopeCompactFrontier", start, 0)
        return result

    def diagnostics(self) -> dict:
        return {
            "backend_identity": self.backend_identity,
            "state_authority": self.state.authority.value,
            "state": self.state.diagnostics(),
            "operators": {"mobile": self.mobile.diagnostics(), "deposition": self.deposition.diagnostics(),
                          "failure_zone": {"backend": self.failure_zone.backend_identity},
                          "large_avalanche": self.large_avalanche.diagnostics()},
            "records": [{"module": r.module, "wall_time_ms": r.wall_time_ms, "transfer": r.transfer,
                         "dirty_tile_count": r.dirty_tile_count} for r in self._records],
        }
