"""FailureZone activation on the device-authoritative state.

Counterpart of DeviceFailureZoneBridge (pkg/runtime/gpu_failure_bridge.py:44-212).
The reference downloads a physically sized region, runs the CPU CAD sweep /
intersection / FailureZone equations on it, and commits the result through
the entrainment contract plus a tool-forcing scatter into ``material_mask``
(:114-136).  The CAD geometry is host code outside this package: ``execute``
takes it as a callable that maps the downloaded patch to the activation
thickness and the affected mask, and keeps the reference's expansion loop,
donor limit and activation modes around it.  The commit itself -- entrainment
over the patch, the contact / 4-neighbourhood forcing rule, the ordered list
of forced cells and the moved volume -- is one libmv2 call
(``mv2_failure_activate``), between the dirty-surface capture and collection.
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass
from time import perf_counter
from typing import Any, Callable

import numpy as np

from . import _lib
from .bulk_state import DeviceBulkState, HostBulkStatePatch
from .config import MaterialScenario, TerrainGrid

FEE_FAILURE_ZONE = "FEE_FAILURE_ZONE"
GEOMETRIC_SWEEP_ONLY = "GEOMETRIC_SWEEP_ONLY_OUTSIDE_FEE_FORCE_DOMAIN"
NO_ACTIVATION = "NONE"


@dataclass(frozen=True)
class FailureGeometry:
    """What the host CAD geometry returns for one patch (the fields of the
    reference's FailureZone / ToolTerrainIntersection that the activation reads)."""

    active_thickness_m: np.ndarray
    affected_mask: np.ndarray
    applicability_status: str = "INSIDE_FEE_DOMAIN"
    penetration_depth_m: np.ndarray | None = None
    intersection: Any = None
    failure_zone: Any = None


@dataclass(frozen=True)
class DeviceFailureZoneResult:
    """Compact FailureZone result and direct resting->mobile transaction data (:25-41)."""

    patch_bbox_yx: tuple[int, int, int, int]
    requested_bbox_yx: tuple[int, int, int, int]
    expansion_count: int
    patch_cell_count: int
    patch_fraction_of_terrain: float
    intersection: Any
    failure_zone: Any
    activated_height_m: np.ndarray
    activated_volume_m3: float
    activation_tool_impulse_on_mobile_terrain_ns: np.ndarray
    activation_mode: str
    forcing_flat_indices: np.ndarray
    timings_ms: dict[str, float]
    dirty_tile_ids: np.ndarray


class DeviceFailureZoneBridge:
    """FailureZone activation committed by HIP kernels on the shared state."""

    backend_identity = "GPU_RUNTIME_COMPACT_HOST_FAILUREZONE_BRIDGE_HIP"

    def __init__(self, state: DeviceBulkState, material: MaterialScenario, grid: TerrainGrid, integrator=None, *,
                 maximum_wedge_length_m: float = 1.0) -> None:
        if state.grid is not grid:
            raise ValueError("[GpuFailureBridge] state/grid identity mismatch")
        if not np.isfinite(maximum_wedge_length_m) or maximum_wedge_length_m <= 0.0:
            raise ValueError("[GpuFailureBridge] maximum_wedge_length_m must be finite/positive")
        self.state, self.material, self.grid, self.integrator = state, material, grid, integrator
        self.maximum_wedge_length_m = float(maximum_wedge_length_m)

    # -- bbox helpers (:160-186) -------------------------------------------------------
    def _pads(self) -> tuple[int, int]:
        return (int(np.ceil(self.maximum_wedge_length_m / self.grid.dy)),
                int(np.ceil(self.maximum_wedge_length_m / self.grid.dx)))

    def physical_request_bbox(self, affected_bbox_grid) -> tuple[int, int, int, int]:
        """(row0, row1, col0, col1) of the sweep's (row0, col0, row1, col1) box
        grown by the maximum wedge run-out; a one-cell query without a sweep."""
        row0, col0, row1, col1 = (int(v) for v in affected_bbox_grid)
        if row1 <= row0 or col1 <= col0:
            return (0, 1, 0, 1)
        pr, pc = self._pads()
        return (max(0, row0 - pr), min(self.grid.ny, row1 + pr), max(0, col0 - pc), min(self.grid.nx, col1 + pc))

    def _expand(self, bbox):
        pr, pc = self._pads()
        return (max(0, bbox[0] - pr), min(self.grid.ny, bbox[1] + pr), max(0, bbox[2] - pc),
                min(self.grid.nx, bbox[3] + pc))

    def _at_global_boundary(self, bbox) -> bool:
        return bbox[0] == 0 or bbox[1] == self.grid.ny or bbox[2] == 0 or bbox[3] == self.grid.nx

    @staticmethod
    def _touches_patch_boundary(field) -> bool:
        active = np.asarray(field) > 1.0e-12
        return bool(active.any() and (active[0].any() or active[-1].any() or active[:, 0].any()
                                      or active[:, -1].any()))

    # -- the reference's execute loop around a host geometry callable (:68-154) ---------
    def execute(self, geometry: Callable[[HostBulkStatePatch, tuple], FailureGeometry],
                requested_bbox_yx) -> DeviceFailureZoneResult:
        requested = tuple(int(v) for v in requested_bbox_yx)
        bbox, expansions = requested, 0
        query_ms = geometry_ms = 0.0
        while True:
            t0 = perf_counter()
            patch = self.state.download_region(bbox, source="failure_zone_geometry")
            query_ms += (perf_counter() - t0) * 1e3
            t0 = perf_counter()
            geo = geometry(patch, bbox)
            geometry_ms += (perf_counter() - t0) * 1e3
            if not self._touches_patch_boundary(geo.active_thickness_m) or self._at_global_boundary(bbox):
                break
            grown = self._expand(bbox)
            if grown == bbox:
                break
            bbox, expansions = grown, expansions + 1
        available = patch.h_resting_derived_m
        activated = np.minimum(np.asarray(geo.active_thickness_m, dtype=np.float64), available)
        mode = FEE_FAILURE_ZONE
        if geo.applicability_status in {"OUTSIDE_FEE_DOMAIN", "PARTIAL_OUTSIDE_FEE_DOMAIN"}:
            pen = np.zeros_like(activated) if geo.penetration_depth_m is None else geo.penetration_depth_m
            activated = np.maximum(activated, np.minimum(pen, available))
            mode = GEOMETRIC_SWEEP_ONLY
        elif not np.any(activated > 0.0):
            mode = NO_ACTIVATION
        return self.activate(bbox, activated, geo.affected_mask, requested_bbox_yx=requested,
                             expansion_count=expansions, activation_mode=mode, intersection=geo.intersection,
                             failure_zone=geo.failure_zone,
                             timings_ms={"terrain_patch_query": query_ms,
                                         "bucket_geometry_and_failure_zone": geometry_ms})

    # -- the device commit (:114-136) ------------------------------------------------------
    def activate(self, bbox_yx, activated_height_m, affected_mask, *, requested_bbox_yx=None,
                 expansion_count: int = 0, activation_mode: str | None = None, intersection=None,
                 failure_zone=None, timings_ms: dict | None = None) -> DeviceFailureZoneResult:
        """Commit one activation patch: entrainment of ``activated_height_m``
        (donor-limited on the device), tool forcing on the contact cells and
        their 4-neighbourhood, dirty tiles of the changed surface."""
        torch = __import__("torch")
        row0, row1, col0, col1 = (int(v) for v in bbox_yx)
        if not (0 <= row0 < row1 <= self.grid.ny and 0 <= col0 < col1 <= self.grid.nx):
            raise ValueError("[GpuFailureBridge] activation bbox outside state")
        shape = (row1 - row0, col1 - col0)
        activated = np.ascontiguousarray(np.asarray(activated_height_m, dtype=np.float64))
        affected = np.ascontiguousarray(np.asarray(affected_mask, dtype=bool).astype(np.int32))
        if activated.shape != shape or affected.shape != shape or not np.all(np.isfinite(activated)):
            raise ValueError("[GpuFailureBridge] activation patch shape/finite check failed")
        state, rt = self.state, self.state.runtime
        t0 = perf_counter()
        state.capture_surface_for_dirty_tracking()
        act_d = state._upload_compact(activated.ravel(), torch.float64)
        aff_d = state._upload_compact(affected.ravel(), torch.int32)
        forcing_d = torch.empty(activated.size, dtype=torch.int32, device=rt.device)
        res = _lib.ActivationResult()
        rt.bind_stream()
        rt.check(rt.lib.mv2_failure_activate(rt.ctx, row0, row1, col0, col1, act_d.data_ptr(), aff_d.data_ptr(),
                                             rt.ptr("material_mask"), forcing_d.data_ptr(), ctypes.byref(res)),
                 value_error=ValueError)
        rt.telemetry.record_call(launches=2, syncs=1)
        count = int(res.forcing_count)
        forcing = forcing_d[:count].cpu().numpy().astype(np.int32) if count else np.empty(0, np.int32)
        rt.telemetry.record_d2h(forcing)
        state.mark_dirty_bbox((row0, row1, col0, col1))
        dirty = state.collect_surface_dirty_tiles()
        timings = dict(timings_ms or {})
        timings["activation_scatter"] = (perf_counter() - t0) * 1e3
        mode = activation_mode or (FEE_FAILURE_ZONE if np.any(activated > 0.0) else NO_ACTIVATION)
        return DeviceFailureZoneResult(
            patch_bbox_yx=(row0, row1, col0, col1),
            requested_bbox_yx=tuple(requested_bbox_yx) if requested_bbox_yx is not None else (row0, row1, col0, col1),
            expansion_count=int(expansion_count), patch_cell_count=int(activated.size),
            patch_fraction_of_terrain=float(activated.size / state.size), intersection=intersection,
            failure_zone=failure_zone, activated_height_m=activated, activated_volume_m3=float(res.activated_volume),
            activation_tool_impulse_on_mobile_terrain_ns=np.zeros(3), activation_mode=mode,
            forcing_flat_indices=forcing, timings_ms=timings, dirty_tile_ids=dirty)

    def clear_tool_forcing(self, flat_indices) -> None:
        """material_mask = 0 on the forced cells after the mobile step (:155-158)."""
        self.state.apply_host_indices("material_mask", flat_indices, 0,
                                      reason="failure_zone_tool_forcing_complete")
