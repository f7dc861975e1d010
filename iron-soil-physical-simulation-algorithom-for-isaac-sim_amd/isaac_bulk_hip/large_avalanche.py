"""LargeAvalanche failure detection, connected components and activation on the MI355X.

Counterpart of DeviceLargeAvalancheBridge (pkg/runtime/gpu_large_avalanche.py:469-901):
the cohesive start/stop failure mask, canonical connected-component labels
(root = minimum linear index; tie between equally large components -> the
smallest root) and the zero-momentum Resting->Mobile activation of the winning
component run as HIP kernels (libmv2 ``mv2_avalanche_label`` /
``mv2_avalanche_mobilize``).  The host keeps the same persistence logic and
classification (:645-688) and reads fixed-size records only.
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

from . import _lib
from .bulk_state import DeviceBulkState
from .config import LargeAvalancheTransitionConfig, MaterialScenario, TerrainGrid


@dataclass(frozen=True)
class DeviceLargeAvalancheResult:
    classification: str
    transitioned: bool
    unstable_cell_count: int
    connected_region_count: int
    largest_connected_cell_count: int
    largest_connected_area_m2: float
    largest_connected_mobilizable_volume_m3: float
    mean_excess_start_deg: float
    maximum_excess_start_deg: float
    persistence_s: float
    transferred_volume_m3: float
    gravity_initiation_impulse_kg_m_s: np.ndarray
    residual_seed_tile_ids: np.ndarray
    union_iterations: int
    parameter_basis: str
    sensitivity_case: str
    newly_activated_volume_m3: float
    reactivated_volume_m3: float
    newly_activated_area_m2: float
    reactivated_area_step_m2: float
    cumulative_resting_to_mobile_m3: float
    cumulative_mobile_to_resting_m3: float
    unique_cells_ever_activated: int
    cells_activated_more_than_once: int
    unique_activated_area_m2: float
    reactivated_area_m2: float
    retired_candidate_area_m2: float
    current_mobile_volume_m3: float
    moving_mobile_volume_m3: float
    maximum_mobile_speed_m_s: float
    active_tile_count: int
    net_terrain_volume_change_m3: float
    net_spatial_transfer_m3: float
    mean_activated_slope_deg: float
    maximum_activated_slope_deg: float
    mean_slope_minus_theta_stop_deg: float
    mean_theta_start_minus_slope_deg: float


def avalanche_config(material: MaterialScenario, config: LargeAvalancheTransitionConfig, *,
                     release_settled_latches: bool) -> _lib.AvalancheConfig:
    start_rad = float(np.deg2rad(material.start_angle_deg))
    stop_rad = float(np.deg2rad(material.stop_angle_deg))
    return _lib.AvalancheConfig(
        tan_start=float(np.tan(start_rad)), tan_stop=float(np.tan(stop_rad)), start_rad=start_rad,
        stop_rad=stop_rad, density=float(material.assumed_bulk_density_kg_m3), gravity=9.81,
        cohesion=float(material.cohesion_proxy_pa), mobilization_depth=config.mobilization_depth_m,
        dry_tolerance=config.dry_tolerance_m, activity_speed=config.mobile_activity_speed_m_s,
        connectivity=int(config.connectivity), release_settled_latches=int(bool(release_settled_latches)),
        weight_mode=1, reserved=0)


def avalanche_fields(state: DeviceBulkState) -> _lib.AvalancheFields:
    a = state.runtime.arrays
    names = {
        "unstable": "avalanche_unstable", "labels": "avalanche_parent", "latch": "avalanche_latch",
        "previous": "avalanche_previous_component", "activation_count": "avalanche_activation_count",
        "counts": "avalanche_component_count", "owned_surface": "avalanche_owned_surface",
        "owned_export_baseline": "avalanche_owned_export_baseline",
        "export_cumulative": "mobile_export_cumulative", "slope": "avalanche_slope",
        "grad_x": "avalanche_gradient_x", "grad_y": "avalanche_gradient_y",
        "severity": "avalanche_severity", "first_activation_time": "avalanche_first_activation_time",
        "last_activation_time": "avalanche_last_activation_time", "r2m_cumulative": "avalanche_r2m_cumulative",
    }
    return _lib.AvalancheFields(**{k: int(a[v].data_ptr()) for k, v in names.items()})


class DeviceLargeAvalancheBridge:
    """Exact-cell HIP connected components with scalar host diagnostics."""

    backend_identity = "GPU_RUNTIME_DEVICE_CONNECTED_LARGE_AVALANCHE_HIP"
    NO_LARGE_EVENT = "NO_LARGE_EVENT"
    LOCAL_STATIC_INSTABILITY = "LOCAL_STATIC_INSTABILITY"
    PERSISTING_LARGE_UNSTABLE_REGION = "PERSISTING_LARGE_UNSTABLE_REGION"
    LARGE_AVALANCHE_MOBILE_PATH = "LARGE_AVALANCHE_MOBILE_PATH"

    def __init__(self, state: DeviceBulkState, material: MaterialScenario, grid: TerrainGrid,
                 integrator, config: LargeAvalancheTransitionConfig) -> None:
        if state.grid is not grid or integrator.shape != grid.shape:
            raise ValueError("[GpuLargeAvalanche] state/grid/integrator mismatch")
        self.state, self.material, self.grid = state, material, grid
        self.integrator, self.config = integrator, config
        self._persistence_s = 0.0
        self._has_previous = False
        self.last_label: _lib.LabelResult | None = None

    def reset(self) -> None:
        self._persistence_s = 0.0
        self._has_previous = False

    def _structs(self, release: bool):
        """The config and field-pointer structs (rebuilt when an array moved)."""
        a = self.state.runtime.arrays
        key = (bool(release), id(self.material), id(self.config), a["avalanche_parent"].data_ptr(),
               a["avalanche_unstable"].data_ptr(),
               a["avalanche_slope"].data_ptr(), a["avalanche_latch"].data_ptr())
        cached = getattr(self, "_struct_cache", None)
        if cached is None or cached[0] != key:
            cached = (key, avalanche_config(self.material, self.config, release_settled_latches=release),
                      avalanche_fields(self.state))
            self._struct_cache = cached
        return cached[1], cached[2]

    def label(self, *, release_settled_latches: bool = True) -> _lib.LabelResult:
        """Failure mask + canonical labels + winner statistics (device part of observe)."""
        rt = self.state.runtime
        rt.bind_stream()
        cfg, fields = self._structs(release_settled_latches)
        out = _lib.LabelResult()
        rt.check(rt.lib.mv2_avalanche_label(rt.ctx, ctypes.byref(cfg), ctypes.byref(fields),
                                            ctypes.byref(out)))
        self.last_label = out
        return out

    def observe_and_maybe_mobilize(self, dt_s: float, *,
                                   release_settled_latches: bool = True) -> DeviceLargeAvalancheResult:
        dt = float(dt_s)
        if not np.isfinite(dt) or dt <= 0.0:
            raise ValueError("[GpuLargeAvalanche] dt_s must be finite/positive")
        state = self.state
        rt = state.runtime
        res = self.label(release_settled_latches=release_settled_latches)
        diag_i = np.asarray(res.diag_i[:])
        diag_f = np.asarray(res.diag_f[:])
        largest = int(diag_i[1])
        area = float(diag_f[4])
        mobilizable = float(diag_f[5])
        mean_excess_deg = float(np.rad2deg(diag_f[0] / largest)) if largest else 0.0
        max_excess_deg = float(np.rad2deg(diag_f[1]))
        cfgc = self.config
        extent_pass = largest >= cfgc.minimum_connected_cells and area >= cfgc.minimum_connected_area_m2
        volume_pass = mobilizable >= cfgc.minimum_mobilizable_volume_m3
        excess_pass = mean_excess_deg >= cfgc.minimum_mean_excess_start_deg
        candidate = extent_pass and volume_pass and excess_pass
        overlap = int(diag_i[4]) > 0
        if candidate:
            self._persistence_s = (self._persistence_s + dt if overlap or not self._has_previous else dt)
        else:
            self._persistence_s = 0.0
        ready = candidate and self._persistence_s >= cfgc.persistence_time_s
        if ready:
            classification = self.LARGE_AVALANCHE_MOBILE_PATH
        elif candidate:
            classification = self.PERSISTING_LARGE_UNSTABLE_REGION
        elif int(diag_i[3]) > 0:
            classification = self.LOCAL_STATIC_INSTABILITY
        else:
            classification = self.NO_LARGE_EVENT
        fields = avalanche_fields(state)
        rt.check(rt.lib.mv2_avalanche_previous(
            rt.ctx, ctypes.byref(fields), int(diag_i[2]), int(candidate),
            int(classification == self.LOCAL_STATIC_INSTABILITY), rt.ptr("frontier_reached"),
            rt.ptr("dirty_tile_flags"), state.tile_size, state.tile_shape[1], state.tile_count))
        self._has_previous = candidate
        transferred = 0.0
        impulse = np.zeros(2)
        mob = np.zeros(16)
        if ready:
            cfg = avalanche_config(self.material, cfgc, release_settled_latches=release_settled_latches)
            rt.check(rt.lib.mv2_avalanche_mobilize(rt.ctx, ctypes.byref(cfg), ctypes.byref(fields),
                                                   int(diag_i[2]), float(state.timestamp_device_s),
                                                   mob.ctypes.data))
            transferred = float(mob[7])
        flags = rt.arrays["dirty_tile_flags"].cpu().numpy()
        seed_tiles = np.flatnonzero(flags).astype(np.int32)
        p = self._persistence_summary()
        return DeviceLargeAvalancheResult(
            classification=classification, transitioned=transferred > 0.0,
            unstable_cell_count=int(diag_i[3]), connected_region_count=int(diag_i[0]),
            largest_connected_cell_count=largest, largest_connected_area_m2=area,
            largest_connected_mobilizable_volume_m3=mobilizable, mean_excess_start_deg=mean_excess_deg,
            maximum_excess_start_deg=max_excess_deg, persistence_s=self._persistence_s,
            transferred_volume_m3=transferred, gravity_initiation_impulse_kg_m_s=impulse,
            residual_seed_tile_ids=seed_tiles, union_iterations=int(res.union_passes),
            parameter_basis=cfgc.parameter_basis, sensitivity_case=cfgc.sensitivity_case,
            newly_activated_volume_m3=float(mob[10]), reactivated_volume_m3=float(mob[11]),
            newly_activated_area_m2=float(mob[12]), reactivated_area_step_m2=float(mob[13]), **p)

    def _persistence_summary(self) -> dict:
        """Lifecycle scalars of reduce_persistence (gpu_large_avalanche.py:378-451,
        :769-839), reduced on the device in one pass (mv2_avalanche_persistence)."""
        st = self.state
        rt = st.runtime
        cfg = avalanche_config(self.material, self.config, release_settled_latches=True)
        fields = avalanche_fields(st)
        di = np.zeros(4, dtype=np.int32)
        df = np.zeros(14)
        tiles = ctypes.c_int32(0)
        rt.bind_stream()
        rt.check(rt.lib.mv2_avalanche_persistence(
            rt.ctx, ctypes.byref(cfg), ctypes.byref(fields), rt.ptr("initial_b_eff"),
            rt.ptr("avalanche_m2r_cumulative"), st.tile_size, st.tile_shape[1], st.tile_count,
            rt.ptr("avalanche_activity_tile_flags"), di.ctypes.data, df.ctypes.data, ctypes.byref(tiles)))
        rt.telemetry.record_d2h(di)
        rt.telemetry.record_d2h(df)
        unique = int(di[0])

        def deg(x):
            return float(np.rad2deg(x))
        return {
            "cumulative_resting_to_mobile_m3": float(df[0]),
            "cumulative_mobile_to_resting_m3": float(df[1]),
            "unique_cells_ever_activated": unique,
            "cells_activated_more_than_once": int(di[1]),
            "unique_activated_area_m2": float(df[2]),
            "reactivated_area_m2": float(df[3]),
            "retired_candidate_area_m2": float(df[12]),
            "current_mobile_volume_m3": float(df[4]),
            "moving_mobile_volume_m3": float(df[13]),
            "maximum_mobile_speed_m_s": float(df[5]),
            "active_tile_count": int(tiles.value),
            "net_terrain_volume_change_m3": float(df[6]),
            "net_spatial_transfer_m3": float(df[7]),
            "mean_activated_slope_deg": deg(df[8] / unique) if unique else 0.0,
            "maximum_activated_slope_deg": deg(df[9]),
            "mean_slope_minus_theta_stop_deg": deg(df[10] / unique) if unique else 0.0,
            "mean_theta_start_minus_slope_deg": deg(df[11] / unique) if unique else 0.0,
        }

    def diagnostics(self) -> dict:
        return {
            "backend_identity": self.backend_identity,
            "field_authority": "DEVICE",
            "host_transfer": "FIXED_SIZE_SCALARS_AND_TILE_FLAGS_ONLY",
            "labelling": "SINGLE_PASS_LOCK_FREE_UNION_FIND_MIN_INDEX_ROOTS",
            "parameter_basis": self.config.parameter_basis,
            "sensitivity_case": self.config.sensitivity_case,
        }


def label_mask(mask, *, connectivity: int = 8, runtime=None, device: str = "cuda:0") -> np.ndarray:
    """Canonical labels of a host or device int mask (min linear index, -1 background)."""
    torch = __import__("torch")
    from .relaxation import _runtime

    rt = runtime or _runtime(device)
    if isinstance(mask, torch.Tensor):
        m = mask.to(device=rt.device, dtype=torch.int32).contiguous()
    else:
        m = torch.from_numpy(np.ascontiguousarray(np.asarray(mask, dtype=np.int32))).to(rt.device)
    if m.dim() != 2:
        raise ValueError("[LargeAvalanche] connected mask must be 2-D")
    ny, nx = m.shape
    labels = torch.empty_like(m)
    rt.bind_stream()
    rt.check(rt.lib.mv2_label_mask(rt.ctx, m.data_ptr(), labels.data_ptr(), int(ny), int(nx),
                                   int(connectivity)))
    return labels.cpu().numpy()
