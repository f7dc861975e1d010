"""Device-authoritative bulk terrain state (the plugin surface of the hot path).

Mirrors DeviceBulkState / DeviceMaterialLedger of the reference
(pkg/runtime/bulk_state_authority.py:416-1406): the same named arrays
(``z_base``, ``b_eff`` with its ``resting`` alias, ``mobile``,
``momentum_x``, ``momentum_y``, ``weights`` and the avalanche/frontier
scratch), the entrain/deposit contracts, the reservoir ledger, explicit host
views and checkpoint restore.  Arrays are torch tensors on the MI355X; the
contracts run as HIP kernels through libmv2.

precision="fp64" stores every field as float64, exactly like the reference.
precision="fp32" stores z_base/b_eff/momentum as float32 and ``mobile`` as
int32 quanta of ``MOBILE_QUANTUM_M`` (2^-24 m): the Mobile V2 step then moves
mass in whole quanta, so its mass ledger is exact.  ``mobile_m()`` and the host
views convert quanta to metres.
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass
from enum import Enum
from time import perf_counter
from typing import Iterable

import numpy as np

from . import _lib
from .config import TerrainGrid, triangle_vertex_weights
from .runtime import HipRuntime

MOBILE_QUANTUM_M = 1.0 / 16777216.0


class BulkStateAuthority(str, Enum):
    HOST = "HOST"
    DEVICE = "DEVICE"


class BulkStateAuthorityError(RuntimeError):
    """Raised when a non-authoritative view attempts to mutate physics state."""


@dataclass(frozen=True)
class DeviceMaterialLedgerSnapshot:
    initial_total_m3: float
    resting_m3: float
    mobile_m3: float
    payload_m3: float
    airborne_m3: float
    outflow_m3: float
    absolute_volume_error_m3: float
    relative_volume_error: float


@dataclass(frozen=True)
class DeviceTransferSnapshot:
    """Per-step transfer accounting with explicit full-field detection
    (bulk_state_authority.py:281-306)."""

    h2d_bytes: int
    d2h_bytes: int
    h2d_transfer_count: int
    d2h_transfer_count: int
    synchronization_count: int
    kernel_launch_count: int
    h2d_blocking_ms: float
    d2h_blocking_ms: float
    full_field_h2d_count: int
    full_field_d2h_count: int

    def to_dict(self) -> dict[str, int | float]:
        return {name: getattr(self, name) for name in self.__dataclass_fields__}


@dataclass(frozen=True)
class HostBulkStatePatch:
    """Compact non-authoritative terrain patch (bulk_state_authority.py:373-413)."""

    bbox_yx: tuple[int, int, int, int]
    H_resting_m: np.ndarray
    H_mobile_m: np.ndarray
    mobile_momentum_m2_s: np.ndarray
    timestamp_device_s: float
    source: str
    z_base_m: np.ndarray | None = None

    def __post_init__(self) -> None:
        row0, row1, col0, col1 = (int(v) for v in self.bbox_yx)
        expected = (row1 - row0, col1 - col0)
        for name in ("H_resting_m", "H_mobile_m", "mobile_momentum_m2_s"):
            arr = np.ascontiguousarray(np.asarray(getattr(self, name), dtype=np.float64))
            arr.setflags(write=False)
            object.__setattr__(self, name, arr)
        if self.H_resting_m.shape != expected or self.H_mobile_m.shape != expected:
            raise ValueError("[HostBulkStatePatch] scalar field shape mismatch")
        if self.mobile_momentum_m2_s.shape != expected + (2,):
            raise ValueError("[HostBulkStatePatch] momentum field shape mismatch")
        base = (np.zeros(expected) if self.z_base_m is None
                else np.ascontiguousarray(np.asarray(self.z_base_m, dtype=np.float64)))
        if base.shape != expected:
            raise ValueError("[HostBulkStatePatch] z_base shape mismatch")
        base.setflags(write=False)
        object.__setattr__(self, "z_base_m", base)

    @property
    def b_eff_m(self) -> np.ndarray:
        return self.H_resting_m

    @property
    def h_resting_derived_m(self) -> np.ndarray:
        return self.b_eff_m - self.z_base_m

    @property
    def H_free_m(self) -> np.ndarray:
        return self.b_eff_m + self.H_mobile_m


@dataclass(frozen=True)
class HostBulkStateView:
    """Read-only host observation taken at an explicit boundary."""

    H_resting_m: np.ndarray
    H_mobile_m: np.ndarray
    mobile_momentum_m2_s: np.ndarray
    timestamp_device_s: float
    reset_generation: int
    source: str
    dirty_tile_ids: np.ndarray
    z_base_m: np.ndarray

    @property
    def b_eff_m(self) -> np.ndarray:
        return self.H_resting_m

    @property
    def H_free_m(self) -> np.ndarray:
        return self.b_eff_m + self.H_mobile_m

    @property
    def h_resting_derived_m(self) -> np.ndarray:
        return self.b_eff_m - self.z_base_m

    def commit(self, *args, **kwargs) -> None:
        raise BulkStateAuthorityError(
            "[HostBulkStateView] host views are non-authoritative and cannot commit physics")


class DeviceBulkState:
    """HIP-resident authoritative state shared by the Mobile V2 operators."""

    backend_identity = "GPU_RUNTIME_DEVICE_AUTHORITATIVE_HIP"
    _INT_FIELDS = ("active_mask", "material_mask", "dirty_mask", "frontier_reached",
                   "avalanche_unstable", "avalanche_parent", "avalanche_component_count",
                   "avalanche_latch", "avalanche_previous_component",
                   "avalanche_activation_count", "deposition_mask")
    _F64_FIELDS = ("avalanche_slope", "avalanche_gradient_x", "avalanche_gradient_y",
                   "avalanche_severity", "avalanche_first_activation_time",
                   "avalanche_last_activation_time", "avalanche_owned_surface",
                   "avalanche_owned_export_baseline", "mobile_export_cumulative",
                   "avalanche_r2m_cumulative", "avalanche_m2r_cumulative", "track_rut",
                   "mobile_flux_export_cumulative")
    _STATE_FIELDS = ("z_base", "b_eff", "mobile", "momentum_x", "momentum_y")
    # transfers of these count as full-field transfers (bulk_state_authority.py:504-516)
    _FULL_FIELD_NAMES = frozenset({"z_base", "b_eff", "mobile", "momentum_x", "momentum_y",
                                   "initial_b_eff", "weights"})
    _MUTABLE_FLOAT = {"b_eff": 0, "mobile": 1, "momentum_x": 2, "momentum_y": 3}
    _COMPACT_INT = ("material_mask", "active_mask", "deposition_mask", "frontier_reached")

    def __init__(self, grid: TerrainGrid, initial_resting_m, *, device: str = "cuda:0",
                 tile_size: int = 64, precision: str = "fp64",
                 runtime: HipRuntime | None = None) -> None:
        if precision not in ("fp64", "fp32"):
            raise ValueError("[DeviceBulkState] precision must be 'fp64' or 'fp32'")
        self.grid = grid
        self.shape = grid.shape
        self.size = int(grid.nx * grid.ny)
        self.tile_size = int(tile_size)
        if self.tile_size < 4:
            raise ValueError("[DeviceBulkState] tile_size must be >= 4")
        self.precision = precision
        self.runtime = runtime or HipRuntime(device)
        self.authority = BulkStateAuthority.DEVICE
        self._timestamp_device_s = 0.0
        self._timestamp_host_view_s = -1.0
        self._reset_generation = 0
        self._dirty_tile_ids: set[int] = set()
        self._full_field_h2d_count = 0
        self._full_field_d2h_count = 0
        self._h2d_blocking_ms = 0.0
        self._d2h_blocking_ms = 0.0
        torch = __import__("torch")
        self._fdtype = torch.float64 if precision == "fp64" else torch.float32
        self._mdtype = torch.float64 if precision == "fp64" else torch.int32
        initial = np.asarray(grid.validate_heightmap(initial_resting_m), dtype=np.float64)
        rt = self.runtime
        rt.upload("z_base", np.zeros(self.shape).ravel(), self._fdtype)
        rt.upload("initial_b_eff", initial.ravel(), self._fdtype)
        rt.upload("b_eff", initial.ravel(), self._fdtype)
        rt.zeros("mobile", self.size, self._mdtype)
        rt.zeros("momentum_x", self.size, self._fdtype)
        rt.zeros("momentum_y", self.size, self._fdtype)
        rt.upload("weights", triangle_vertex_weights(grid.ny, grid.nx, grid.dx, grid.dy).ravel(),
                  torch.float64)
        rt.arrays["initial_resting"] = rt.arrays["initial_b_eff"]
        rt.arrays["resting"] = rt.arrays["b_eff"]
        for name in self._INT_FIELDS:
            rt.zeros(name, self.size, torch.int32)
        for name in self._F64_FIELDS:
            rt.zeros(name, self.size, torch.float64)
        rt.zeros("dirty_tile_flags", self.tile_count, torch.int32)
        rt.zeros("avalanche_activity_tile_flags", self.tile_count, torch.int32)
        rt.zeros("surface_before", self.size, torch.float64)
        # per-cell amounts of the last deposition step (metres; FP32: mobile quanta)
        rt.zeros("deposition_work", self.size, self._mdtype)
        self._bind()
        self.begin_physics_step()

    # -- binding -----------------------------------------------------------------
    def _bind(self) -> None:
        rt = self.runtime
        prec = _lib.MV2_FP64 if self.precision == "fp64" else _lib.MV2_FP32
        rc = rt.lib.mv2_bind(rt.ctx, prec, self.grid.ny, self.grid.nx, float(self.grid.dx),
                             float(self.grid.dy), rt.ptr("z_base"), rt.ptr("b_eff"), rt.ptr("mobile"),
                             rt.ptr("momentum_x"), rt.ptr("momentum_y"))
        rt.check(rc)

    @property
    def timestamp_device_s(self) -> float:
        return self._timestamp_device_s

    @property
    def reset_generation(self) -> int:
        return self._reset_generation

    @property
    def tile_shape(self) -> tuple[int, int]:
        return ((self.shape[0] + self.tile_size - 1) // self.tile_size,
                (self.shape[1] + self.tile_size - 1) // self.tile_size)

    @property
    def tile_count(self) -> int:
        ty, tx = self.tile_shape
        return int(ty * tx)

    def begin_physics_step(self) -> None:
        """Reset the per-step transfer counters (initialisation/reset excluded)."""
        self.runtime.telemetry.reset()
        self._full_field_h2d_count = 0
        self._full_field_d2h_count = 0
        self._h2d_blocking_ms = 0.0
        self._d2h_blocking_ms = 0.0

    def transfer_snapshot(self) -> DeviceTransferSnapshot:
        """Transfers since begin_physics_step (bulk_state_authority.py:656-681)."""
        t = self.runtime.telemetry
        return DeviceTransferSnapshot(
            h2d_bytes=int(t.h2d_bytes), d2h_bytes=int(t.d2h_bytes),
            h2d_transfer_count=int(t.h2d_transfer_count), d2h_transfer_count=int(t.d2h_transfer_count),
            synchronization_count=int(t.synchronization_count), kernel_launch_count=int(t.kernel_launch_count),
            h2d_blocking_ms=self._h2d_blocking_ms, d2h_blocking_ms=self._d2h_blocking_ms,
            full_field_h2d_count=self._full_field_h2d_count, full_field_d2h_count=self._full_field_d2h_count)

    def assert_normal_step_transfer_budget(self) -> None:
        """A normal physics step may move no complete terrain field (:683-689)."""
        snap = self.transfer_snapshot()
        if snap.full_field_h2d_count or snap.full_field_d2h_count:
            raise BulkStateAuthorityError(
                "[DeviceBulkState] normal GPU physics performed a full-field transfer: "
                f"H2D={snap.full_field_h2d_count}, D2H={snap.full_field_d2h_count}")

    def advance_time(self, dt_s: float) -> None:
        dt = float(dt_s)
        if not np.isfinite(dt) or dt < 0.0:
            raise ValueError("[DeviceBulkState] dt_s must be finite/non-negative")
        self._timestamp_device_s += dt

    # -- field access ----------------------------------------------------------------
    def mobile_m(self):
        """Mobile thickness in metres as a device float64 tensor."""
        m = self.runtime.arrays["mobile"]
        return m.double() * MOBILE_QUANTUM_M if self.precision == "fp32" else m

    def set_fields(self, *, z_base=None, b_eff=None, mobile=None, momentum_x=None,
                   momentum_y=None) -> None:
        """Explicit full-field initialisation (acceptance/checkpoint boundary)."""
        torch = __import__("torch")
        rt = self.runtime
        for name, value in (("z_base", z_base), ("b_eff", b_eff), ("momentum_x", momentum_x),
                            ("momentum_y", momentum_y)):
            if value is None:
                continue
            arr = np.asarray(value, dtype=np.float64).reshape(self.shape)
            rt.arrays[name].copy_(torch.from_numpy(arr.ravel()).to(rt.device, self._fdtype))
            rt.telemetry.record_h2d(arr)
            self._full_field_h2d_count += 1
        if mobile is not None:
            arr = np.asarray(mobile, dtype=np.float64).reshape(self.shape)
            if np.any(arr < 0.0):
                raise ValueError("MobileV2 h must be nonnegative")
            if self.precision == "fp32":
                q = np.rint(arr * (1.0 / MOBILE_QUANTUM_M))
                if np.any(q > np.iinfo(np.int32).max):
                    raise ValueError("[DeviceBulkState] mobile thickness exceeds the FP32 storage range")
                t = torch.from_numpy(q.astype(np.int32).ravel())
            else:
                t = torch.from_numpy(arr.ravel())
            rt.arrays["mobile"].copy_(t.to(rt.device))
            rt.telemetry.record_h2d(arr)
            self._full_field_h2d_count += 1

    # -- contracts -------------------------------------------------------------------
    def _compact(self, flat_indices, values, what: str):
        torch = __import__("torch")
        indices = np.asarray(flat_indices, dtype=np.int64).reshape(-1)
        depth = np.asarray(values, dtype=np.float64).reshape(-1)
        if depth.shape != indices.shape or not np.all(np.isfinite(depth)):
            raise ValueError(f"[DeviceBulkState] invalid {what} request")
        if np.any(indices < 0) or np.any(indices >= self.size):
            raise ValueError(f"[DeviceBulkState] {what} index outside state")
        if np.unique(indices).size != indices.size:
            raise ValueError(f"[DeviceBulkState] duplicate {what} indices")
        idx_t = torch.from_numpy(indices.astype(np.int32)).to(self.runtime.device)
        dep_t = torch.from_numpy(depth).to(self.runtime.device)
        self.runtime.telemetry.record_h2d(indices.astype(np.int32))
        self.runtime.telemetry.record_h2d(depth)
        return indices, idx_t, dep_t

    def entrain_host_indices(self, flat_indices, requested_depth_m, *, reason: str) -> float:
        """Resting -> Mobile transfer (bulk_state_authority.py:920-960)."""
        if not reason:
            raise ValueError("[DeviceBulkState] entrainment reason must be non-empty")
        indices, idx_t, dep_t = self._compact(flat_indices, requested_depth_m, "entrainment")
        if not indices.size:
            return 0.0
        rt = self.runtime
        rt.bind_stream()
        out = ctypes.c_double(0.0)
        rc = rt.lib.mv2_entrain(rt.ctx, idx_t.data_ptr(), dep_t.data_ptr(), int(indices.size), 1,
                                ctypes.byref(out))
        rt.check(rc)
        rt.telemetry.record_call()
        rows, cols = np.divmod(indices, self.shape[1])
        self.mark_dirty_bbox((int(rows.min()), int(rows.max()) + 1, int(cols.min()), int(cols.max()) + 1))
        return float(out.value)

    def deposit_host_indices(self, flat_indices, requested_depth_m, *, density_kg_m3: float,
                             reason: str) -> dict:
        """Mobile -> Resting transfer with reported impulse (bulk_state_authority.py:962-1002)."""
        if not reason:
            raise ValueError("[DeviceBulkState] deposition reason must be non-empty")
        density = float(density_kg_m3)
        if not density > 0.0:
            raise ValueError("[DeviceBulkState] invalid deposition indices/density")
        indices, idx_t, dep_t = self._compact(flat_indices, requested_depth_m, "deposition")
        if not indices.size:
            return {"deposited_volume_m3": 0.0, "removed_mobile_momentum_kg_m_s": np.zeros(3)}
        rt = self.runtime
        rt.bind_stream()
        out = np.zeros(3)
        rc = rt.lib.mv2_deposit(rt.ctx, idx_t.data_ptr(), dep_t.data_ptr(), int(indices.size), 1,
                                density, out.ctypes.data)
        rt.check(rc)
        rt.telemetry.record_call()
        rows, cols = np.divmod(indices, self.shape[1])
        self.mark_dirty_bbox((int(rows.min()), int(rows.max()) + 1, int(cols.min()), int(cols.max()) + 1))
        return {"deposited_volume_m3": float(out[0]),
                "removed_mobile_momentum_kg_m_s": np.asarray([out[1], out[2], 0.0])}

    # -- compact host bridges (bulk_state_authority.py:804-918, :1036-1137) ------------
    def _upload_compact(self, values, dtype):
        """A compact command/patch buffer (never a terrain field) to the device."""
        torch = __import__("torch")
        start = perf_counter()
        host = np.ascontiguousarray(values)
        t = torch.from_numpy(host).to(self.runtime.device, non_blocking=False)
        self.runtime.telemetry.record_h2d(host)
        self._h2d_blocking_ms += (perf_counter() - start) * 1e3
        return t if t.dtype == dtype else t.to(dtype)

    def apply_host_patch(self, field: str, bbox_yx, values, *, reason: str) -> None:
        """Commit a compact host-generated patch of one float field (:804-845)."""
        torch = __import__("torch")
        if field == "resting":
            field = "b_eff"
        if field not in self._MUTABLE_FLOAT:
            raise ValueError("[DeviceBulkState] unsupported mutable field")
        if not reason:
            raise ValueError("[DeviceBulkState] patch reason must be non-empty")
        row0, row1, col0, col1 = (int(v) for v in bbox_yx)
        if not (0 <= row0 <= row1 <= self.shape[0] and 0 <= col0 <= col1 <= self.shape[1]):
            raise ValueError("[DeviceBulkState] patch bbox outside state")
        patch = np.asarray(values, dtype=np.float64)
        if patch.shape != (row1 - row0, col1 - col0) or not np.all(np.isfinite(patch)):
            raise ValueError("[DeviceBulkState] patch values shape/finite check failed")
        if patch.size == 0:
            return
        if field == "mobile" and self.precision == "fp32":
            q = np.rint(patch / MOBILE_QUANTUM_M)
            if np.any(q < np.iinfo(np.int32).min) or np.any(q > np.iinfo(np.int32).max):
                raise ValueError("[DeviceBulkState] mobile thickness exceeds the FP32 storage range")
        dev = self._upload_compact(patch.ravel(), torch.float64)
        rt = self.runtime
        rt.bind_stream()
        rt.check(rt.lib.mv2_patch_scatter(rt.ctx, self._MUTABLE_FLOAT[field], row0, row1, col0, col1,
                                          dev.data_ptr()))
        rt.telemetry.record_call(syncs=0)
        self.mark_dirty_bbox((row0, row1, col0, col1))

    def apply_host_indices(self, field: str, flat_indices, values, *, reason: str) -> None:
        """Scatter compact activation/mask data into an int32 field (:847-885)."""
        torch = __import__("torch")
        if field not in self._COMPACT_INT:
            raise ValueError("[DeviceBulkState] unsupported compact integer field")
        if not reason:
            raise ValueError("[DeviceBulkState] compact scatter reason must be non-empty")
        indices = np.asarray(flat_indices, dtype=np.int64).reshape(-1)
        if np.any(indices < 0) or np.any(indices >= self.size):
            raise ValueError("[DeviceBulkState] compact scatter index outside state")
        scalar = np.isscalar(values)
        if not scalar:
            data = np.asarray(values, dtype=np.int32).reshape(-1)
            if data.shape != indices.shape:
                raise ValueError("[DeviceBulkState] compact scatter values shape mismatch")
        if not indices.size:
            return
        rt = self.runtime
        idx = self._upload_compact(indices.astype(np.int32), torch.int32)
        vals = None if scalar else self._upload_compact(data, torch.int32)
        rt.bind_stream()
        rt.check(rt.lib.mv2_index_scatter_int(rt.ctx, rt.ptr(field), idx.data_ptr(),
                                              None if vals is None else vals.data_ptr(),
                                              int(values) if scalar else 0, int(indices.size)))
        rt.telemetry.record_call(syncs=0)

    def add_host_indices(self, field: str, flat_indices, values, *, reason: str) -> None:
        """Conservatively add compact source terms to a float field (:887-918);
        repeated indices accumulate (atomic adds)."""
        torch = __import__("torch")
        if field == "resting":
            field = "b_eff"
        if field not in self._MUTABLE_FLOAT:
            raise ValueError("[DeviceBulkState] unsupported compact float field")
        if not reason:
            raise ValueError("[DeviceBulkState] compact add reason must be non-empty")
        indices = np.asarray(flat_indices, dtype=np.int64).reshape(-1)
        data = np.asarray(values, dtype=np.float64).reshape(-1)
        if data.shape != indices.shape or not np.all(np.isfinite(data)):
            raise ValueError("[DeviceBulkState] compact add values shape/finite check failed")
        if np.any(indices < 0) or np.any(indices >= self.size):
            raise ValueError("[DeviceBulkState] compact add index outside state")
        if not indices.size:
            return
        if field == "mobile" and self.precision == "fp32" and np.any(np.abs(data) / MOBILE_QUANTUM_M > 2.0 ** 31):
            raise ValueError("[DeviceBulkState] mobile thickness exceeds the FP32 storage range")
        rt = self.runtime
        idx = self._upload_compact(indices.astype(np.int32), torch.int32)
        vals = self._upload_compact(data, torch.float64)
        rt.bind_stream()
        rt.check(rt.lib.mv2_index_add(rt.ctx, self._MUTABLE_FLOAT[field], idx.data_ptr(), vals.data_ptr(),
                                      int(indices.size)))
        rt.telemetry.record_call(syncs=0)

    def _gather_region(self, row0, row1, col0, col1, source: str) -> HostBulkStatePatch:
        torch = __import__("torch")
        rt = self.runtime
        h, w = row1 - row0, col1 - col0
        start = perf_counter()
        dev = torch.empty(5 * h * w, dtype=torch.float64, device=rt.device)
        rt.bind_stream()
        rt.check(rt.lib.mv2_region_gather(rt.ctx, row0, row1, col0, col1, dev.data_ptr()))
        host = dev.cpu().numpy().reshape(5, h, w)
        rt.telemetry.record_call(syncs=1)
        rt.telemetry.record_d2h(host)
        self._d2h_blocking_ms += (perf_counter() - start) * 1e3
        self._timestamp_host_view_s = self._timestamp_device_s
        return HostBulkStatePatch((row0, row1, col0, col1), host[0], host[2], np.stack((host[3], host[4]), -1),
                                  self._timestamp_device_s, source, host[1])

    def download_patch(self, bbox_yx, *, source: str) -> HostBulkStatePatch:
        """A bounded patch (at most one tile of cells) for a host geometry operator (:1036-1085)."""
        if not source:
            raise ValueError("[DeviceBulkState] patch source must be non-empty")
        row0, row1, col0, col1 = (int(v) for v in bbox_yx)
        if not (0 <= row0 < row1 <= self.shape[0] and 0 <= col0 < col1 <= self.shape[1]):
            raise ValueError("[DeviceBulkState] patch bbox outside state")
        if (row1 - row0) * (col1 - col0) > self.tile_size * self.tile_size:
            raise ValueError("[DeviceBulkState] patch exceeds compact bridge capacity; split into tiles")
        return self._gather_region(row0, row1, col0, col1, source)

    def download_region(self, bbox_yx, *, source: str) -> HostBulkStatePatch:
        """A geometry-driven compact region of any extent short of the whole
        terrain (:1087-1137), gathered by one launch and one transfer (the
        reference stitches one download per tile and field)."""
        row0, row1, col0, col1 = (int(v) for v in bbox_yx)
        if not (0 <= row0 < row1 <= self.shape[0] and 0 <= col0 < col1 <= self.shape[1]):
            raise ValueError("[DeviceBulkState] region bbox outside state")
        if (row1 - row0) * (col1 - col0) >= self.size:
            raise BulkStateAuthorityError(
                "GPU_RUNTIME_FULL_FIELD_TERRAIN_REQUEST: compact bridge requested whole terrain")
        if not source:
            raise ValueError("[DeviceBulkState] patch source must be non-empty")
        return self._gather_region(row0, row1, col0, col1, source)

    def reservoir_reduction(self, density_kg_m3: float) -> dict:
        density = float(density_kg_m3)
        if not np.isfinite(density) or density <= 0.0:
            raise ValueError("[DeviceBulkState] density must be finite/positive")
        rt = self.runtime
        rt.bind_stream()
        out = np.zeros(4)
        rt.check(rt.lib.mv2_reservoirs(rt.ctx, 1, density, out.ctypes.data))
        rt.telemetry.record_d2h(out)
        return {"resting_volume_m3": float(out[0]), "mobile_volume_m3": float(out[1]),
                "mobile_momentum_kg_m_s": np.asarray([out[2], out[3], 0.0])}

    def reservoir_reduction_into(self, out, density_kg_m3: float) -> None:
        """The sums of ``reservoir_reduction`` written asynchronously into the
        device float64 tensor ``out[4]`` (resting, mobile, density * px,
        density * py) without allocating device memory or synchronising --
        usable under HIP graph capture.  No counterpart in the reference."""
        torch = __import__("torch")
        density = float(density_kg_m3)
        if not np.isfinite(density) or density <= 0.0:
            raise ValueError("[DeviceBulkState] density must be finite/positive")
        if (out.dtype != torch.float64 or out.device != self.runtime.device or out.numel() < 4
                or not out.is_contiguous()):
            raise ValueError("[DeviceBulkState] out must be a contiguous float64 device tensor of >= 4 values")
        rt = self.runtime
        rt.bind_stream()
        rt.check(rt.lib.mv2_reservoirs_noalloc(rt.ctx, 1, density, out.data_ptr()))

    # -- dirty tiles -----------------------------------------------------------------
    def mark_dirty_tiles(self, tile_ids: Iterable[int]) -> None:
        for tile_id in np.asarray(tuple(tile_ids), dtype=np.int64).ravel():
            if tile_id < 0 or tile_id >= self.tile_count:
                raise ValueError("[DeviceBulkState] dirty tile ID outside grid")
            self._dirty_tile_ids.add(int(tile_id))

    def mark_dirty_bbox(self, bbox_yx) -> None:
        row0, row1, col0, col1 = (int(v) for v in bbox_yx)
        if row1 <= row0 or col1 <= col0:
            return
        ty, tx = self.tile_shape
        fy, ly = max(0, row0 // self.tile_size), min(ty - 1, (row1 - 1) // self.tile_size)
        fx, lx = max(0, col0 // self.tile_size), min(tx - 1, (col1 - 1) // self.tile_size)
        self.mark_dirty_tiles(r * tx + c for r in range(fy, ly + 1) for c in range(fx, lx + 1))

    def capture_surface_for_dirty_tracking(self) -> None:
        """Snapshot b_eff + mobile before a mutating operator (bulk_state_authority.py:733-745)."""
        rt = self.runtime
        rt.bind_stream()
        rt.check(rt.lib.mv2_surface_capture(rt.ctx, rt.ptr("surface_before")))

    def collect_surface_dirty_tiles(self, *, tolerance_m: float = 1.0e-10) -> np.ndarray:
        """Tiles whose surface moved by more than the tolerance since the capture
        (:747-781): flagged and compacted to ascending ids on the device."""
        tolerance = float(tolerance_m)
        if not np.isfinite(tolerance) or tolerance < 0.0:
            raise ValueError("[DeviceBulkState] dirty tolerance must be finite/non-negative")
        rt = self.runtime
        rt.bind_stream()
        ids = np.zeros(self.tile_count, dtype=np.int32)
        count = ctypes.c_int(0)
        rt.check(rt.lib.mv2_surface_dirty_tiles(
            rt.ctx, rt.ptr("surface_before"), tolerance, self.tile_size, self.tile_shape[1],
            self.tile_count, rt.ptr("dirty_tile_flags"), ids.ctypes.data, ctypes.byref(count)))
        tile_ids = ids[:count.value].copy()
        rt.telemetry.record_d2h(tile_ids)
        self.mark_dirty_tiles(tile_ids)
        return tile_ids

    def sample_surface_bilinear(self, row_col, *, source: str) -> np.ndarray:
        """Bilinear free-surface heights at fractional (row, col) queries (:1138-1181)."""
        torch = __import__("torch")
        if not source:
            raise ValueError("[DeviceBulkState] terrain query source must be non-empty")
        queries = np.asarray(row_col, dtype=np.float64)
        if queries.ndim != 2 or queries.shape[1] != 2 or not np.all(np.isfinite(queries)):
            raise ValueError("[DeviceBulkState] terrain queries must be finite [N,2]")
        count = int(queries.shape[0])
        if count == 0:
            return np.empty(0, dtype=np.float64)
        rt = self.runtime
        q = torch.from_numpy(np.ascontiguousarray(queries)).to(rt.device)
        out = torch.empty(count, dtype=torch.float64, device=rt.device)
        rt.bind_stream()
        rt.check(rt.lib.mv2_surface_sample(rt.ctx, q.data_ptr(), count, out.data_ptr()))
        values = out.cpu().numpy()
        rt.telemetry.record_h2d(queries)
        rt.telemetry.record_d2h(values)
        return values

    def download_dirty_tiles(self, tile_ids=None) -> dict[int, np.ndarray]:
        """Current surface of the given (default: the recorded dirty) tiles with
        their one-cell overlap (:1280-1328), gathered by one launch."""
        torch = __import__("torch")
        selected = (np.asarray(tuple(tile_ids), dtype=np.int64).ravel() if tile_ids is not None
                    else self.consume_dirty_tile_ids().astype(np.int64))
        if selected.size == 0:
            return {}
        if np.any(selected < 0) or np.any(selected >= self.tile_count):
            raise ValueError("[DeviceBulkState] dirty tile ID outside grid")
        rt = self.runtime
        ts = self.tile_size
        stride = (ts + 1) * (ts + 1)
        result: dict[int, np.ndarray] = {}
        tiles_x = self.tile_shape[1]
        for start in range(0, selected.size, 65535):
            chunk = selected[start:start + 65535]
            ids = torch.from_numpy(chunk.astype(np.int32)).to(rt.device)
            out = torch.empty(chunk.size * stride, dtype=torch.float64, device=rt.device)
            rt.bind_stream()
            rt.check(rt.lib.mv2_surface_gather_tiles(rt.ctx, ids.data_ptr(), int(chunk.size), ts, tiles_x,
                                                     self.tile_count, out.data_ptr()))
            host = out.cpu().numpy()
            rt.telemetry.record_d2h(host)
            for j, tile in enumerate(chunk.tolist()):
                r0, c0 = (tile // tiles_x) * ts, (tile % tiles_x) * ts
                h = min(r0 + ts + 1, self.shape[0]) - r0
                w = min(c0 + ts + 1, self.shape[1]) - c0
                result[int(tile)] = host[j * stride:j * stride + h * w].reshape(h, w).copy()
        self._timestamp_host_view_s = self._timestamp_device_s
        return result

    def consume_dirty_tile_ids(self) -> np.ndarray:
        out = np.asarray(sorted(self._dirty_tile_ids), dtype=np.int32)
        self._dirty_tile_ids.clear()
        return out

    # -- explicit boundaries ---------------------------------------------------------
    def explicit_host_view(self, *, source: str) -> HostBulkStateView:
        allowed = {"debug", "acceptance", "checkpoint", "reset", "visualization"}
        if source not in allowed:
            raise BulkStateAuthorityError(
                "[DeviceBulkState] full host views are restricted to explicit boundaries, "
                f"got source={source!r}")
        f = self.download_fields(self._STATE_FIELDS)
        resting, z, mobile = f["b_eff"], f["z_base"], f["mobile"]
        q = np.stack((f["momentum_x"], f["momentum_y"]), axis=-1)
        self._timestamp_host_view_s = self._timestamp_device_s
        return HostBulkStateView(resting, mobile, q, self._timestamp_device_s,
                                 self._reset_generation, source, self.consume_dirty_tile_ids(), z)

    # -- packed full-field transfers (explicit boundaries only) ---------------------
    def _is_float_field(self, name: str) -> bool:
        return name in self._STATE_FIELDS or name in self._F64_FIELDS

    def download_fields(self, names) -> dict[str, np.ndarray]:
        """Named fields as host arrays -- float fields in float64 metres (FP32
        mobile quanta times 2^-24, exact), int fields int32 -- packed on the
        device and copied in one pinned transfer per dtype."""
        torch = __import__("torch")
        rt = self.runtime
        names = list(names)
        floats = [n for n in names if self._is_float_field(n)]
        ints = [n for n in names if n in self._INT_FIELDS]
        unknown = sorted(set(names) - set(floats) - set(ints))
        if unknown:
            raise BulkStateAuthorityError(f"[DeviceBulkState] unknown fields: {unknown}")
        out: dict[str, np.ndarray] = {}
        start = perf_counter()
        self._full_field_d2h_count += sum(1 for n in names if n in self._FULL_FIELD_NAMES)
        for group, dtype in ((floats, torch.float64), (ints, torch.int32)):
            if not group:
                continue
            parts = []
            for name in group:
                t = rt.arrays[name]
                if name == "mobile" and self.precision == "fp32":
                    t = t.double() * MOBILE_QUANTUM_M
                parts.append(t.reshape(-1).to(dtype))
            packed = torch.cat(parts)
            host = torch.empty(packed.shape, dtype=dtype, pin_memory=True)
            host.copy_(packed)
            arr = host.numpy()
            rt.telemetry.record_d2h(arr)
            for k, name in enumerate(group):
                out[name] = arr[k * self.size:(k + 1) * self.size].reshape(self.shape).copy()
        self._d2h_blocking_ms += (perf_counter() - start) * 1e3
        return out

    def _upload_fields(self, fields: dict) -> None:
        torch = __import__("torch")
        rt = self.runtime
        floats = [n for n in fields if self._is_float_field(n)]
        ints = [n for n in fields if n in self._INT_FIELDS]
        start = perf_counter()
        self._full_field_h2d_count += sum(1 for n in fields if n in self._FULL_FIELD_NAMES)
        for group, dtype in ((floats, np.float64), (ints, np.int32)):
            if not group:
                continue
            packed = np.concatenate([np.ascontiguousarray(np.asarray(fields[n], dtype=dtype)).ravel()
                                     for n in group])
            rt.telemetry.record_h2d(packed)
            dev = torch.from_numpy(packed).pin_memory().to(rt.device, non_blocking=True)
            for k, name in enumerate(group):
                view = dev[k * self.size:(k + 1) * self.size]
                target = rt.arrays[name]
                if name == "mobile" and self.precision == "fp32":
                    q = torch.round(view * (1.0 / MOBILE_QUANTUM_M))
                    if bool(((q > np.iinfo(np.int32).max) | (q < np.iinfo(np.int32).min)).any()):
                        raise ValueError("[DeviceBulkState] mobile thickness exceeds the FP32 storage range")
                    target.copy_(q.to(torch.int32))
                else:
                    target.copy_(view.to(target.dtype))
        rt.synchronize()
        self._h2d_blocking_ms += (perf_counter() - start) * 1e3

    def restore_checkpoint_fields(self, fields: dict, *, timestamp_s: float, source: str) -> None:
        """Explicit checkpoint boundary (bulk_state_authority.py:1219-1278):
        validated fields uploaded packed, one transfer per dtype."""
        if source not in {"acceptance", "checkpoint"}:
            raise BulkStateAuthorityError(
                "[DeviceBulkState] checkpoint restore requires an explicit boundary")
        known = {"resting"} | set(self._STATE_FIELDS) | set(self._F64_FIELDS) | set(self._INT_FIELDS)
        unknown = sorted(set(fields) - known)
        if unknown:
            raise BulkStateAuthorityError(f"[DeviceBulkState] unsupported checkpoint fields: {unknown}")
        values = {}
        for name, value in fields.items():
            key = "b_eff" if name == "resting" else name
            arr = np.asarray(value)
            if arr.shape != self.shape or not np.all(np.isfinite(arr)):
                raise BulkStateAuthorityError(f"[DeviceBulkState] invalid checkpoint field {key!r}")
            values[key] = arr
        t = float(timestamp_s)
        if not np.isfinite(t) or t < 0.0:
            raise BulkStateAuthorityError("[DeviceBulkState] checkpoint timestamp must be finite/non-negative")
        self._upload_fields(values)
        self._timestamp_device_s = t
        self._timestamp_host_view_s = -1.0
        self.mark_dirty_tiles(range(self.tile_count))
        self.begin_physics_step()

    def reset(self) -> None:
        rt = self.runtime
        rt.arrays["b_eff"].copy_(rt.arrays["initial_b_eff"])
        for name in ("z_base", "mobile", "momentum_x", "momentum_y", "dirty_tile_flags",
                     "avalanche_activity_tile_flags", "deposition_work", *self._INT_FIELDS, *self._F64_FIELDS):
            rt.arrays[name].zero_()
        rt.synchronize()
        self._timestamp_device_s = 0.0
        self._timestamp_host_view_s = -1.0
        self._dirty_tile_ids = set(range(self.tile_count))
        self._reset_generation += 1
        self.begin_physics_step()

    def assert_consistent(self) -> None:
        if self.authority is not BulkStateAuthority.DEVICE:
            raise BulkStateAuthorityError("[DeviceBulkState] state authority is not DEVICE")
        if self.runtime.arrays["resting"] is not self.runtime.arrays["b_eff"]:
            raise BulkStateAuthorityError("[DeviceBulkState] resting alias diverged from b_eff")

    def diagnostics(self) -> dict:
        self.assert_consistent()
        return {
            "backend_identity": self.backend_identity,
            "state_authority": self.authority.value,
            "authoritative_geometry": ["z_base", "b_eff", "mobile"],
            "precision": self.precision,
            "mobile_storage": "int32 quanta of 2^-24 m" if self.precision == "fp32" else "float64 m",
            "shape_yx": list(self.shape),
            "tile_size": self.tile_size,
            "timestamp_device_s": self._timestamp_device_s,
            "reset_generation": self._reset_generation,
            "resident_array_names": sorted(self.runtime.arrays),
        }


class DeviceMaterialLedger:
    """Scalar material ledger whose terrain entries come from device reductions."""

    def __init__(self, state: DeviceBulkState, density_kg_m3: float, *, payload_m3: float = 0.0,
                 airborne_m3: float = 0.0, outflow_m3: float = 0.0,
                 initial_total_m3: float | None = None) -> None:
        values = state.reservoir_reduction(density_kg_m3)
        self._density_kg_m3 = float(density_kg_m3)
        computed = float(values["resting_volume_m3"] + values["mobile_volume_m3"]
                         + payload_m3 + airborne_m3 + outflow_m3)
        self._initial_total_m3 = computed if initial_total_m3 is None else float(initial_total_m3)
        if not np.isfinite(self._initial_total_m3) or self._initial_total_m3 < 0.0:
            raise ValueError("[DeviceMaterialLedger] initial total must be finite/non-negative")

    @property
    def initial_total_m3(self) -> float:
        return self._initial_total_m3

    def snapshot(self, state: DeviceBulkState, *, payload_m3: float, airborne_m3: float,
                 outflow_m3: float) -> DeviceMaterialLedgerSnapshot:
        s = state.reservoir_reduction(self._density_kg_m3)
        current = float(s["resting_volume_m3"] + s["mobile_volume_m3"] + float(payload_m3)
                        + float(airborne_m3) + float(outflow_m3))
        error = current - self._initial_total_m3
        return DeviceMaterialLedgerSnapshot(
            initial_total_m3=self._initial_total_m3, resting_m3=float(s["resting_volume_m3"]),
            mobile_m3=float(s["mobile_volume_m3"]), payload_m3=float(payload_m3),
            airborne_m3=float(airborne_m3), outflow_m3=float(outflow_m3),
            absolute_volume_error_m3=error,
            relative_volume_error=abs(error) / max(abs(self._initial_total_m3), 1.0e-12))
