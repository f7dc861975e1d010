"""Angle-of-repose relaxation on the MI355X.

* ``relax_critical_slope`` / ``relax_device``: slope_model.relax_critical_slope
  (slope_model.py:273-343) -- the plain relaxation and the error judge of the
  learned warm start -- one HIP launch per iteration (LDS tiles, all 8
  red-black sub-phases of an iteration per tile, device-side convergence
  test, launches enqueued in batches with one host read per batch).
* ``HipMinimumSlopeAdapter``: MinimumSlopeAdapter (pkg/solvers/
  minimum_slope_adapter.py:14-320) on canonical H[y, x], closed boundary.
* ``CompactTileFrontier`` / ``CompactActiveEdgeBatch``: the host-side edge
  batch builder (pkg/solvers/sparse_tile_slope.py:20-230), and
  ``HipCompactActiveEdgeOperator``: WarpCompactActiveEdgeOperator
  (pkg/solvers/warp_compact_frontier.py:144-405) with HIP kernels, plus
  ``advance_rounds``: the residual projection loop of the frame scheduler
  (pkg/runtime/v2_physics_core.py:1172-1254) run on the device.
* ``GpuResidualFrontier``: that loop's seeds, active tile set and failure
  modes (v2_physics_core.py:1142-1254) around ``advance_rounds``.
"""

from __future__ import annotations

import ctypes
from collections import OrderedDict
from dataclasses import dataclass, field
from types import MappingProxyType

import numpy as np

from . import _lib
from .config import SolverConfig, TerrainGrid
from .runtime import HipRuntime


@dataclass(frozen=True)
class RelaxationStats:
    iterations: int
    converged: bool
    max_slope: float
    volume_before: float
    volume_after: float


@dataclass(frozen=True)
class RelaxationResult:
    heightmap_stable: np.ndarray
    heightmap_sequence: tuple
    iteration_count: int
    volume_before_m3: float
    volume_after_m3: float
    boundary_outflow_m3: float
    converged: bool
    diagnostics: dict = field(default_factory=dict, compare=False)

    def __post_init__(self) -> None:
        stable = np.asarray(self.heightmap_stable, dtype=np.float64)
        if stable.ndim != 2 or not np.all(np.isfinite(stable)) or np.any(stable < 0.0):
            raise ValueError("[TerrainRelaxationSolver] stable heightmap must be finite, non-negative 2-D")
        object.__setattr__(self, "diagnostics", MappingProxyType(dict(self.diagnostics)))


def legacy_critical_differences(dx: float, dy: float, critical_angle_deg: float) -> np.ndarray:
    """critical_slope * distance for _neighbor_pairs (slope_model.py:744-750)."""
    critical_slope = float(np.tan(np.deg2rad(critical_angle_deg)))
    diagonal = float(np.hypot(dx, dy))
    return np.asarray([critical_slope * dx, critical_slope * dy,
                       critical_slope * diagonal, critical_slope * diagonal], dtype=np.float64)


_RUNTIMES: dict[str, HipRuntime] = {}


def _runtime(device: str) -> HipRuntime:
    rt = _RUNTIMES.get(device)
    if rt is None:
        rt = HipRuntime(device)
        _RUNTIMES[device] = rt
    return rt


_CRIT_CACHE: dict = {}


def relax_device(height_yx, dx: float, dy: float, critical_angle_deg: float = 34.0,
                 max_iterations: int = 20_000, tolerance: float = 1e-8, *,
                 runtime: HipRuntime | None = None) -> tuple[int, bool, float]:
    """Relax a device ``H[y, x]`` tensor (float64 or float32) in place.

    Semantics are those of MinimumSlopeAdapter.solve with a closed boundary:
    relax_critical_slope applied to the legacy ``H[x, y]`` view.  Returns
    (iterations, converged, max excess of the last iteration).
    """
    torch = __import__("torch")
    if height_yx.dim() != 2 or min(height_yx.shape) < 2:
        raise ValueError("height must be a 2-D array with both dimensions >= 2")
    if not height_yx.is_contiguous():
        raise ValueError("[MiniSlope] device height must be contiguous")
    if height_yx.dtype == torch.float64:
        prec = _lib.MV2_FP64
    elif height_yx.dtype == torch.float32:
        prec = _lib.MV2_FP32
    else:
        raise TypeError("[MiniSlope] device height must be float64 or float32")
    if not 0.0 < critical_angle_deg < 90.0:
        raise ValueError("critical_angle_deg must be between 0 and 90")
    rt = runtime or _runtime(str(height_yx.device))
    rt.bind_stream()
    key = (float(dx), float(dy), float(critical_angle_deg))
    crit = _CRIT_CACHE.get(key)
    if crit is None:
        crit = _CRIT_CACHE.setdefault(key, legacy_critical_differences(dx, dy, critical_angle_deg))
    res = _lib.RelaxResult()
    ny, nx = height_yx.shape
    rc = rt.lib.mv2_relax(rt.ctx, prec, height_yx.data_ptr(), int(ny), int(nx), crit.ctypes.data,
                          int(max_iterations), float(tolerance), ctypes.byref(res))
    rt.check(rc)
    return int(res.iterations), bool(res.converged), float(res.max_excess_last)


def max_neighbor_slope(height: np.ndarray, grid_spacing: tuple[float, float]) -> float:
    """Maximum slope between 8-connected cells (slope_model.py:261-270)."""
    dx, dy = grid_spacing
    h = np.asarray(height, dtype=np.float64)
    diag = float(np.hypot(dx, dy))
    pairs = ((h[:-1, :], h[1:, :], dx), (h[:, :-1], h[:, 1:], dy),
             (h[:-1, :-1], h[1:, 1:], diag), (h[:-1, 1:], h[1:, :-1], diag))
    return max(float(np.max(np.abs(a - b) / d)) for a, b, d in pairs)


def relax_critical_slope(height: np.ndarray, grid_spacing: tuple[float, float],
                         critical_angle_deg: float = 34.0, max_iterations: int = 20_000,
                         tolerance: float = 1e-8, *, device: str = "cuda:0",
                         dtype=np.float64) -> tuple[np.ndarray, RelaxationStats]:
    """Host-array entry with the reference signature (legacy ``H[x, y]`` layout)."""
    torch = __import__("torch")
    h = np.asarray(height)
    if h.ndim != 2 or min(h.shape) < 2:
        raise ValueError("height must be a 2-D array with both dimensions >= 2")
    if np.any(h < 0):
        raise ValueError("soil heights must be non-negative")
    if not 0.0 < critical_angle_deg < 90.0:
        raise ValueError("critical_angle_deg must be between 0 and 90")
    dx, dy = grid_spacing
    if dx <= 0 or dy <= 0:
        raise ValueError("grid spacing must be positive")
    host = np.asarray(h, dtype=np.float64)
    volume_before = float(host.sum() * (dx * dy))
    tdtype = torch.float64 if np.dtype(dtype) == np.float64 else torch.float32
    yx = torch.from_numpy(np.ascontiguousarray(host.T)).to(device=device, dtype=tdtype).contiguous()
    iterations, converged, _ = relax_device(yx, dx, dy, critical_angle_deg, max_iterations, tolerance)
    out = np.ascontiguousarray(yx.cpu().numpy().astype(np.float64).T)
    stats = RelaxationStats(iterations, converged, max_neighbor_slope(out, (dx, dy)),
                            volume_before, float(out.sum() * (dx * dy)))
    return out, stats


class HipMinimumSlopeAdapter:
    """MinimumSlopeAdapter on canonical ``H[y, x]`` backed by the HIP relaxation.

    Closed boundaries: one device relaxation to convergence.  Open boundaries
    (minimum_slope_adapter.py:160-181, :185-212): the field padded with a
    one-cell ring at ``boundary_height_m`` stays on the device; every
    iteration the ring is reset to that height and one relaxation iteration
    runs on the padded field, until an iteration converges; material that
    left the interior is the boundary outflow."""

    backend_name = "GPU_HIP_FULL_DOMAIN_MINIMUM_SLOPE"

    def __init__(self, grid: TerrainGrid, config: SolverConfig, *, device: str = "cuda:0",
                 precision: str = "fp64") -> None:
        if grid.valid_mask is not None and not np.all(grid.valid_mask):
            raise ValueError("[MinimumSlopeAdapter] legacy solver does not support holes in valid_mask")
        if config.boundary_condition not in ("closed", "open"):
            raise ValueError("[Config:Solver] boundary_condition must be 'closed' or 'open'")
        self.grid = grid
        self.config = config
        self.device = device
        self.precision = precision
        self._last_result: RelaxationResult | None = None

    @property
    def last_result(self):
        return self._last_result

    def reset(self) -> None:
        self._last_result = None

    def _solve(self, heightmap, max_iterations: int) -> RelaxationResult:
        torch = __import__("torch")
        initial = np.array(self.grid.validate_heightmap(heightmap), dtype=np.float64, copy=True)
        if np.any(initial < 0.0):
            raise ValueError("soil heights must be non-negative")
        tdtype = torch.float64 if self.precision == "fp64" else torch.float32
        cfg = self.config
        if cfg.boundary_condition == "closed":
            yx = torch.from_numpy(initial).to(device=self.device, dtype=tdtype).contiguous()
            iterations, converged, last = relax_device(
                yx, self.grid.dx, self.grid.dy, cfg.critical_angle_deg, max_iterations, cfg.tolerance)
            stable = yx.cpu().numpy().astype(np.float64)
        else:
            # the padded field lives on the device; the ring is reset before
            # every single-iteration call, as the reference pads every call
            ring = float(cfg.boundary_height_m)
            if ring < 0.0:
                raise ValueError("soil heights must be non-negative")
            padded = np.pad(initial, 1, mode="constant", constant_values=ring)
            yx = torch.from_numpy(padded).to(device=self.device, dtype=tdtype).contiguous()
            runtime = _runtime(str(yx.device))
            iterations, converged, last = 0, False, 0.0
            for iterations in range(1, max_iterations + 1):
                yx[0, :] = ring
                yx[-1, :] = ring
                yx[:, 0] = ring
                yx[:, -1] = ring
                _, conv, last = relax_device(yx, self.grid.dx, self.grid.dy, cfg.critical_angle_deg, 1,
                                             cfg.tolerance, runtime=runtime)
                if conv:
                    converged = True
                    break
            stable = yx[1:-1, 1:-1].cpu().numpy().astype(np.float64)
        before = self.grid.compute_volume(initial)
        after = self.grid.compute_volume(stable)
        if cfg.boundary_condition == "open":
            if after > before + cfg.conservation_tolerance_m3:
                raise RuntimeError(
                    "[MinimumSlopeAdapter] open boundary created terrain volume; "
                    f"before={before}, after={after}, shape={initial.shape}")
            outflow = max(before - after, 0.0)
        else:
            outflow = 0.0
        balance = before - after - outflow
        if cfg.boundary_condition == "closed" and abs(balance) > cfg.conservation_tolerance_m3:
            raise RuntimeError(
                "[MinimumSlopeAdapter] closed-boundary volume conservation failed; "
                f"shape={initial.shape}, before={before}, after={after}, error={balance}, "
                f"tolerance={cfg.conservation_tolerance_m3}")
        result = RelaxationResult(
            stable, (initial.astype(np.float32), stable.astype(np.float32)), iterations, before, after,
            outflow, converged,
            {"adapter": "HipMinimumSlopeAdapter", "legacy_entrypoint": "slope_model.relax_critical_slope",
             "canonical_axis_order": "yx", "legacy_axis_order": "xy",
             "boundary_condition": cfg.boundary_condition, "critical_angle_deg": cfg.critical_angle_deg,
             "maximum_iterations": max_iterations, "volume_balance_error_m3": balance,
             "max_excess_last_iteration_m": last, "precision": self.precision})
        self._last_result = result
        return result

    def step(self, heightmap) -> RelaxationResult:
        return self._solve(heightmap, 1)

    def solve(self, heightmap) -> RelaxationResult:
        return self._solve(heightmap, self.config.max_iterations)


# ---------------------------------------------------------------------------
# compact tile frontier (sparse_tile_slope.py:20-230) and its device operator
# ---------------------------------------------------------------------------
@dataclass(frozen=True)
class CompactActiveEdgeBatch:
    tile_ids: np.ndarray
    a_i: np.ndarray
    a_j: np.ndarray
    b_i: np.ndarray
    b_j: np.ndarray
    direction_index: int
    phase: int

    def __post_init__(self) -> None:
        arrays = []
        for value in (self.tile_ids, self.a_i, self.a_j, self.b_i, self.b_j):
            a = np.ascontiguousarray(value, dtype=np.int32)
            a.setflags(write=False)
            arrays.append(a)
        if len({a.size for a in arrays}) != 1:
            raise ValueError("[SparseTileMiniSlope] edge buffer lengths differ")
        for name, a in zip(("tile_ids", "a_i", "a_j", "b_i", "b_j"), arrays):
            object.__setattr__(self, name, a)

    @property
    def edge_count(self) -> int:
        return int(self.a_i.size)


class CompactTileFrontier:
    """Unique edge-owner batches for compact active tile ids."""

    def __init__(self, shape: tuple[int, int], tile_size: int) -> None:
        self.shape = int(shape[0]), int(shape[1])
        self.tile_size = int(tile_size)
        self.tile_shape = ((self.shape[0] + tile_size - 1) // tile_size,
                           (self.shape[1] + tile_size - 1) // tile_size)
        li, lj = np.indices((tile_size, tile_size), dtype=np.int32)
        self._local_i, self._local_j = li.ravel(), lj.ravel()
        self._cache: OrderedDict = OrderedDict()

    @property
    def tile_count(self) -> int:
        return int(self.tile_shape[0] * self.tile_shape[1])

    def tiles_around_cells(self, i, j, *, stencil_cells: int = 1) -> np.ndarray:
        si = np.asarray(i, dtype=np.int64).ravel()
        sj = np.asarray(j, dtype=np.int64).ravel()
        if si.size == 0:
            return np.empty(0, dtype=np.int32)
        off = np.arange(-stencil_cells, stencil_cells + 1, dtype=np.int64)
        ci = np.repeat((si[:, None] + off[None, :]), off.size, axis=1).ravel()
        cj = np.tile((sj[:, None] + off[None, :]), (1, off.size)).ravel()
        valid = (ci >= 0) & (ci < self.shape[0]) & (cj >= 0) & (cj < self.shape[1])
        return np.unique(((ci[valid] // self.tile_size) * self.tile_shape[1]
                          + cj[valid] // self.tile_size).astype(np.int32))

    def edge_batch(self, active_tile_ids, *, direction_index: int, di: int, dj: int,
                   phase: int) -> CompactActiveEdgeBatch:
        tile_ids = np.unique(np.asarray(active_tile_ids, dtype=np.int32))
        key = (tile_ids.tobytes(), int(di), int(dj), int(phase))
        cached = self._cache.get(key)
        if cached is not None:
            self._cache.move_to_end(key)
            return cached
        if tile_ids.size == 0:
            e = np.empty(0, dtype=np.int32)
            batch = CompactActiveEdgeBatch(e, e, e, e, e, direction_index, phase)
        else:
            tr = tile_ids // self.tile_shape[1]
            tc = tile_ids % self.tile_shape[1]
            a_i = (tr[:, None] * self.tile_size + self._local_i[None, :]).ravel()
            a_j = (tc[:, None] * self.tile_size + self._local_j[None, :]).ravel()
            owners = np.repeat(tile_ids, self._local_i.size)
            b_i, b_j = a_i + int(di), a_j + int(dj)
            coord = a_i if di else a_j
            valid = ((a_i < self.shape[0]) & (a_j < self.shape[1]) & (b_i >= 0)
                     & (b_i < self.shape[0]) & (b_j >= 0) & (b_j < self.shape[1])
                     & ((coord & 1) == int(phase)))
            batch = CompactActiveEdgeBatch(owners[valid], a_i[valid], a_j[valid], b_i[valid],
                                           b_j[valid], direction_index, phase)
        self._cache[key] = batch
        self._cache.move_to_end(key)
        while len(self._cache) > 32:
            self._cache.popitem(last=False)
        return batch


def neighbor_pairs(dx: float, dy: float) -> list[tuple[int, int, float]]:
    """slope_model._neighbor_pairs (slope_model.py:744-750)."""
    return [(1, 0, float(dx)), (0, 1, float(dy)), (1, 1, float(np.hypot(dx, dy))),
            (1, -1, float(np.hypot(dx, dy)))]


@dataclass(frozen=True)
class FrontierRoundsResult:
    rounds: int
    converged: bool
    active_tile_ids: np.ndarray
    moved_edge_count: int
    cleared_reached: bool


@dataclass(frozen=True)
class FrontierPhaseResult:
    moved_edge_count: int
    triggered_tile_ids: np.ndarray
    unstable_owner_tile_ids: np.ndarray


class HipCompactActiveEdgeOperator:
    """GPU-resident height/reach state with cached device edge buffers."""

    backend_name = "GPU_HIP_COMPACT_ACTIVE_EDGE_FRONTIER"

    def __init__(self, shape, *, device: str = "cuda:0", runtime: HipRuntime | None = None,
                 tile_size: int = 64, batch_cache_limit: int = 32) -> None:
        torch = __import__("torch")
        self.shape = int(shape[0]), int(shape[1])
        self.tile_size = int(tile_size)
        if self.tile_size < 1:
            raise ValueError("[WarpMiniSlope] tile_size must be positive")
        self.batch_cache_limit = int(batch_cache_limit)
        if self.batch_cache_limit < 1:
            raise ValueError("[WarpMiniSlope] batch_cache_limit must be positive")
        self.tile_shape = ((self.shape[0] + self.tile_size - 1) // self.tile_size,
                           (self.shape[1] + self.tile_size - 1) // self.tile_size)
        self.tile_count = self.tile_shape[0] * self.tile_shape[1]
        self.runtime = runtime or HipRuntime(device)
        self._height_key, self._weights_key, self._reached_key = "height", "frontier_weights", "reached"
        self._cache: OrderedDict = OrderedDict()
        self._initialized = False
        dev = self.runtime.device
        self._moved = torch.zeros(self.tile_count, dtype=torch.int32, device=dev)
        self._unstable = torch.zeros(self.tile_count, dtype=torch.int32, device=dev)

    def initialize(self, height_xy, reached_xy, vertex_weights_xy_m2=None) -> None:
        torch = __import__("torch")
        h = np.asarray(height_xy, dtype=np.float64)
        r = np.asarray(reached_xy, dtype=np.int32)
        if h.shape != self.shape or r.shape != self.shape:
            raise ValueError("[WarpMiniSlope] resident state shape mismatch")
        w = np.ones(self.shape) if vertex_weights_xy_m2 is None else np.asarray(vertex_weights_xy_m2, np.float64)
        if w.shape != self.shape or not np.all(np.isfinite(w)) or np.any(w <= 0.0):
            raise ValueError("[WarpMiniSlope] vertex weights must be finite/positive")
        self.runtime.upload(self._height_key, h.ravel(), torch.float64)
        self.runtime.upload(self._weights_key, w.ravel(), torch.float64)
        self.runtime.upload(self._reached_key, r.ravel(), torch.int32)
        self._initialized = True

    def bind_device_state(self, state) -> None:
        if getattr(state, "runtime", None) is not self.runtime:
            raise ValueError("[WarpMiniSlope] state/runtime ownership mismatch")
        if tuple(getattr(state, "shape", ())) != self.shape:
            raise ValueError("[WarpMiniSlope] state shape mismatch")
        if int(getattr(state, "tile_size", -1)) != self.tile_size:
            raise ValueError("[WarpMiniSlope] state/frontier tile-size mismatch")
        if getattr(state, "precision", "fp64") != "fp64":
            raise ValueError("[WarpMiniSlope] the weighted frontier sweep binds fp64 state")
        self._height_key, self._weights_key, self._reached_key = "b_eff", "weights", "frontier_reached"
        self._initialized = True

    def _resident_batch(self, batch: CompactActiveEdgeBatch):
        torch = __import__("torch")
        key = (batch.a_i.tobytes() + batch.a_j.tobytes(), batch.direction_index, batch.phase)
        cached = self._cache.get(key)
        if cached is not None:
            self._cache.move_to_end(key)
            return cached
        dev = self.runtime.device
        # the batch's index arrays may be read-only views: upload copies
        arrays = tuple(torch.from_numpy(np.array(v, order="C", copy=True)).to(dev)
                       for v in (batch.tile_ids, batch.a_i, batch.a_j, batch.b_i, batch.b_j))
        self._cache[key] = arrays
        while len(self._cache) > self.batch_cache_limit:
            self._cache.popitem(last=False)
        return arrays

    def relax_phase(self, batch: CompactActiveEdgeBatch, *, critical_difference_m: float) -> FrontierPhaseResult:
        if not self._initialized:
            raise RuntimeError("[WarpMiniSlope] initialize resident state first")
        empty = np.empty(0, dtype=np.int32)
        if batch.edge_count == 0:
            return FrontierPhaseResult(0, empty, empty)
        rt = self.runtime
        rt.bind_stream()
        tile, ai, aj, bi, bj = self._resident_batch(batch)
        moved = ctypes.c_int32(0)
        rc = rt.lib.mv2_frontier_relax(
            rt.ctx, rt.ptr(self._height_key), rt.ptr(self._weights_key), rt.ptr(self._reached_key),
            ai.data_ptr(), aj.data_ptr(), bi.data_ptr(), bj.data_ptr(), batch.edge_count,
            self.shape[0], self.shape[1], self.tile_size, self.tile_shape[1], self.tile_count,
            float(critical_difference_m), self._moved.data_ptr(), ctypes.byref(moved))
        rt.check(rc)
        flags = self._moved.cpu().numpy()
        return FrontierPhaseResult(int(moved.value), np.flatnonzero(flags).astype(np.int32), empty)

    def scan_phase(self, batch: CompactActiveEdgeBatch, *, critical_difference_m: float,
                   tolerance_m: float) -> FrontierPhaseResult:
        if not self._initialized:
            raise RuntimeError("[WarpMiniSlope] initialize resident state first")
        empty = np.empty(0, dtype=np.int32)
        if batch.edge_count == 0:
            return FrontierPhaseResult(0, empty, empty)
        rt = self.runtime
        rt.bind_stream()
        tile, ai, aj, bi, bj = self._resident_batch(batch)
        rc = rt.lib.mv2_frontier_scan(
            rt.ctx, rt.ptr(self._height_key), rt.ptr(self._reached_key), tile.data_ptr(),
            ai.data_ptr(), aj.data_ptr(), bi.data_ptr(), bj.data_ptr(), batch.edge_count,
            self.shape[1], self.tile_count, float(critical_difference_m), float(tolerance_m),
            self._unstable.data_ptr())
        rt.check(rc)
        flags = self._unstable.cpu().numpy()
        return FrontierPhaseResult(0, empty, np.flatnonzero(flags).astype(np.int32))

    def run_phase(self, batch, *, critical_difference_m: float, tolerance_m: float) -> FrontierPhaseResult:
        moved = self.relax_phase(batch, critical_difference_m=critical_difference_m)
        unstable = self.scan_phase(batch, critical_difference_m=critical_difference_m,
                                   tolerance_m=tolerance_m)
        return FrontierPhaseResult(moved.moved_edge_count, moved.triggered_tile_ids,
                                   unstable.unstable_owner_tile_ids)

    def advance_rounds(self, active_tiles, *, neighbors, critical_tangent: float, tolerance_m: float,
                       max_rounds: int) -> FrontierRoundsResult:
        """Residual projection rounds on the device (v2_physics_core.py:1183-1254).

        ``active_tiles``: tile ids (host) or a device int32 tensor of
        ``tile_count`` flags, which is then updated in place to the tiles
        still active.  ``neighbors``: (di, dj, distance) per direction, the
        critical difference of a direction being ``critical_tangent *
        distance``.  Stops when the active set is empty (``reached`` is then
        zeroed) or after ``max_rounds`` rounds."""
        torch = __import__("torch")
        if not self._initialized:
            raise RuntimeError("[WarpMiniSlope] initialize resident state first")
        pairs = list(neighbors)
        if not 1 <= len(pairs) <= 4:
            raise ValueError("[WarpMiniSlope] 1 to 4 neighbour directions")
        if int(max_rounds) < 0:
            raise ValueError("[WarpMiniSlope] max_rounds must be non-negative")
        rt = self.runtime
        if isinstance(active_tiles, torch.Tensor):
            flags = active_tiles
            if (flags.dtype != torch.int32 or flags.numel() != self.tile_count
                    or flags.device != rt.device or not flags.is_contiguous()):
                raise ValueError("[WarpMiniSlope] tile flags must be a contiguous device int32 "
                                 "tensor of tile_count entries")
        else:
            ids = np.unique(np.asarray(active_tiles, dtype=np.int64).ravel())
            if ids.size and (ids[0] < 0 or ids[-1] >= self.tile_count):
                raise ValueError("[WarpMiniSlope] tile ID outside grid")
            host = np.zeros(self.tile_count, dtype=np.int32)
            host[ids] = 1
            flags = torch.from_numpy(host).to(rt.device)
        cfg = _lib.FrontierRoundConfig()
        cfg.rows, cfg.cols = self.shape
        cfg.tile_size, cfg.tiles_x, cfg.tile_count = self.tile_size, self.tile_shape[1], self.tile_count
        cfg.n_directions = len(pairs)
        for d, (di, dj, distance) in enumerate(pairs):
            cfg.di[d], cfg.dj[d] = int(di), int(dj)
            cfg.critical_difference[d] = float(critical_tangent) * float(distance)
        cfg.tolerance = float(tolerance_m)
        cfg.max_rounds = int(max_rounds)
        out = _lib.FrontierRoundResult()
        rt.bind_stream()
        rc = rt.lib.mv2_frontier_rounds(rt.ctx, ctypes.byref(cfg), rt.ptr(self._height_key),
                                        rt.ptr(self._weights_key), rt.ptr(self._reached_key),
                                        flags.data_ptr(), ctypes.byref(out))
        rt.check(rc)
        active = (np.empty(0, dtype=np.int32) if out.converged
                  else np.flatnonzero(flags.cpu().numpy()).astype(np.int32))
        return FrontierRoundsResult(int(out.rounds), bool(out.converged), active, int(out.moved_edges),
                                    bool(out.cleared_reached))

    def download_state(self):
        h = self.runtime.download(self._height_key).reshape(self.shape)
        r = self.runtime.download(self._reached_key).reshape(self.shape).astype(bool)
        return h, r

    def diagnostics(self) -> dict:
        return self.runtime.diagnostics(self.backend_name)


WarpCompactActiveEdgeOperator = HipCompactActiveEdgeOperator


class GpuResidualFrontier:
    """Residual MiniSlope projection of the frame scheduler on the device.

    Mirrors V2PhysicsCore's GPU frontier state (pkg/runtime/v2_physics_core.py:
    338-343, 966-973, 1142-1254): cell seeds mark ``frontier_reached`` and
    activate the tiles around them, residual seed tiles join the active set,
    and ``advance()`` runs the bounded rounds -- here all on the device --
    raising the reference's NUMERICAL_NONCONVERGENCE errors."""

    def __init__(self, operator: HipCompactActiveEdgeOperator, *, stop_angle_deg: float, dx: float,
                 dy: float, tolerance_m: float = 1.0e-8, max_rounds: int = 4096,
                 numerical_safety_max_iterations: int = 1_000_000) -> None:
        torch = __import__("torch")
        self.operator = operator
        self.critical_tangent = float(np.tan(np.deg2rad(stop_angle_deg)))
        self.neighbors = neighbor_pairs(dx, dy)
        self.tolerance_m = float(tolerance_m)
        self.max_rounds = int(max_rounds)
        self.numerical_safety_max_iterations = int(numerical_safety_max_iterations)
        self.iterations = 0
        self.frontier = CompactTileFrontier(operator.shape, operator.tile_size)
        self._flags = torch.zeros(operator.tile_count, dtype=torch.int32, device=operator.runtime.device)

    @property
    def active_tile_ids(self) -> np.ndarray:
        return np.flatnonzero(self._flags.cpu().numpy()).astype(np.int32)

    def add_tiles(self, tile_ids) -> None:
        """Union of residual seed tiles into the active set (v2_physics_core.py:966-973)."""
        ids = np.unique(np.asarray(tile_ids, dtype=np.int64).ravel())
        if ids.size == 0:
            return
        if ids[0] < 0 or ids[-1] >= self.operator.tile_count:
            raise ValueError("[WarpMiniSlope] tile ID outside grid")
        torch = __import__("torch")
        self._flags[torch.from_numpy(ids).to(self._flags.device)] = 1

    def seed_cells(self, flat_indices) -> None:
        """_seed_gpu_residual_frontier (v2_physics_core.py:1152-1170)."""
        torch = __import__("torch")
        indices = np.unique(np.asarray(flat_indices, dtype=np.int64).ravel())
        if indices.size == 0:
            return
        rows_total, cols_total = self.operator.shape
        if indices[0] < 0 or indices[-1] >= rows_total * cols_total:
            raise ValueError("[WarpMiniSlope] seed index outside grid")
        reached = self.operator.runtime.arrays[self.operator._reached_key]
        reached[torch.from_numpy(indices).to(reached.device)] = 1
        rows, cols = np.divmod(indices, cols_total)
        self.add_tiles(self.frontier.tiles_around_cells(rows, cols))

    def clear(self) -> None:
        """Discard the candidate seeds and the reach mask (v2_physics_core.py:991-1000)."""
        self._flags.zero_()
        self.operator.runtime.arrays[self.operator._reached_key].zero_()

    def advance(self) -> FrontierRoundsResult:
        """_advance_gpu_residual_frontier (v2_physics_core.py:1172-1254)."""
        safety_left = self.numerical_safety_max_iterations - self.iterations
        budget = max(0, min(self.max_rounds, safety_left))
        result = self.operator.advance_rounds(self._flags, neighbors=self.neighbors,
                                              critical_tangent=self.critical_tangent,
                                              tolerance_m=self.tolerance_m, max_rounds=budget)
        self.iterations += result.rounds
        if not result.converged:
            if safety_left < self.max_rounds:
                raise RuntimeError("NUMERICAL_NONCONVERGENCE: GPU residual MiniSlope exceeded "
                                   "the numerical safety limit")
            raise RuntimeError("NUMERICAL_NONCONVERGENCE: V3 GPU residual projection exceeded "
                               f"{result.rounds} bounded numerical rounds")
        return result
