"""Configuration and geometry types that cross the Mobile V2 boundary.

Field names, defaults and validation messages follow the reference so that a
caller's configuration objects carry over unchanged:
  MobileV2Config                 pkg/experimental/mobile_v2_reference.py:16-39
  MaterialScenario               pkg/bulk_state/models.py:51-110
  TerrainGrid (subset)           pkg/terrain/terrain_grid.py:26-113, :171-195
  TerrainVolumeIntegrator        pkg/bulk_state/volume_integrator.py:34-140
  SolverConfig (relaxation part) pkg/config/loader.py:137-272
  LargeAvalancheTransitionConfig pkg/bulk_interaction/large_avalanche.py:38-144
(pkg/ = src/isaac_bulk_pipeline/ in the reference repository.)
"""

from __future__ import annotations

from dataclasses import dataclass, field, fields, replace
from enum import Enum

import numpy as np

UNCALIBRATED_LABEL = "UNCALIBRATED_SCENARIO_PARAMETER"
LITERATURE_REDUCED_ORDER_UNCALIBRATED = "LITERATURE_BASED_REDUCED_ORDER_UNCALIBRATED"


@dataclass(frozen=True)
class MobileV2Config:
    dx_m: float = 0.05
    dy_m: float = 0.05
    gravity_m_s2: float = 9.81
    earth_pressure_coefficient: float = 0.45
    cfl: float = 0.18
    dry_tolerance_m: float = 1.0e-10
    start_angle_deg: float = 34.0
    stop_angle_deg: float = 30.0
    basal_friction_coefficient: float = 0.55
    maximum_substeps: int = 4096
    # Explicit opt-in for a two-dimensional CFL above the reference's 0.25
    # guard (positivity is then checked, not guaranteed).
    allow_cfl_above_quarter: bool = False

    def __post_init__(self) -> None:
        values = np.asarray([
            self.dx_m, self.dy_m, self.gravity_m_s2, self.earth_pressure_coefficient,
            self.cfl, self.dry_tolerance_m, self.basal_friction_coefficient,
        ], dtype=np.float64)
        if not np.all(np.isfinite(values)) or np.any(values <= 0.0):
            raise ValueError("MobileV2 configuration must be finite/positive")
        if self.cfl > 0.25 and not self.allow_cfl_above_quarter:
            raise ValueError("MobileV2 two-dimensional CFL must be <= 0.25")

    @property
    def tan_start(self) -> float:
        return float(np.tan(np.deg2rad(self.start_angle_deg)))


@dataclass(frozen=True)
class MaterialScenario:
    name: str
    assumed_bulk_density_kg_m3: float
    internal_friction_angle_deg: float
    cohesion_proxy_pa: float
    tool_friction_coefficient: float
    start_angle_deg: float
    stop_angle_deg: float
    mobile_friction_coefficient: float
    calibration_status: str = UNCALIBRATED_LABEL

    def __post_init__(self) -> None:
        if not isinstance(self.name, str) or not self.name.strip():
            raise ValueError("[MaterialScenario] name must be non-empty")
        if self.calibration_status != UNCALIBRATED_LABEL:
            raise ValueError(
                "[MaterialScenario] accepts only explicitly uncalibrated parameters "
                f"labelled {UNCALIBRATED_LABEL!r}")
        for name in ("assumed_bulk_density_kg_m3", "internal_friction_angle_deg",
                     "cohesion_proxy_pa", "tool_friction_coefficient", "start_angle_deg",
                     "stop_angle_deg", "mobile_friction_coefficient"):
            value = float(getattr(self, name))
            if not np.isfinite(value) or value < 0.0:
                raise ValueError(f"[BulkState] {name} must be finite and non-negative")
            object.__setattr__(self, name, value)
        if self.assumed_bulk_density_kg_m3 <= 0.0:
            raise ValueError("[MaterialScenario] assumed density must be > 0")
        for name in ("internal_friction_angle_deg", "start_angle_deg", "stop_angle_deg"):
            if not getattr(self, name) < 90.0:
                raise ValueError(f"[MaterialScenario] {name} must be < 90 degrees")
        if self.start_angle_deg <= self.stop_angle_deg:
            raise ValueError(
                "[MaterialScenario] start_angle_deg must exceed stop_angle_deg to define hysteresis")


@dataclass
class TerrainGrid:
    """Canonical ``H[y, x]`` grid (the geometry subset the hot path uses)."""

    nx: int
    ny: int
    dx: float
    dy: float
    origin_x: float = 0.0
    origin_y: float = 0.0
    terrain_prim_path: str = "/World/Terrain"
    valid_mask: np.ndarray | None = None

    def __post_init__(self) -> None:
        if not isinstance(self.nx, int) or not isinstance(self.ny, int):
            raise TypeError("[TerrainGrid] nx and ny must be integers")
        if self.nx < 2 or self.ny < 2:
            raise ValueError(f"[TerrainGrid] nx and ny must be >= 2; nx={self.nx}, ny={self.ny}")
        for name in ("dx", "dy", "origin_x", "origin_y"):
            if not np.isfinite(getattr(self, name)):
                raise ValueError(f"[TerrainGrid] {name} must be finite; value={getattr(self, name)}")
        if self.dx <= 0.0 or self.dy <= 0.0:
            raise ValueError(
                f"[TerrainGrid] cell spacing must be positive metres; dx={self.dx}, dy={self.dy}")
        if not self.terrain_prim_path.startswith("/"):
            raise ValueError("[TerrainGrid] terrain_prim_path must be an absolute USD Prim path")
        if self.valid_mask is not None:
            mask = np.asarray(self.valid_mask, dtype=bool)
            if mask.shape != self.shape:
                raise ValueError("[TerrainGrid] valid_mask shape must equal (ny, nx)")
            self.valid_mask = np.ascontiguousarray(mask.copy())

    @property
    def shape(self) -> tuple[int, int]:
        return self.ny, self.nx

    @property
    def cell_area(self) -> float:
        return float(self.dx * self.dy)

    def validate_heightmap(self, heightmap) -> np.ndarray:
        height = np.asarray(heightmap)
        if height.shape != self.shape:
            raise ValueError(
                "[TerrainGrid] heightmap must use H[y,x] with shape (ny,nx); "
                f"expected={self.shape}, received={height.shape}, prim_path={self.terrain_prim_path}")
        if not np.issubdtype(height.dtype, np.number):
            raise TypeError(f"[TerrainGrid] heightmap dtype must be numeric; dtype={height.dtype}")
        if not np.all(np.isfinite(height)):
            raise ValueError("[TerrainGrid] heightmap contains NaN or Inf")
        return height

    def compute_volume(self, heightmap) -> float:
        height = self.validate_heightmap(heightmap)
        values = height if self.valid_mask is None else height[self.valid_mask]
        return float(values.sum(dtype=np.float64) * self.cell_area)


class SurfaceTopology(str, Enum):
    TRIANGLE_A_C = "TRIANGLE_A_C"
    QUAD_TRAPEZOID = "QUAD_TRAPEZOID"


def triangle_vertex_weights(ny: int, nx: int, dx: float, dy: float) -> np.ndarray:
    """Triangle-A-C control-area weights (volume_integrator.py:76-91)."""
    weights = np.zeros((ny, nx), dtype=np.float64)
    factor = dx * dy / 6.0
    weights[:-1, :-1] += 2.0 * factor
    weights[:-1, 1:] += factor
    weights[1:, 1:] += 2.0 * factor
    weights[1:, :-1] += factor
    return weights


@dataclass(frozen=True)
class TerrainVolumeIntegrator:
    nx: int
    ny: int
    dx_m: float
    dy_m: float
    topology: SurfaceTopology = SurfaceTopology.TRIANGLE_A_C
    _weights: np.ndarray = field(init=False, repr=False, compare=False)

    def __post_init__(self) -> None:
        if self.nx < 2 or self.ny < 2:
            raise ValueError("[TerrainVolumeIntegrator] nx and ny must be >= 2")
        if not (self.dx_m > 0.0 and self.dy_m > 0.0):
            raise ValueError("[TerrainVolumeIntegrator] dx_m and dy_m must be finite and positive")
        topology = SurfaceTopology(self.topology)
        if topology is SurfaceTopology.TRIANGLE_A_C:
            w = triangle_vertex_weights(self.ny, self.nx, self.dx_m, self.dy_m)
        else:
            w = np.zeros((self.ny, self.nx))
            f = self.dx_m * self.dy_m / 4.0
            w[:-1, :-1] += f
            w[:-1, 1:] += f
            w[1:, 1:] += f
            w[1:, :-1] += f
        w.setflags(write=False)
        object.__setattr__(self, "topology", topology)
        object.__setattr__(self, "_weights", w)

    @classmethod
    def from_grid(cls, grid: TerrainGrid, *, topology=SurfaceTopology.TRIANGLE_A_C):
        return cls(int(grid.nx), int(grid.ny), float(grid.dx), float(grid.dy), topology)

    @property
    def shape(self) -> tuple[int, int]:
        return self.ny, self.nx

    @property
    def vertex_weights_m2(self) -> np.ndarray:
        out = np.array(self._weights, copy=True)
        out.setflags(write=False)
        return out

    def integrate(self, height_m) -> float:
        h = np.asarray(height_m, dtype=np.float64)
        if h.shape != self.shape:
            raise ValueError("[TerrainVolumeIntegrator] field shape mismatch")
        return float(np.sum(h * self._weights, dtype=np.float64))


@dataclass(frozen=True)
class SolverConfig:
    """Relaxation settings (the MiniSlope fields of the reference SolverConfig)."""

    critical_angle_deg: float = 34.0
    max_iterations: int = 1000
    numerical_safety_max_iterations: int = 1_000_000
    tolerance: float = 1e-8
    boundary_condition: str = "closed"
    boundary_height_m: float = 0.0
    conservation_tolerance_m3: float = 1e-8

    def __post_init__(self) -> None:
        if not 0.0 < self.critical_angle_deg < 90.0:
            raise ValueError("critical_angle_deg must be between 0 and 90")
        if self.max_iterations < 1:
            raise ValueError("[Config:Solver] max_iterations must be >= 1")
        if not (self.tolerance >= 0.0 and np.isfinite(self.tolerance)):
            raise ValueError("[Config:Solver] tolerance must be finite and non-negative")
        if self.boundary_condition not in ("closed", "open"):
            raise ValueError("[Config:Solver] boundary_condition must be 'closed' or 'open'")


@dataclass(frozen=True)
class LargeAvalancheTransitionConfig:
    minimum_connected_cells: int = 32
    minimum_connected_area_m2: float = 0.08
    minimum_mobilizable_volume_m3: float = 0.004
    minimum_mean_excess_start_deg: float = 0.0
    persistence_time_s: float = 0.12
    mobilization_depth_m: float = 0.08
    gravity_velocity_length_m: float = 0.25
    velocity_efficiency: float = 0.45
    mobile_activity_speed_m_s: float = 0.05
    mobile_activity_volume_m3: float = 1.0e-5
    dry_tolerance_m: float = 1.0e-8
    settled_mobile_volume_m3: float = 1.0e-5
    settled_speed_m_s: float = 0.05
    connectivity: int = 8
    parameter_basis: str = LITERATURE_REDUCED_ORDER_UNCALIBRATED
    sensitivity_case: str = "nominal_uncalibrated"

    def __post_init__(self) -> None:
        if not isinstance(self.minimum_connected_cells, int) or self.minimum_connected_cells < 1:
            raise ValueError("[LargeAvalanche] minimum_connected_cells must be >= 1")
        if self.connectivity not in (4, 8):
            raise ValueError("[LargeAvalanche] connectivity must be 4 or 8")
        numeric = np.asarray([
            self.minimum_connected_area_m2, self.minimum_mobilizable_volume_m3,
            self.minimum_mean_excess_start_deg, self.persistence_time_s,
            self.mobilization_depth_m, self.gravity_velocity_length_m, self.velocity_efficiency,
            self.mobile_activity_speed_m_s, self.mobile_activity_volume_m3, self.dry_tolerance_m,
            self.settled_mobile_volume_m3, self.settled_speed_m_s,
        ], dtype=np.float64)
        if not np.all(np.isfinite(numeric)) or np.any(numeric < 0.0):
            raise ValueError("[LargeAvalanche] numeric configuration must be finite/non-negative")
        if self.persistence_time_s <= 0.0:
            raise ValueError("[LargeAvalanche] persistence_time_s must be positive")
        if self.mobilization_depth_m <= 0.0 or self.gravity_velocity_length_m <= 0.0:
            raise ValueError("[LargeAvalanche] failure depth/velocity length must be positive")
        if not 0.0 < self.velocity_efficiency <= 1.0:
            raise ValueError("[LargeAvalanche] velocity_efficiency must lie in (0, 1]")
        if self.parameter_basis != LITERATURE_REDUCED_ORDER_UNCALIBRATED:
            raise ValueError("[LargeAvalanche] transition parameters must remain explicitly "
                             "literature-based reduced-order / uncalibrated")
        if not self.sensitivity_case:
            raise ValueError("[LargeAvalanche] sensitivity_case must be non-empty")

    def sensitivity_ensemble(self, relative_fraction: float = 0.25):
        f = float(relative_fraction)
        if not np.isfinite(f) or not 0.0 < f < 1.0:
            raise ValueError("[LargeAvalanche] sensitivity fraction must lie in (0, 1)")
        return (
            replace(self, persistence_time_s=self.persistence_time_s * (1.0 - f),
                    minimum_connected_area_m2=self.minimum_connected_area_m2 * (1.0 - f),
                    minimum_mobilizable_volume_m3=self.minimum_mobilizable_volume_m3 * (1.0 - f),
                    mobilization_depth_m=self.mobilization_depth_m * (1.0 + f),
                    velocity_efficiency=min(1.0, self.velocity_efficiency * (1.0 + f)),
                    sensitivity_case="more_mobile_uncalibrated"),
            replace(self, sensitivity_case="nominal_uncalibrated"),
            replace(self, persistence_time_s=self.persistence_time_s * (1.0 + f),
                    minimum_connected_area_m2=self.minimum_connected_area_m2 * (1.0 + f),
                    minimum_mobilizable_volume_m3=self.minimum_mobilizable_volume_m3 * (1.0 + f),
                    mobilization_depth_m=self.mobilization_depth_m * (1.0 - f),
                    velocity_efficiency=self.velocity_efficiency * (1.0 - f),
                    sensitivity_case="less_mobile_uncalibrated"),
        )

    @classmethod
    def from_mapping(cls, value: dict) -> "LargeAvalancheTransitionConfig":
        if not isinstance(value, dict):
            raise TypeError("[LargeAvalanche] configuration must be a mapping")
        names = {f.name for f in fields(cls)}
        unknown = sorted(set(value) - names)
        if unknown:
            raise ValueError(f"[LargeAvalanche] unknown configuration fields: {unknown}")
        return cls(**value)
