"""Mobile V2 granular-terrain step on the MI355X.

Drop-in counterparts of the reference's device solvers:
  HipProductionMobileV2Solver  <- WarpProductionMobileV2Solver
                                  (pkg/bulk_interaction/warp_mobile_v2.py:91-253)
  HipMobileV2ReferenceSolver   <- WarpMobileV2ReferenceSolver
                                  (pkg/experimental/mobile_v2_warp.py:238-344)
with the same records (WarpMobileStep, warp_mobile_layer.py:349-371;
MobileV2Step, mobile_v2_reference.py:63-90).  The whole step -- ledger,
static early exit, every CFL substep and the closing ledger -- is one HIP
launch (libmv2 ``mv2_step``); the host reads one fixed-size record.
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

from . import _lib
from .bulk_state import DeviceBulkState
from .config import MobileV2Config, TerrainGrid


@dataclass(frozen=True)
class WarpMobileStep:
    volume_before_m3: float
    volume_after_m3: float
    momentum_before_terrain_kg_m_s: np.ndarray
    momentum_after_terrain_kg_m_s: np.ndarray
    gravity_pressure_impulse_terrain_ns: np.ndarray
    basal_friction_impulse_terrain_ns: np.ndarray
    tool_impulse_on_mobile_terrain_ns: np.ndarray
    numerical_dissipative_impulse_terrain_ns: np.ndarray
    maximum_speed_m_s: float
    substeps: int
    cfl_limited: bool
    kinetic_energy_before_j: float
    kinetic_energy_after_j: float
    gravity_pressure_work_j: float
    basal_friction_work_j: float
    tool_work_j: float
    transport_numerical_energy_residual_j: float
    donor_export_m3: float
    receiver_import_m3: float
    transport_mass_residual_m3: float
    advected_momentum_crossings_kg_m_s: np.ndarray


@dataclass(frozen=True)
class MobileV2State:
    b_eff_m: np.ndarray
    h_m: np.ndarray
    q_m2_s: np.ndarray

    def __post_init__(self) -> None:
        b = np.asarray(self.b_eff_m, dtype=np.float64)
        h = np.asarray(self.h_m, dtype=np.float64)
        q = np.asarray(self.q_m2_s, dtype=np.float64)
        if b.ndim != 2 or h.shape != b.shape or q.shape != b.shape + (2,):
            raise ValueError("MobileV2 state shape mismatch")
        if not np.all(np.isfinite(b)) or not np.all(np.isfinite(h)) or not np.all(np.isfinite(q)):
            raise ValueError("MobileV2 state must be finite")
        if np.any(h < -1.0e-14):
            raise ValueError("MobileV2 h must be nonnegative")

    @property
    def H_free_m(self) -> np.ndarray:
        return np.asarray(self.b_eff_m) + np.asarray(self.h_m)


@dataclass(frozen=True)
class MobileV2Step:
    state: MobileV2State
    dt_s: float
    substeps: int
    maximum_cfl: float
    mass_before_m3: float
    mass_after_m3: float
    mass_residual_m3: float
    momentum_before_m4_s: np.ndarray
    momentum_after_m4_s: np.ndarray
    conservative_topography_wall_impulse_m4_s: np.ndarray
    friction_impulse_m4_s: np.ndarray
    external_impulse_m4_s: np.ndarray
    momentum_accounting_residual_m4_s: np.ndarray
    kinetic_before_j_per_density: float
    kinetic_after_j_per_density: float
    gravitational_before_j_per_density: float
    gravitational_after_j_per_density: float
    internal_before_j_per_density: float
    internal_after_j_per_density: float
    energy_before_j_per_density: float
    energy_after_j_per_density: float
    friction_dissipation_j_per_density: float
    external_work_j_per_density: float
    entrainment_deposition_energy_transfer_j_per_density: float
    numerical_energy_residual_j_per_density: float
    minimum_h_m: float


def _step_config(config: MobileV2Config, dt: float, *, density: float, limit_mode: int,
                 weight_mode: int, external=(0.0, 0.0), diagnostics: int = 0) -> _lib.StepConfig:
    """C step configuration.  ``diagnostics`` 1 keeps only the five per-substep
    sums the production record reads (warp_mobile_v2.py:220-241)."""
    return _lib.StepConfig(
        dt=float(dt), gravity=config.gravity_m_s2, earth_pressure=config.earth_pressure_coefficient,
        cfl=config.cfl, dry_tolerance=config.dry_tolerance_m,
        basal_friction=config.basal_friction_coefficient, tan_start=config.tan_start,
        external_ax=float(external[0]), external_ay=float(external[1]), density=float(density),
        maximum_substeps=int(config.maximum_substeps), limit_mode=int(limit_mode),
        weight_mode=int(weight_mode), static_exit=1, diagnostics=int(diagnostics))


_RECORD_BYTES = ctypes.sizeof(_lib.StepRecord)
_CONFIG_CACHE: dict = {}


def run_step(state: DeviceBulkState, config: MobileV2Config, dt: float, *, density: float,
             limit_mode: int, weight_mode: int, external=(0.0, 0.0), diagnostics: int = 0,
             error_prefix: str = "") -> _lib.StepRecord:
    """Run one device step on ``state`` and return the raw record."""
    rt = state.runtime
    rt.bind_stream()
    # configs are frozen dataclasses: the C struct of a (config, dt, ...) key is reused
    key = (config, float(dt), float(density), int(limit_mode), int(weight_mode), tuple(external),
           int(diagnostics))
    try:
        cfg = _CONFIG_CACHE.get(key)
    except TypeError:                      # an unhashable caller config: no caching
        key, cfg = None, None
    if cfg is None:
        if len(_CONFIG_CACHE) > 64:
            _CONFIG_CACHE.clear()
        cfg = _step_config(config, dt, density=density, limit_mode=limit_mode,
                           weight_mode=weight_mode, external=external, diagnostics=diagnostics)
        if key is not None:
            _CONFIG_CACHE[key] = cfg
    rec = _lib.StepRecord()
    rc = rt.lib.mv2_step(rt.ctx, ctypes.byref(cfg), ctypes.byref(rec))
    rt.telemetry.record_call(launches=int(rec.launches) + 4, syncs=1)
    rt.telemetry.record_d2h_bytes(_RECORD_BYTES)
    if rc != _lib.MV2_OK:
        msg = _lib.last_error(rt.ctx)
        if rc == _lib.MV2_EINVAL:
            raise ValueError(msg)
        if rc == _lib.MV2_ENUMERIC:
            raise RuntimeError(f"{error_prefix}{msg}")
        _lib.raise_for(rc, rt.ctx)
    return rec


def step_from_record(rec: _lib.StepRecord, area: float, density: float) -> WarpMobileStep:
    """WarpMobileStep of a raw step record (mobile_v2_warp.py:300-344 field
    meanings): ledger sums are per unit density and cell area on the device,
    impulses scale by density x area.  Scalars come from ctypes slices as
    Python floats; the seven 3-vectors are rows of one array (distinct rows,
    so no two fields share elements)."""
    before = rec.summary_before[:]
    if rec.static_exit:
        vec = np.array(((before[1], before[2], 0.0), (before[1], before[2], 0.0)) + ((0.0, 0.0, 0.0),) * 5)
        fields = dict(
            volume_before_m3=before[0], volume_after_m3=before[0],
            momentum_before_terrain_kg_m_s=vec[0], momentum_after_terrain_kg_m_s=vec[1],
            gravity_pressure_impulse_terrain_ns=vec[2], basal_friction_impulse_terrain_ns=vec[3],
            tool_impulse_on_mobile_terrain_ns=vec[4], numerical_dissipative_impulse_terrain_ns=vec[5],
            maximum_speed_m_s=before[3], substeps=0, cfl_limited=False,
            kinetic_energy_before_j=before[4], kinetic_energy_after_j=before[4],
            gravity_pressure_work_j=0.0, basal_friction_work_j=0.0, tool_work_j=0.0,
            transport_numerical_energy_residual_j=0.0, donor_export_m3=0.0,
            receiver_import_m3=0.0, transport_mass_residual_m3=0.0,
            advected_momentum_crossings_kg_m_s=vec[6])
        return _frozen_step(fields)
    after = rec.summary_after[:]
    diag = rec.diag[:]
    gx, gy = diag[0] * density * area, diag[1] * density * area
    fx, fy = diag[4] * density * area, diag[5] * density * area
    friction_work = 0.5 * diag[7] * density * area
    vec = np.array(((before[1], before[2], 0.0), (after[1], after[2], 0.0), (gx, gy, 0.0), (fx, fy, 0.0),
                    (0.0, 0.0, 0.0), (after[1] - before[1] - gx - fx, after[2] - before[2] - gy - fy, 0.0),
                    (0.0, 0.0, 0.0)))
    fields = dict(
        volume_before_m3=before[0], volume_after_m3=after[0],
        momentum_before_terrain_kg_m_s=vec[0], momentum_after_terrain_kg_m_s=vec[1],
        gravity_pressure_impulse_terrain_ns=vec[2], basal_friction_impulse_terrain_ns=vec[3],
        tool_impulse_on_mobile_terrain_ns=vec[4], numerical_dissipative_impulse_terrain_ns=vec[5],
        maximum_speed_m_s=after[3], substeps=int(rec.substeps),
        cfl_limited=bool(rec.cfl_limited), kinetic_energy_before_j=before[4],
        kinetic_energy_after_j=after[4],
        gravity_pressure_work_j=(after[5] + after[6]) - (before[5] + before[6]),
        basal_friction_work_j=friction_work, tool_work_j=0.0,
        transport_numerical_energy_residual_j=(after[4] + after[5] + after[6])
        - (before[4] + before[5] + before[6]) + friction_work,
        donor_export_m3=0.0, receiver_import_m3=0.0,
        transport_mass_residual_m3=after[0] - before[0],
        advected_momentum_crossings_kg_m_s=vec[6])
    return _frozen_step(fields)


_STEP_FIELDS = tuple(WarpMobileStep.__dataclass_fields__)


def _frozen_step(fields: dict) -> WarpMobileStep:
    """The frozen record filled through its instance dict: the generated
    __init__ sets each field with object.__setattr__, which costs more than
    the rest of a step call's host work.  Same fields, same frozen class."""
    assert len(fields) == len(_STEP_FIELDS)
    step = object.__new__(WarpMobileStep)
    step.__dict__.update(fields)
    return step


class HipProductionMobileV2Solver:
    """Accepted V2 equations consuming the persistent device fields (step_resident)."""

    backend_name = "GPU_HIP_MOBILE_V2_PRODUCTION_SHARED_STATE"

    def __init__(self, *, runtime=None, config: MobileV2Config | None = None) -> None:
        self.runtime = runtime
        self.config = config or MobileV2Config()
        self.state: DeviceBulkState | None = None
        self.material = None
        self.grid = None
        self.last_record: _lib.StepRecord | None = None

    def bind_device_state(self, state: DeviceBulkState, material, grid, integrator) -> None:
        del integrator
        if self.runtime is None:
            self.runtime = state.runtime
        if state.runtime is not self.runtime:
            raise ValueError("[MobileV2Production] runtime mismatch")
        required = {"b_eff", "mobile", "momentum_x", "momentum_y", "weights"}
        missing = required - set(self.runtime.arrays)
        if missing:
            raise ValueError(f"[MobileV2Production] missing state: {sorted(missing)}")
        self.state, self.material, self.grid = state, material, grid

    def set_kernel_timing(self, enable: bool) -> None:
        """Bracket each substep launch with HIP events (record.substep_kernel_ms)."""
        rt = self.runtime if self.runtime is not None else self.state.runtime
        rt.check(rt.lib.mv2_set_timing(rt.ctx, int(bool(enable))))

    def set_launch_graphs(self, enable: bool) -> None:
        """Replay the step's first launch round from a cached HIP graph (default
        off: plain launches measured faster on ROCm 7.2).  Results are identical."""
        rt = self.runtime if self.runtime is not None else self.state.runtime
        rt.check(rt.lib.mv2_set_graphs(rt.ctx, int(bool(enable))))

    def graph_replays(self) -> int:
        """Number of step calls on this runtime served by a graph replay."""
        rt = self.runtime if self.runtime is not None else self.state.runtime
        return int(rt.lib.mv2_graph_replays(rt.ctx))

    def set_speculation(self, enable: bool) -> None:
        """FP32: run the next step's begin pass and first substep in one launch on
        the CFL maximum predicted by this step's end pass (default on).  A
        prediction that does not hold bit for bit restarts the step from its
        input, so results are identical either way."""
        rt = self.runtime if self.runtime is not None else self.state.runtime
        rt.check(rt.lib.mv2_set_speculation(rt.ctx, int(bool(enable))))

    def speculation_stats(self) -> tuple[int, int]:
        """(steps that speculated, steps whose prediction did not hold) on this runtime."""
        rt = self.runtime if self.runtime is not None else self.state.runtime
        out = (ctypes.c_longlong * 2)()
        rt.check(rt.lib.mv2_speculation_stats(rt.ctx, out))
        return int(out[0]), int(out[1])

    def step_resident(self, dt_s: float, **_) -> WarpMobileStep:
        if self.state is None:
            raise RuntimeError("[MobileV2Production] bind state first")
        state = self.state
        density = float(self.material.assumed_bulk_density_kg_m3)
        rec = run_step(state, self.config, float(dt_s), density=density, limit_mode=1,
                       weight_mode=1, diagnostics=1, error_prefix="[MobileV2Production] ")
        self.last_record = rec
        return step_from_record(rec, state.grid.dx * state.grid.dy, density)

    def diagnostics(self) -> dict:
        return {
            "backend_name": self.backend_name,
            "state": ["z_base", "b_eff", "mobile", "momentum_x", "momentum_y"],
            "pressure_flux": "0.5*K*g*h^2_CONSERVATIVE_FACE_FLUX",
            "K_classification": "ENGINEERING_CLOSURE_UNCALIBRATED",
            "precision": None if self.state is None else self.state.precision,
            "substep_loop": "DEVICE_RESIDENT_COOPERATIVE_LAUNCH",
        }


# Name under which callers of the reference bind the production solver.
WarpProductionMobileV2Solver = HipProductionMobileV2Solver


class HipMobileV2ReferenceSolver:
    """Standalone device step of a host ``MobileV2State`` (never bound to production state)."""

    def __init__(self, config: MobileV2Config | None = None, *, device: str = "cuda:0",
                 precision: str = "fp64") -> None:
        self.config = config or MobileV2Config()
        self.device = device
        self.precision = precision
        self._state: DeviceBulkState | None = None

    def _device_state(self, shape) -> DeviceBulkState:
        ny, nx = shape
        if self._state is None or self._state.shape != (ny, nx):
            grid = TerrainGrid(int(nx), int(ny), self.config.dx_m, self.config.dy_m)
            runtime = None if self._state is None else self._state.runtime
            self._state = DeviceBulkState(grid, np.zeros((ny, nx)), device=self.device,
                                          precision=self.precision, runtime=runtime)
        return self._state

    def step(self, state: MobileV2State, dt_s: float, *, external_acceleration_m_s2=None) -> MobileV2Step:
        dt = float(dt_s)
        if not np.isfinite(dt) or dt <= 0.0:
            raise ValueError("dt must be finite/positive")
        ext = (0.0, 0.0)
        per_cell = None
        if external_acceleration_m_s2 is not None:
            # broadcast to q.shape like the reference (mobile_v2_reference.py:204-210)
            e = np.asarray(external_acceleration_m_s2, dtype=np.float64)
            if e.shape == (2,):
                ext = (float(e[0]), float(e[1]))
            else:
                per_cell = np.broadcast_to(e, np.shape(state.q_m2_s))
                if not np.all(np.isfinite(per_cell)):
                    raise ValueError("MobileV2 external acceleration must be finite")
        dev = self._device_state(state.h_m.shape)
        q = np.asarray(state.q_m2_s, dtype=np.float64)
        dev.set_fields(z_base=np.zeros(state.h_m.shape), b_eff=state.b_eff_m, mobile=state.h_m,
                       momentum_x=q[..., 0], momentum_y=q[..., 1])
        rt = dev.runtime
        if per_cell is not None:
            fdt = rt.arrays["momentum_x"].dtype
            ex = rt.upload("v2_external_x", np.ascontiguousarray(per_cell[..., 0]).ravel(), fdt)
            ey = rt.upload("v2_external_y", np.ascontiguousarray(per_cell[..., 1]).ravel(), fdt)
            rt.check(rt.lib.mv2_set_external(rt.ctx, ex.data_ptr(), ey.data_ptr(),
                                             float(np.max(np.abs(per_cell), initial=0.0))))
        try:
            rec = run_step(dev, self.config, dt, density=1.0, limit_mode=0, weight_mode=0,
                           external=ext, error_prefix="Warp ")
        finally:
            if per_cell is not None:
                rt.check(rt.lib.mv2_set_external(rt.ctx, None, None, 0.0))
                rt.release("v2_external_x")
                rt.release("v2_external_y")
        view = dev.explicit_host_view(source="acceptance")
        final = MobileV2State(np.array(state.b_eff_m, dtype=np.float64, copy=True),
                              view.H_mobile_m, view.mobile_momentum_m2_s)
        b, a = np.asarray(rec.summary_before[:]), np.asarray(rec.summary_after[:])
        d = np.asarray(rec.diag[:])
        area = self.config.dx_m * self.config.dy_m
        k0, g0, i0 = b[4], b[5], b[6]
        k1, g1, i1 = a[4], a[5], a[6]
        e0, e1 = k0 + g0 + i0, k1 + g1 + i1
        cons = d[8:10] * area
        fric = d[4:6] * area
        ext_imp = d[2:4] * area
        fdis = float(0.5 * d[7] * area)
        ework = float(d[6] * area)
        p0, p1 = b[1:3] * 1.0, a[1:3] * 1.0
        return MobileV2Step(
            final, dt, int(rec.substeps), float(rec.maximum_cfl), float(b[0]), float(a[0]),
            float(a[0] - b[0]), p0, p1, cons, fric, ext_imp, p1 - p0 - (cons + fric + ext_imp),
            float(k0), float(k1), float(g0), float(g1), float(i0), float(i1), float(e0), float(e1),
            fdis, ework, 0.0, float(e1 - e0 - ework + fdis), float(min(a[7], 0.0)))


WarpMobileV2ReferenceSolver = HipMobileV2ReferenceSolver
