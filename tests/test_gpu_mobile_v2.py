"""GPU parity of the Mobile V2 step: HIP kernels through the C-ABI vs the oracle."""

from __future__ import annotations

import json
from pathlib import Path

import numpy as np
import pytest

import contract_cases
import oracle

pytestmark = pytest.mark.gpu

GOLD = json.loads((Path(__file__).parent / "golden" / "mobile_v2_cpu_contracts.json").read_text())


def _solver(precision="fp64", **cfg):
    from isaac_bulk_hip import HipMobileV2ReferenceSolver, MobileV2Config

    return HipMobileV2ReferenceSolver(MobileV2Config(**cfg), precision=precision)


@pytest.mark.parametrize("name", sorted(contract_cases.cases()))
def test_fp64_contract_cases_bit_exact(name):
    from isaac_bulk_hip import MobileV2State

    b, h, q, dt = contract_cases.cases()[name]
    step = _solver().step(MobileV2State(b, h, q), dt)
    h1, qx1, qy1, rec = oracle.v2_step(b, h, q[..., 0], q[..., 1], dt)
    assert step.substeps == rec["substeps"] == GOLD["contract_cases"][name]["substeps"]
    assert step.maximum_cfl == rec["maximum_cfl"]
    np.testing.assert_array_equal(step.state.h_m, h1)
    np.testing.assert_array_equal(step.state.q_m2_s[..., 0], qx1)
    np.testing.assert_array_equal(step.state.q_m2_s[..., 1], qy1)
    e = GOLD["contract_cases"][name]["energy"]
    assert step.energy_after_j_per_density == pytest.approx(e["after_j_per_density"], rel=1e-12, abs=1e-15)
    assert step.friction_dissipation_j_per_density == pytest.approx(
        e["friction_dissipation_j_per_density"], rel=1e-9, abs=1e-15)
    assert abs(step.mass_residual_m3) < 1e-15


@pytest.mark.parametrize("cfl", [0.25, 0.18])
def test_fp64_pile_128_one_second_bit_exact(cfl):
    """128^2 Gaussian pile with a 32^2 mobile patch, 1.0 s, identical fields."""
    from isaac_bulk_hip import MobileV2State
    from isaac_bulk_hip.workloads import pile_with_patch

    b, h, qx, qy = pile_with_patch(128, seed=7)
    mu = float(np.tan(np.deg2rad(35.0)))
    step = _solver(cfl=cfl, basal_friction_coefficient=mu).step(
        MobileV2State(b, h, np.stack((qx, qy), -1)), 1.0)
    h1, qx1, qy1, rec = oracle.v2_step(b, h, qx, qy, 1.0, cfl=cfl, mu=mu)
    assert rec["status"] == 0 and step.substeps == rec["substeps"] > 10
    np.testing.assert_array_equal(step.state.h_m, h1)
    np.testing.assert_array_equal(step.state.q_m2_s[..., 0], qx1)
    np.testing.assert_array_equal(step.state.q_m2_s[..., 1], qy1)


def test_fp32_pile_128_within_tolerance_and_mass_exact():
    """FP32 fields (int32 mobile quanta) vs the FP64 oracle: L_inf <= 1e-3 m, RMSE <= 1e-5 m."""
    from isaac_bulk_hip import MobileV2State
    from isaac_bulk_hip.workloads import pile_with_patch

    b, h, qx, qy = pile_with_patch(128, seed=7)
    mu = float(np.tan(np.deg2rad(35.0)))
    qm = 2.0 ** -24
    hq = np.rint(h / qm) * qm                    # the oracle starts from the same quantised h
    step = _solver("fp32", cfl=0.25, basal_friction_coefficient=mu).step(
        MobileV2State(b.astype(np.float32).astype(np.float64), hq, np.zeros(h.shape + (2,))), 1.0)
    h1, qx1, qy1, rec = oracle.v2_step(b.astype(np.float32).astype(np.float64), hq, qx, qy, 1.0,
                                       cfl=0.25, mu=mu)
    err = step.state.h_m - h1
    assert np.max(np.abs(err)) <= 1e-3
    assert np.sqrt(np.mean(err ** 2)) <= 1e-5
    # mass moves in whole quanta: the total is exactly conserved
    assert np.sum(np.rint(step.state.h_m / qm)) == np.sum(np.rint(hq / qm))
    assert abs(step.mass_residual_m3) <= 1e-10 * step.mass_before_m3
    assert abs(step.substeps - rec["substeps"]) <= max(2, rec["substeps"] // 20)


def test_fp32_band_heights_give_identical_fields(monkeypatch):
    """The FP32 substep's band height (rows per wave, chosen per grid) only
    changes how cells are grouped into waves: every band gives the same bits."""
    from isaac_bulk_hip import MobileV2State
    from isaac_bulk_hip.workloads import pile_with_patch

    b, h, _, _ = pile_with_patch(150, seed=3)
    state = MobileV2State(b.astype(np.float32).astype(np.float64), np.rint(h * 2.0 ** 24) * 2.0 ** -24,
                          np.zeros(h.shape + (2,)))
    results = {}
    for band in (2, 3, 4, 5, 6):
        monkeypatch.setenv("MV2_FRAME_BAND_ROWS", str(band))
        step = _solver("fp32", cfl=0.25).step(state, 0.25)
        results[band] = (step.state.h_m, step.state.q_m2_s, step.substeps)
    monkeypatch.delenv("MV2_FRAME_BAND_ROWS")
    auto = _solver("fp32", cfl=0.25).step(state, 0.25)
    for band, (hh, qq, n) in results.items():
        np.testing.assert_array_equal(hh, auto.state.h_m, err_msg=str(band))
        np.testing.assert_array_equal(qq, auto.state.q_m2_s, err_msg=str(band))
        assert n == auto.substeps > 3


def test_production_step_resident_matches_oracle_and_ledgers():
    from isaac_bulk_hip import (DeviceBulkState, MaterialScenario, MobileV2Config, TerrainGrid,
                                TerrainVolumeIntegrator, WarpProductionMobileV2Solver)
    from isaac_bulk_hip.config import triangle_vertex_weights
    from isaac_bulk_hip.workloads import pile_with_patch

    b, h, qx, qy = pile_with_patch(64, seed=11)
    grid = TerrainGrid(64, 64, 0.05, 0.05)
    state = DeviceBulkState(grid, b)
    state.set_fields(mobile=h)
    material = MaterialScenario("iron_ore_fines_frozen", 1370.0, 29.8, 800.0, 0.5, 38.0, 30.0, 0.35)
    cfg = MobileV2Config(start_angle_deg=38.0, stop_angle_deg=30.0, basal_friction_coefficient=0.35)
    solver = WarpProductionMobileV2Solver(runtime=state.runtime, config=cfg)
    solver.bind_device_state(state, material, grid, TerrainVolumeIntegrator.from_grid(grid))
    rec = solver.step_resident(1.0 / 60.0)
    h1, qx1, qy1, orec = oracle.v2_step(b, h, qx, qy, 1.0 / 60.0, mu=0.35, start_angle_deg=38.0,
                                        limit_mode=1)
    view = state.explicit_host_view(source="acceptance")
    np.testing.assert_array_equal(view.H_mobile_m, h1)
    np.testing.assert_array_equal(view.mobile_momentum_m2_s[..., 0], qx1)
    assert rec.substeps == orec["substeps"]
    w = triangle_vertex_weights(64, 64, 0.05, 0.05)
    assert rec.volume_before_m3 == pytest.approx(float(np.sum(h * w)), rel=1e-13)
    assert rec.volume_after_m3 == pytest.approx(float(np.sum(h1 * w)), rel=1e-13)
    # the scheme conserves the uniform-weight mass; Triangle-A-C weights halve
    # boundary cells, so the weighted volume changes when mass reaches them
    assert rec.transport_mass_residual_m3 == pytest.approx(float(np.sum(h1 * w) - np.sum(h * w)),
                                                           rel=1e-9, abs=1e-15)
    assert abs(np.sum(view.H_mobile_m) - np.sum(h)) <= 1e-13 * np.sum(h)
    g = rec.gravity_pressure_impulse_terrain_ns
    assert g[0] == pytest.approx(orec["diag"][0] * 1370.0 * 0.0025, rel=1e-9, abs=1e-12)


def test_static_state_takes_zero_substeps():
    from isaac_bulk_hip import MobileV2State

    b, h, q, dt = contract_cases.cases()["STATIC_SLOPE"]
    step = _solver().step(MobileV2State(b, h, q), dt)
    assert step.substeps == 0
    np.testing.assert_array_equal(step.state.h_m, h)


def test_maximum_substeps_raises_runtime_error():
    from isaac_bulk_hip import MobileV2State

    b, h, q, dt = contract_cases.cases()["MOUND"]
    with pytest.raises(RuntimeError, match="maximum_substeps"):
        _solver(maximum_substeps=1).step(MobileV2State(b, h, q), dt)


def test_gate_a_contracts_on_device():
    from isaac_bulk_hip import DeviceBulkState, DeviceMaterialLedger, TerrainGrid

    grid = TerrainGrid(16, 16, 0.05, 0.05, terrain_prim_path="/Terrain")
    state = DeviceBulkState(grid, np.ones(grid.shape), tile_size=4)
    ledger = DeviceMaterialLedger(state, 1370.0)
    before = state.explicit_host_view(source="acceptance")
    idx = np.asarray([7 * 16 + 7, 7 * 16 + 8])
    entrained = state.entrain_host_indices(idx, np.asarray([0.03, 2.0]), reason="gate_a_entrainment")
    after_e = state.explicit_host_view(source="acceptance")
    assert entrained == GOLD["gate_a"]["entrained_volume_m3"]
    assert np.max(np.abs(before.H_free_m - after_e.H_free_m)) == 0.0
    deposited = state.deposit_host_indices(idx, np.asarray([0.01, 0.25]), density_kg_m3=1370.0,
                                           reason="gate_a_deposition")
    assert deposited["deposited_volume_m3"] == GOLD["gate_a"]["deposited_volume_m3"]
    after_d = state.explicit_host_view(source="acceptance")
    assert np.max(np.abs(after_e.H_free_m - after_d.H_free_m)) == 0.0
    snap = ledger.snapshot(state, payload_m3=0.0, airborne_m3=0.0, outflow_m3=0.0)
    assert abs(snap.absolute_volume_error_m3) <= 2e-12
    assert state.runtime.arrays["resting"] is state.runtime.arrays["b_eff"]
