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


def test_fp32_production_diagnostics_mode():
    """FP32 steps with the production record's five per-substep sums
    (diagnostics = 1, warp_mobile_v2.py:220-241) against the full mode on the
    same frames: fields and substep counts identical, the five sums identical
    bit for bit, the other five 0; the production record's impulses and
    friction work track the FP64 oracle's within FP32 tolerance."""
    from isaac_bulk_hip import DeviceBulkState, MobileV2Config, TerrainGrid
    from isaac_bulk_hip.mobile_v2 import run_step
    from isaac_bulk_hip.workloads import pile_with_patch

    n, dt = 96, 1.0 / 60.0
    b, h, qx, qy = pile_with_patch(n, seed=11)
    b = b.astype(np.float32).astype(np.float64)
    h = np.rint(h * 2.0 ** 24) * 2.0 ** -24
    cfg = MobileV2Config(start_angle_deg=38.0, stop_angle_deg=30.0, basal_friction_coefficient=0.35)
    runs = {}
    for mode in (0, 1):
        state = DeviceBulkState(TerrainGrid(n, n, 0.05, 0.05), b, precision="fp32")
        state.set_fields(mobile=h)
        recs = [run_step(state, cfg, dt, density=1.0, limit_mode=1, weight_mode=1, diagnostics=mode)
                for _ in range(3)]
        view = state.explicit_host_view(source="acceptance")
        runs[mode] = (view.H_mobile_m.copy(), view.mobile_momentum_m2_s.copy(),
                      [np.array(r.diag[:]) for r in recs], [r.substeps for r in recs])
    kept, dropped = [0, 1, 4, 5, 7], [2, 3, 6, 8, 9]
    np.testing.assert_array_equal(runs[0][0], runs[1][0])
    np.testing.assert_array_equal(runs[0][1], runs[1][1])
    assert runs[0][3] == runs[1][3] and min(runs[0][3]) > 0
    for d0, d1 in zip(runs[0][2], runs[1][2]):
        np.testing.assert_array_equal(d0[kept], d1[kept])
        assert np.all(d1[dropped] == 0.0)
    # the first frame's sums against the FP64 oracle from the same state (the
    # device's sum 7 is twice the kinetic energy removed by friction, the
    # record halves it)
    _, _, _, orec = oracle.v2_step(b, h, qx, qy, dt, mu=0.35, start_angle_deg=38.0, limit_mode=1)
    for d in kept:
        got = runs[1][2][0][d] * (0.5 if d == 7 else 1.0)
        assert got == pytest.approx(orec["diag"][d], rel=2e-3, abs=1e-9), d
    with pytest.raises(ValueError, match="diagnostics"):
        state = DeviceBulkState(TerrainGrid(n, n, 0.05, 0.05), b, precision="fp32")
        state.set_fields(mobile=h)
        run_step(state, cfg, dt, density=1.0, limit_mode=1, weight_mode=1, diagnostics=2)


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


def test_c1_cfl_0_4_through_the_opt_in_bit_exact():
    """C1 (128^2 pile, mu = tan 35 deg, 1.0 s) at CFL 0.4: the reference
    rejects CFL > 0.25 (mobile_v2_reference.py:37-38), so it runs through the
    explicit opt-in; the FP64 fields equal the oracle's at the same CFL bit for
    bit (or both stop with the same positivity error)."""
    from isaac_bulk_hip import MobileV2Config, MobileV2State
    from isaac_bulk_hip.workloads import pile_with_patch

    with pytest.raises(ValueError, match="0.25"):
        MobileV2Config(cfl=0.4)
    b, h, qx, qy = pile_with_patch(128, seed=7)
    mu = float(np.tan(np.deg2rad(35.0)))
    solver = _solver(cfl=0.4, basal_friction_coefficient=mu, allow_cfl_above_quarter=True)
    h1, qx1, qy1, rec = oracle.v2_step(b, h, qx, qy, 1.0, cfl=0.4, mu=mu)
    if rec["status"] == 1:
        with pytest.raises(RuntimeError, match="positivity"):
            solver.step(MobileV2State(b, h, np.stack((qx, qy), -1)), 1.0)
        return
    step = solver.step(MobileV2State(b, h, np.stack((qx, qy), -1)), 1.0)
    assert rec["status"] == 0 and step.substeps == rec["substeps"] > 5
    assert step.maximum_cfl == rec["maximum_cfl"] and rec["maximum_cfl"] > 0.25
    np.testing.assert_array_equal(step.state.h_m, h1)
    np.testing.assert_array_equal(step.state.q_m2_s[..., 0], qx1)
    np.testing.assert_array_equal(step.state.q_m2_s[..., 1], qy1)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_per_cell_external_acceleration_matches_oracle(precision):
    """A per-cell external acceleration (the reference's array broadcast to
    q.shape, mobile_v2_reference.py:204-210): FP64 fields bit-exact against
    the oracle, FP32 within the frame tolerances; the impulse and work of the
    push agree with the oracle's ledgers."""
    from isaac_bulk_hip import MobileV2State
    from isaac_bulk_hip.workloads import pile_with_patch

    b, h, qx, qy = pile_with_patch(96, seed=5)
    rng = np.random.default_rng(3)
    ext = np.zeros((96, 96, 2))
    ext[..., 0] = rng.uniform(-8.0, 8.0, (96, 96))
    ext[40:, :, 1] = 6.0
    qm = 2.0 ** -24
    if precision == "fp32":
        b = b.astype(np.float32).astype(np.float64)
        h = np.rint(h / qm) * qm
    step = _solver(precision).step(MobileV2State(b, h, np.zeros(h.shape + (2,))), 0.2,
                                   external_acceleration_m_s2=ext)
    h1, qx1, qy1, rec = oracle.v2_step(b, h, qx, qy, 0.2, external=ext)
    assert rec["status"] == 0 and rec["substeps"] > 3
    if precision == "fp64":
        assert step.substeps == rec["substeps"]
        np.testing.assert_array_equal(step.state.h_m, h1)
        np.testing.assert_array_equal(step.state.q_m2_s[..., 0], qx1)
        np.testing.assert_array_equal(step.state.q_m2_s[..., 1], qy1)
        assert step.external_work_j_per_density == pytest.approx(rec["external_work"], rel=1e-9)
    else:
        assert abs(step.substeps - rec["substeps"]) <= 1
        for got, ref in ((step.state.h_m, h1), (step.state.q_m2_s[..., 0], qx1), (step.state.q_m2_s[..., 1], qy1)):
            assert np.max(np.abs(got - ref)) <= 1e-3 and np.sqrt(np.mean((got - ref) ** 2)) <= 1e-5
    # the device record counts wet cells (the Warp kernel's convention); the
    # reference's ledger adds h a dt of cells at or below the dry tolerance
    np.testing.assert_allclose(step.external_impulse_m4_s, rec["external_impulse"], rtol=1e-5, atol=1e-12)
    # with no external acceleration the same state moves differently
    plain = _solver(precision).step(MobileV2State(b, h, np.zeros(h.shape + (2,))), 0.2)
    assert not np.array_equal(plain.state.q_m2_s, step.state.q_m2_s)


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_step_graph_replays_match_plain_launches(precision):
    """Steps served by a replayed HIP graph of the launch sequence equal the
    same steps launched kernel by kernel, bit for bit and frame by frame, with
    eager torch work and synchronizations between the replays."""
    import torch
    from isaac_bulk_hip import (DeviceBulkState, MaterialScenario, MobileV2Config, TerrainGrid,
                                TerrainVolumeIntegrator, WarpProductionMobileV2Solver)
    from isaac_bulk_hip.workloads import cut_face_terrain

    n = 130
    b, h, _ = cut_face_terrain(n, 0.05, seed=3)
    grid = TerrainGrid(n, n, 0.05, 0.05)
    material = MaterialScenario("iron_ore_fines_uncalibrated", 1370.0, 29.8, 0.0, 0.5, 34.0, 30.0, 0.55)
    runs = []
    for graphs in (True, False):
        state = DeviceBulkState(grid, b, precision=precision)
        state.set_fields(mobile=h)
        solver = WarpProductionMobileV2Solver(runtime=state.runtime, config=MobileV2Config())
        solver.bind_device_state(state, material, grid, TerrainVolumeIntegrator.from_grid(grid))
        solver.set_launch_graphs(graphs)
        runs.append((state, solver))
    for frame in range(12):
        recs = []
        for state, solver in runs:
            recs.append(solver.step_resident(1.0 / 60.0))
            torch.cuda.synchronize()
            _ = (state.runtime.arrays["mobile"].double().sum() + 1.0).item()   # eager work between replays
        assert recs[0].substeps == recs[1].substeps > 0, frame
        assert recs[0].volume_after_m3 == recs[1].volume_after_m3, frame
        fields = [state.runtime.arrays for state, _ in runs]
        for name in ("mobile", "momentum_x", "momentum_y", "b_eff"):
            assert torch.equal(fields[0][name], fields[1][name]), (frame, name)
    # the first calls see each key once; later ones replay
    assert runs[0][1].graph_replays() >= 6 and runs[1][1].graph_replays() == 0
