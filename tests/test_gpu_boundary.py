"""GPU tests of the DeviceBulkState compact bridges, the transfer budget, the
FailureZone activation, the operator chain and the FP32 exchange contract,
through the C-ABI, against the oracle's restatements."""

from __future__ import annotations

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

Q = oracle.QUANTUM


def _state(n=40, precision="fp64", seed=0, tile_size=16):
    from isaac_bulk_hip import DeviceBulkState, TerrainGrid
    from isaac_bulk_hip.workloads import pile_with_patch

    b, h, _, _ = pile_with_patch(n, seed=seed)
    grid = TerrainGrid(n, n, 0.05, 0.05)
    state = DeviceBulkState(grid, b + 0.3, precision=precision, tile_size=tile_size)
    rng = np.random.default_rng(seed)
    state.set_fields(mobile=h, momentum_x=0.01 * rng.standard_normal((n, n)) * (h > 0),
                     momentum_y=0.01 * rng.standard_normal((n, n)) * (h > 0))
    state.begin_physics_step()
    return state, grid


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_patch_scatter_gather_and_regions(precision):
    from isaac_bulk_hip import BulkStateAuthorityError

    state, grid = _state(precision=precision)
    rng = np.random.default_rng(1)
    view0 = state.explicit_host_view(source="acceptance")
    vals = rng.uniform(0.4, 1.2, (7, 9))
    state.apply_host_patch("resting", (3, 10, 5, 14), vals, reason="test")
    mob = np.rint(rng.uniform(0.0, 0.05, (4, 3)) / Q) * Q           # whole quanta: exact in both modes
    state.apply_host_patch("mobile", (20, 24, 0, 3), mob, reason="test")
    qx = rng.normal(0.0, 0.1, (2, 5))
    state.apply_host_patch("momentum_x", (30, 32, 30, 35), qx, reason="test")
    state.apply_host_patch("b_eff", (0, 0, 0, 5), np.zeros((0, 5)), reason="empty")
    view = state.explicit_host_view(source="acceptance")
    cast = (lambda a: a) if precision == "fp64" else (lambda a: a.astype(np.float32).astype(np.float64))
    exp_b = view0.b_eff_m.copy()
    exp_b[3:10, 5:14] = cast(vals)
    np.testing.assert_array_equal(view.b_eff_m, exp_b)
    exp_m = view0.H_mobile_m.copy()
    exp_m[20:24, 0:3] = mob
    np.testing.assert_array_equal(view.H_mobile_m, exp_m)
    np.testing.assert_array_equal(view.mobile_momentum_m2_s[30:32, 30:35, 0], cast(qx))
    # compact reads equal the host view's slices
    p = state.download_patch((2, 12, 4, 16), source="geometry")
    np.testing.assert_array_equal(p.b_eff_m, view.b_eff_m[2:12, 4:16])
    np.testing.assert_array_equal(p.H_mobile_m, view.H_mobile_m[2:12, 4:16])
    np.testing.assert_array_equal(p.mobile_momentum_m2_s, view.mobile_momentum_m2_s[2:12, 4:16])
    np.testing.assert_array_equal(p.z_base_m, view.z_base_m[2:12, 4:16])
    r = state.download_region((1, 39, 0, 33), source="failure_zone_geometry")
    np.testing.assert_array_equal(r.H_free_m, view.H_free_m[1:39, 0:33])
    with pytest.raises(ValueError, match="capacity"):
        state.download_patch((0, 17, 0, 16), source="geometry")     # more than one 16 x 16 tile
    with pytest.raises(BulkStateAuthorityError, match="FULL_FIELD"):
        state.download_region((0, 40, 0, 40), source="geometry")
    with pytest.raises(ValueError, match="outside"):
        state.apply_host_patch("b_eff", (0, 41, 0, 1), np.zeros((41, 1)), reason="test")
    with pytest.raises(ValueError, match="unsupported"):
        state.apply_host_patch("z_base", (0, 1, 0, 1), np.zeros((1, 1)), reason="test")


def test_index_scatter_and_add():
    import torch

    for precision in ("fp64", "fp32"):
        state, grid = _state(precision=precision)
        a = state.runtime.arrays
        idx = np.array([5, 77, 1599, 300])
        state.apply_host_indices("material_mask", idx, 1, reason="forcing")
        state.apply_host_indices("frontier_reached", idx[:2], np.array([7, -3]), reason="seed")
        mm = a["material_mask"].cpu().numpy()
        assert set(np.flatnonzero(mm).tolist()) == set(idx.tolist())
        fr = a["frontier_reached"].cpu().numpy()
        assert fr[5] == 7 and fr[77] == -3 and np.count_nonzero(fr) == 2
        with pytest.raises(ValueError, match="unsupported"):
            state.apply_host_indices("avalanche_latch", idx, 1, reason="x")
        with pytest.raises(ValueError, match="outside"):
            state.apply_host_indices("material_mask", [1600], 1, reason="x")
        before = state.explicit_host_view(source="acceptance")
        # dyadic values: every order of the atomic sums gives the same bits
        add_idx = np.array([10, 10, 10, 42, 900])
        add_val = np.array([0.25, 0.125, -0.0625, 2.0 ** -10, 0.5])
        state.add_host_indices("resting", add_idx, add_val, reason="landing")
        mob_val = np.array([3, 5, 7, 11, 13]) * Q
        state.add_host_indices("mobile", add_idx, mob_val, reason="landing")
        after = state.explicit_host_view(source="acceptance")
        exp_b = before.b_eff_m.ravel().copy()
        np.add.at(exp_b, add_idx, add_val)
        if precision == "fp64":
            np.testing.assert_array_equal(after.b_eff_m.ravel(), exp_b)
        else:
            # float32 atomics round after each add (in the order they land)
            np.testing.assert_allclose(after.b_eff_m.ravel(), exp_b, rtol=0, atol=4 * 2.0 ** -22)
        exp_m = before.H_mobile_m.ravel().copy()
        np.add.at(exp_m, add_idx, mob_val)
        np.testing.assert_array_equal(after.H_mobile_m.ravel(), exp_m)
        torch.cuda.synchronize()


def test_transfer_snapshot_and_normal_step_budget():
    from isaac_bulk_hip import BulkStateAuthorityError

    state, grid = _state()
    state.begin_physics_step()
    state.apply_host_indices("material_mask", [1, 2], 1, reason="forcing")
    state.download_region((0, 8, 0, 8), source="geometry")
    state.reservoir_reduction(1370.0)
    snap = state.transfer_snapshot()
    assert snap.full_field_h2d_count == 0 and snap.full_field_d2h_count == 0
    assert snap.h2d_transfer_count >= 1 and snap.d2h_transfer_count >= 2
    state.assert_normal_step_transfer_budget()
    state.explicit_host_view(source="debug")                         # five full fields
    assert state.transfer_snapshot().full_field_d2h_count == 5
    with pytest.raises(BulkStateAuthorityError, match="full-field transfer"):
        state.assert_normal_step_transfer_budget()
    state.begin_physics_step()
    state.assert_normal_step_transfer_budget()
    assert set(state.transfer_snapshot().to_dict()) >= {"full_field_h2d_count", "d2h_blocking_ms"}


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_failure_zone_activation_matches_oracle(precision):
    """Activation patch on a seeded terrain: b/h per cell, the moved volume,
    the forcing index set (row-major) and material_mask equal the oracle's
    restatement of gpu_failure_bridge.py:114-136; H_free is bit-invariant;
    clear_tool_forcing resets exactly those cells."""
    from isaac_bulk_hip import DeviceFailureZoneBridge, MaterialScenario, TerrainVolumeIntegrator

    state, grid = _state(n=64, precision=precision, seed=3)
    rng = np.random.default_rng(7)
    bbox = (10, 41, 5, 35)
    shape = (31, 30)
    act = np.where(rng.random(shape) < 0.4, rng.uniform(0.0, 0.04, shape), 0.0)
    act[0, 0] = 5.0                                                  # donor-limited
    aff = rng.random(shape) < 0.15
    view0 = state.explicit_host_view(source="acceptance")
    a = state.runtime.arrays
    w = TerrainVolumeIntegrator.from_grid(grid).vertex_weights_m2
    if precision == "fp64":
        ob, om, ovol, oforce = oracle.failure_activation(view0.z_base_m, view0.b_eff_m, view0.H_mobile_m, w, bbox,
                                                         act, aff)
    else:
        ob, om, ovol, oforce = oracle.failure_activation(
            a["z_base"].cpu().numpy().reshape(64, 64), a["b_eff"].cpu().numpy().reshape(64, 64),
            a["mobile"].cpu().numpy().reshape(64, 64), w, bbox, act, aff, fp32=True)
        om = om.astype(np.float64) * Q
    material = MaterialScenario("iron_ore_fines_uncalibrated", 1370.0, 29.8, 0.0, 0.5, 34.0, 30.0, 0.55)
    bridge = DeviceFailureZoneBridge(state, material, grid, TerrainVolumeIntegrator.from_grid(grid))
    res = bridge.activate(bbox, act, aff)
    view = state.explicit_host_view(source="acceptance")
    np.testing.assert_array_equal(view.b_eff_m, np.asarray(ob, dtype=np.float64))
    np.testing.assert_array_equal(view.H_mobile_m, om)
    if precision == "fp32":
        np.testing.assert_array_equal(view.H_free_m, view0.H_free_m)      # whole quanta on the float grid
    else:
        # (b - d) + (m + d) equals b + m up to one rounding
        np.testing.assert_allclose(view.H_free_m, view0.H_free_m, rtol=2.3e-16, atol=0)
    assert res.forcing_flat_indices.tolist() == oforce.tolist() and oforce.size > 0
    assert res.activated_volume_m3 == pytest.approx(ovol, rel=1e-13)
    mm = a["material_mask"].cpu().numpy()
    assert np.flatnonzero(mm).tolist() == oforce.tolist()
    # the free surface does not move (collected dirty tiles stay empty beyond
    # rounding); the patch's tiles are recorded as touched
    assert res.activation_mode == "FEE_FAILURE_ZONE"
    if precision == "fp32":
        assert res.dirty_tile_ids.size == 0
    assert {(10 // 16) * 4 + 5 // 16, (40 // 16) * 4 + 34 // 16} <= set(view.dirty_tile_ids.tolist())
    bridge.clear_tool_forcing(res.forcing_flat_indices)
    assert not a["material_mask"].cpu().numpy().any()


def test_failure_zone_execute_with_host_geometry():
    """The reference's execute loop around a host geometry callable: the
    region grows while the activation touches its edge, the activation is
    donor-limited on the patch, and the commit equals a direct activation."""
    from isaac_bulk_hip import DeviceFailureZoneBridge, FailureGeometry, MaterialScenario

    state, grid = _state(n=64, seed=4)
    material = MaterialScenario("iron_ore_fines_uncalibrated", 1370.0, 29.8, 0.0, 0.5, 34.0, 30.0, 0.55)
    bridge = DeviceFailureZoneBridge(state, material, grid, maximum_wedge_length_m=0.2)
    calls = []

    def geometry(patch, bbox):
        calls.append(bbox)
        rows = np.arange(bbox[0], bbox[1])[:, None]
        cols = np.arange(bbox[2], bbox[3])[None, :]
        wedge = (np.abs(rows - 30) <= 6) & (np.abs(cols - 30) <= 6)   # a 13 x 13 wedge
        return FailureGeometry(np.where(wedge, 0.02, 0.0), wedge)

    view0 = state.explicit_host_view(source="acceptance")
    res = bridge.execute(geometry, bridge.physical_request_bbox((28, 28, 33, 33)))
    assert len(calls) >= 2 and res.expansion_count == len(calls) - 1
    r0, r1, c0, c1 = res.patch_bbox_yx
    assert r0 <= 24 and r1 >= 37 and c0 <= 24 and c1 >= 37
    view = state.explicit_host_view(source="acceptance")
    np.testing.assert_allclose(view.H_free_m, view0.H_free_m, rtol=2.3e-16, atol=0)
    moved = view0.b_eff_m - view.b_eff_m
    assert moved[24:37, 24:37].min() > 0.0 and np.count_nonzero(moved) == 169


def test_chain_records_mobile_and_activation():
    from isaac_bulk_hip import GpuBulkOperatorChain, MaterialScenario, TerrainVolumeIntegrator

    state, grid = _state(n=48, seed=5)
    material = MaterialScenario("iron_ore_fines_uncalibrated", 1370.0, 29.8, 0.0, 0.5, 34.0, 30.0, 0.55)
    chain = GpuBulkOperatorChain(state, material, grid, TerrainVolumeIntegrator.from_grid(grid))
    state.begin_physics_step()
    act = np.full((6, 6), 0.01)
    res = chain.activate_failure_zone((10, 16, 10, 16), act, np.ones((6, 6), bool))
    step = chain.step_mobile(1.0 / 60.0)
    chain.clear_failure_tool_forcing(res)
    chain.step_deposition(1.0 / 60.0)
    chain.advance_large_avalanche(1.0 / 60.0)
    state.assert_normal_step_transfer_budget()
    mods = [r.module for r in chain.records]
    assert mods == ["FailureZoneActivation", "Mobile", "Deposition", "LargeAvalancheDeviceTransition"]
    assert step.substeps > 0 and "full_field_d2h_count" in chain.records[0].transfer
    assert chain.diagnostics()["records"][1]["dirty_tile_count"] >= 1


def test_fp32_entrain_deposit_keep_free_surface_and_ledger():
    """FP32 storage: per cell b_eff + mobile bit-invariant, the amounts equal
    the oracle's rule, and the reported volumes equal the change of the
    resting and of the mobile volume."""
    from isaac_bulk_hip import TerrainVolumeIntegrator

    state, grid = _state(n=48, precision="fp32", seed=6)
    a = state.runtime.arrays
    w = TerrainVolumeIntegrator.from_grid(grid).vertex_weights_m2.ravel()
    rng = np.random.default_rng(9)
    idx = rng.choice(48 * 48, 400, replace=False)
    req = rng.uniform(0.0, 0.02, idx.size)
    b0 = a["b_eff"].cpu().numpy().copy()
    m0 = a["mobile"].cpu().numpy().copy()
    vol = state.entrain_host_indices(idx, req, reason="dig")
    b1 = a["b_eff"].cpu().numpy().copy()
    m1 = a["mobile"].cpu().numpy().copy()
    for i, r in zip(idx, req):
        assert m1[i] - m0[i] == oracle.entrain_quanta_f32(b0[i], min(r, max(float(b0[i]), 0.0)),
                                                          2 ** 31 - 1 - int(m0[i]))
    np.testing.assert_array_equal(b1.astype(np.float64) + m1 * Q, b0.astype(np.float64) + m0 * Q)
    assert vol == pytest.approx(float(np.sum((b0.astype(np.float64) - b1) * w)), rel=1e-13)
    assert vol == pytest.approx(float(np.sum((m1 - m0) * Q * w)), rel=1e-13)
    dep = state.deposit_host_indices(idx, req * 0.7, density_kg_m3=1370.0, reason="settle")
    b2 = a["b_eff"].cpu().numpy().copy()
    m2 = a["mobile"].cpu().numpy().copy()
    crossing = 0
    for i, r in zip(idx, req * 0.7):
        dq, bn = oracle.deposit_quanta_f32(b1[i], min(int(np.floor(r / Q)), int(m1[i])))
        assert m1[i] - m2[i] == dq and b2[i] == bn
        crossing += float(bn) != float(b1[i]) + dq * Q
    # H_free changes only in the rare rounded crossings of the deposit rule
    mism = (b2.astype(np.float64) + m2 * Q) != (b1.astype(np.float64) + m1 * Q)
    assert int(mism.sum()) == crossing <= idx.size // 20
    assert dep["deposited_volume_m3"] == pytest.approx(float(np.sum((m1 - m2) * Q * w)), rel=1e-13)
