"""GPU parity of the Mobile -> Resting deposition step (mv2_deposition_step
through WarpDepositionOperator) against the oracle's restatement."""

from __future__ import annotations

import ctypes

import numpy as np
import pytest

import deposition_cases as dc
import oracle

pytestmark = pytest.mark.gpu

Q = 1.0 / 16777216.0
TAN_STOP = float(np.tan(np.deg2rad(dc.STOP_DEG)))


def _material(cohesion=dc.COHESION):
    from isaac_bulk_hip import MaterialScenario

    return MaterialScenario("deposition", dc.DENSITY, 29.8, cohesion, 0.5, 38.0, dc.STOP_DEG, 0.35)


def _setup(b, h, qx, qy, forcing, cum, *, dx, dy, precision):
    import torch

    from isaac_bulk_hip import DeviceBulkState, TerrainGrid, WarpDepositionOperator

    ny, nx = b.shape
    grid = TerrainGrid(nx, ny, dx, dy, terrain_prim_path="/Terrain")
    state = DeviceBulkState(grid, b, tile_size=8, precision=precision)
    state.set_fields(mobile=h, momentum_x=qx, momentum_y=qy)
    rt = state.runtime
    rt.arrays["material_mask"].copy_(torch.from_numpy(forcing.ravel().astype(np.int32)).to(rt.device))
    rt.arrays["avalanche_m2r_cumulative"].copy_(torch.from_numpy(cum.ravel()).to(rt.device))
    op = WarpDepositionOperator(runtime=rt)
    op.bind_device_state(state)
    return state, op


def _fields(state):
    rt = state.runtime
    rt.synchronize()
    shape = state.shape
    return {name: rt.arrays[name].cpu().numpy().reshape(shape) for name in (
        "b_eff", "mobile", "momentum_x", "momentum_y", "deposition_work",
        "deposition_removed_momentum_x", "deposition_removed_momentum_y", "avalanche_m2r_cumulative")}


def _oracle(b, h, qx, qy, forcing, cum, *, dx, dy, dt, cohesion, settle, fp32):
    w = oracle.vertex_weights(*b.shape, dx, dy)
    return oracle.deposition(b, h, qx, qy, cum, w, dx=dx, dy=dy, dt=dt, density=dc.DENSITY,
                             gravity=dc.GRAVITY, cohesion=cohesion, tan_stop=TAN_STOP,
                             settle_subcell_tail=settle, forcing=forcing, fp32=fp32)


def _case(ny, nx, seed, *, dx, dy, dt, cohesion, settle, precision):
    b, h, qx, qy, forcing, cum = dc.fields(ny, nx, seed, dx=dx, dy=dy)
    print(f"[deposition] {precision} {ny}x{nx} seed={seed} cohesion={cohesion} settle={settle}",
          flush=True)
    state, op = _setup(b, h, qx, qy, forcing, cum, dx=dx, dy=dy, precision=precision)
    step = op.step_resident(_material(cohesion), dt, settle_subcell_tail=settle)
    got = _fields(state)
    if precision == "fp64":
        ref = _oracle(b, h, qx, qy, forcing, cum, dx=dx, dy=dy, dt=dt, cohesion=cohesion,
                      settle=settle, fp32=False)
    else:
        hq = np.rint(h / Q).astype(np.int32)
        ref = _oracle(b.astype(np.float32), hq, qx.astype(np.float32), qy.astype(np.float32),
                      forcing, cum, dx=dx, dy=dy, dt=dt, cohesion=cohesion, settle=settle, fp32=True)
    return step, got, ref, h


FIELD_ORDER = ("b_eff", "mobile", "momentum_x", "momentum_y", "deposition_work",
               "deposition_removed_momentum_x", "deposition_removed_momentum_y",
               "avalanche_m2r_cumulative")


def _assert_parity(step, got, ref):
    for name, expected in zip(FIELD_ORDER, ref[:8]):
        np.testing.assert_array_equal(got[name], expected, err_msg=name)
    sums = ref[8]
    assert step.deposited_volume_m3 == pytest.approx(sums[0], rel=1e-12, abs=1e-300)
    np.testing.assert_allclose(step.removed_mobile_momentum_terrain_kg_m_s[:2], sums[1:],
                               rtol=1e-12, atol=1e-300)
    assert step.removed_mobile_momentum_terrain_kg_m_s[2] == 0.0


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("cohesion,settle", [(0.0, False), (50.0, False), (800.0, False),
                                             (50.0, True)])
def test_ragged_grid_bit_exact(precision, cohesion, settle):
    """37 x 53 with dy != dx: every field bit-identical to the oracle, the
    three weighted sums within 1e-12 relative (summation order differs)."""
    step, got, ref, h = _case(37, 53, 21, dx=0.05, dy=0.07, dt=1.0 / 60.0, cohesion=cohesion,
                              settle=settle, precision=precision)
    _assert_parity(step, got, ref)
    assert step.deposited_volume_m3 > 0.0
    assert step.parameter_status == "LITERATURE_BASED_REDUCED_ORDER_UNCALIBRATED"


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_full_size_701_bit_exact_and_balanced(precision):
    """701 x 701 (the full-size terrain): fields bit-identical to the oracle
    and the deposited volume equal to the mobile volume that left."""
    step, got, ref, h = _case(701, 701, 3, dx=0.05, dy=0.05, dt=1.0 / 60.0, cohesion=50.0,
                              settle=False, precision=precision)
    _assert_parity(step, got, ref)
    w = oracle.vertex_weights(701, 701, 0.05, 0.05)
    if precision == "fp32":
        before = float(np.sum(np.rint(h / Q) * Q * w))
        after = float(np.sum(got["mobile"].astype(np.float64) * Q * w))
    else:
        before, after = float(np.sum(h * w)), float(np.sum(got["mobile"] * w))
    assert abs(before - after - step.deposited_volume_m3) <= 1e-10 * max(before, 1.0)


def test_reference_subcell_tail_case_on_device():
    """tests/test_device_bulk_state.py:195-225 on the device state."""
    import torch

    from isaac_bulk_hip import (DeviceBulkState, TerrainGrid, TerrainVolumeIntegrator,
                                WarpDepositionOperator)

    grid = TerrainGrid(16, 16, 0.05, 0.05, terrain_prim_path="/Terrain")
    integrator = TerrainVolumeIntegrator.from_grid(grid)
    resting = np.zeros(grid.shape)
    resting[8, 8] = 0.20
    mobile = np.zeros(grid.shape)
    mobile[8, 9] = 0.01
    qx = np.zeros(grid.shape)
    qx[8, 9] = 0.02
    state = DeviceBulkState(grid, resting, tile_size=4)
    state.set_fields(mobile=mobile, momentum_x=qx)
    op = WarpDepositionOperator(runtime=state.runtime)
    op.bind_device_state(state)
    result = op.step_resident(_material(), 1.0 / 60.0, settle_subcell_tail=True)
    view = state.explicit_host_view(source="acceptance")
    assert result.deposited_volume_m3 == pytest.approx(integrator.integrate(mobile), abs=1e-13)
    np.testing.assert_allclose(view.H_resting_m, resting + mobile, atol=1e-13)
    np.testing.assert_allclose(view.H_mobile_m, 0.0, atol=1e-13)
    assert torch.count_nonzero(state.runtime.arrays["momentum_x"]).item() == 0


def test_dry_field_and_repeated_calls():
    """No mobile material: nothing moves and every per-cell output is zeroed
    (outputs of an earlier call do not leak into the next)."""
    b, h, qx, qy, forcing, cum = dc.fields(24, 24, 9)
    state, op = _setup(b, h, qx, qy, forcing, cum, dx=0.05, dy=0.05, precision="fp64")
    first = op.step_resident(_material(50.0), 1.0 / 60.0)
    assert first.deposited_volume_m3 > 0.0
    state.set_fields(mobile=np.zeros_like(h), momentum_x=np.zeros_like(h), momentum_y=np.zeros_like(h))
    before = _fields(state)
    second = op.step_resident(_material(50.0), 1.0 / 60.0)
    after = _fields(state)
    assert second.deposited_volume_m3 == 0.0
    np.testing.assert_array_equal(after["b_eff"], before["b_eff"])
    for name in ("deposition_work", "deposition_removed_momentum_x", "deposition_removed_momentum_y"):
        np.testing.assert_array_equal(after[name], 0.0)


def test_errors_follow_the_reference():
    from isaac_bulk_hip import DeviceBulkState, HipRuntime, TerrainGrid, WarpDepositionOperator

    with pytest.raises(ValueError):
        WarpDepositionOperator(settling_rate_m_s=0.0, runtime=object())
    grid = TerrainGrid(8, 8, 0.05, 0.05, terrain_prim_path="/Terrain")
    state = DeviceBulkState(grid, np.ones(grid.shape), tile_size=4)
    op = WarpDepositionOperator(runtime=state.runtime)
    with pytest.raises(RuntimeError):
        op.step_resident(_material(), 1.0 / 60.0)
    op.bind_device_state(state)
    for dt in (0.0, -1.0, float("nan")):
        with pytest.raises(ValueError):
            op.step_resident(_material(), dt)
    other = WarpDepositionOperator(runtime=HipRuntime("cuda:0"))
    with pytest.raises(ValueError):
        other.bind_device_state(state)
    # the C-ABI rejects missing output fields
    from isaac_bulk_hip import _lib

    rt = state.runtime
    cfg = _lib.DepositionConfig(1.0 / 60.0, 0.3, 0.18, 1370.0, 9.81, 0.0, 0.5, 0, 1)
    out = np.zeros(3)
    rc = rt.lib.mv2_deposition_step(rt.ctx, ctypes.byref(cfg), None, None, None, None, None,
                                    out.ctypes.data)
    assert rc == _lib.MV2_EINVAL
