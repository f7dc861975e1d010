"""GPU parity of the LargeAvalanche persistence reduction
(mv2_avalanche_persistence, reduce_persistence) against the oracle."""

from __future__ import annotations

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

Q = 1.0 / 16777216.0


def _history(ny, nx, seed):
    rng = np.random.default_rng(seed)
    initial = 1.0 + rng.uniform(0.0, 0.5, (ny, nx))
    resting = initial + rng.normal(0.0, 0.02, (ny, nx))
    mobile = np.where(rng.random((ny, nx)) < 0.3, rng.uniform(0.0, 0.1, (ny, nx)), 0.0)
    qx = mobile * rng.uniform(-0.4, 0.4, (ny, nx))
    qy = mobile * rng.uniform(-0.4, 0.4, (ny, nx))
    count = rng.integers(0, 4, (ny, nx)).astype(np.int32) * (rng.random((ny, nx)) < 0.4)
    slope = rng.uniform(0.0, 1.2, (ny, nx))
    unstable = (rng.random((ny, nx)) < 0.02).astype(np.int32)
    latch = (rng.random((ny, nx)) < 0.3).astype(np.int32)
    r2m = rng.uniform(0.0, 1e-4, (ny, nx)) * (count > 0)
    m2r = rng.uniform(0.0, 1e-4, (ny, nx))
    return initial, resting, mobile, qx, qy, count.astype(np.int32), slope, unstable, latch, r2m, m2r


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("shape,tile", [((97, 131), 16), ((701, 701), 64)])
def test_persistence_matches_oracle(precision, shape, tile):
    import torch

    from isaac_bulk_hip import (DeviceBulkState, DeviceLargeAvalancheBridge, LargeAvalancheTransitionConfig,
                                MaterialScenario, TerrainGrid, TerrainVolumeIntegrator)

    ny, nx = shape
    print(f"[persistence] {precision} {ny}x{nx} tile {tile}", flush=True)
    initial, resting, mobile, qx, qy, count, slope, unstable, latch, r2m, m2r = _history(ny, nx, 5)
    grid = TerrainGrid(nx, ny, 0.05, 0.05, terrain_prim_path="/Terrain")
    state = DeviceBulkState(grid, initial, tile_size=tile, precision=precision)
    state.set_fields(b_eff=resting, mobile=mobile, momentum_x=qx, momentum_y=qy)
    rt = state.runtime
    for name, arr in (("avalanche_activation_count", count), ("avalanche_unstable", unstable),
                      ("avalanche_latch", latch)):
        rt.arrays[name].copy_(torch.from_numpy(arr.ravel().astype(np.int32)).to(rt.device))
    for name, arr in (("avalanche_slope", slope), ("avalanche_r2m_cumulative", r2m),
                      ("avalanche_m2r_cumulative", m2r)):
        rt.arrays[name].copy_(torch.from_numpy(arr.ravel()).to(rt.device))
    material = MaterialScenario("persistence", 1370.0, 29.8, 0.0, 0.5, 38.0, 30.0, 0.35)
    cfg = LargeAvalancheTransitionConfig()
    bridge = DeviceLargeAvalancheBridge(state, material, grid, TerrainVolumeIntegrator.from_grid(grid), cfg)
    got = bridge._persistence_summary()
    if precision == "fp32":
        initial_o = initial.astype(np.float32).astype(np.float64)
        resting_o = resting.astype(np.float32).astype(np.float64)
        mobile_o = np.rint(mobile / Q) * Q
        qx_o = qx.astype(np.float32).astype(np.float64)
        qy_o = qy.astype(np.float32).astype(np.float64)
    else:
        initial_o, resting_o, mobile_o, qx_o, qy_o = initial, resting, mobile, qx, qy
    w = oracle.vertex_weights(ny, nx, 0.05, 0.05)
    stop, start = np.deg2rad(30.0), np.deg2rad(38.0)
    di, df, flags = oracle.persistence(initial_o, resting_o, mobile_o, qx_o, qy_o, w, slope, unstable, latch,
                                       count, r2m, m2r, tile_size=tile, start_rad=start, stop_rad=stop,
                                       dry=cfg.dry_tolerance_m, activity_speed=cfg.mobile_activity_speed_m_s)
    np.testing.assert_array_equal(rt.arrays["avalanche_activity_tile_flags"].cpu().numpy(), flags)
    assert got["active_tile_count"] == int(np.count_nonzero(flags))
    assert got["unique_cells_ever_activated"] == di[0]
    assert got["cells_activated_more_than_once"] == di[1]
    assert got["maximum_mobile_speed_m_s"] == df[5]
    assert got["maximum_activated_slope_deg"] == float(np.rad2deg(df[9]))
    sums = {"cumulative_resting_to_mobile_m3": df[0], "cumulative_mobile_to_resting_m3": df[1],
            "unique_activated_area_m2": df[2], "reactivated_area_m2": df[3],
            "current_mobile_volume_m3": df[4], "net_terrain_volume_change_m3": df[6],
            "net_spatial_transfer_m3": df[7], "retired_candidate_area_m2": df[12],
            "moving_mobile_volume_m3": df[13],
            "mean_activated_slope_deg": float(np.rad2deg(df[8] / di[0])),
            "mean_slope_minus_theta_stop_deg": float(np.rad2deg(df[10] / di[0])),
            "mean_theta_start_minus_slope_deg": float(np.rad2deg(df[11] / di[0]))}
    for key, expected in sums.items():
        scale = max(abs(expected), float(np.sum(w)) * 1e-3)
        assert abs(got[key] - expected) <= 1e-11 * scale, key
