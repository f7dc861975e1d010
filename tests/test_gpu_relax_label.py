"""GPU parity of relaxation, compact frontier sweeps and labelling vs the oracle."""

from __future__ import annotations

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


def _device(a, dtype):
    import torch

    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda:0", dtype)


@pytest.mark.parametrize("shape,angle,seed", [((40, 37), 34.0, 0), ((65, 130), 30.0, 1),
                                              ((2, 2), 34.0, 2), ((97, 33), 38.0, 3)])
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_relax_bit_exact(shape, angle, seed, precision):
    import torch
    from isaac_bulk_hip import relax_device

    rng = np.random.default_rng(seed)
    h = rng.uniform(0.0, 1.5, size=shape)
    dtype = torch.float64 if precision == "fp64" else torch.float32
    npdt = np.float64 if precision == "fp64" else np.float32
    t = _device(h.astype(npdt), dtype)
    it, conv, last = relax_device(t, 0.05, 0.05, angle, 5000, 1e-8 if precision == "fp64" else 1e-6)
    crit = float(np.tan(np.deg2rad(angle)))
    o, oit, oconv, olast = oracle.relax_yx(h.astype(npdt), 0.05, 0.05, crit, 5000,
                                           1e-8 if precision == "fp64" else 1e-6, dtype=npdt)
    assert (it, conv) == (oit, oconv)
    np.testing.assert_array_equal(t.cpu().numpy(), o)


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_relax_converges_at_either_iteration_parity(precision):
    """A launch applies one iteration, then two (FP32).  Convergence on the
    first of two re-applies that one iteration to the launch's input; on the
    second it ends there.  Both must return the oracle's field, iteration
    count and last maximum excess exactly.  The seeds converge after 6, 7, 8
    and 5 iterations (checked against the oracle below)."""
    import torch
    from isaac_bulk_hip import relax_device

    dtype = torch.float64 if precision == "fp64" else torch.float32
    npdt = np.float64 if precision == "fp64" else np.float32
    crit = float(np.tan(np.deg2rad(34.0)))
    counts = set()
    for seed in (0, 2, 5, 6, 0):                 # the repeat follows a call that needed more launches
        h = np.random.default_rng(100 + seed).uniform(0.0, 0.3, size=(48, 56)).astype(npdt)
        t = _device(h, dtype)
        it, conv, last = relax_device(t, 0.05, 0.05, 34.0, 5000, 1e-3)
        o, oit, oconv, olast = oracle.relax_yx(h, 0.05, 0.05, crit, 5000, 1e-3, dtype=npdt)
        assert (it, conv, last) == (oit, oconv, olast), seed
        np.testing.assert_array_equal(t.cpu().numpy(), o)
        counts.add(it)
    assert counts == {5, 6, 7, 8}
    # a second call on the relaxed field converges in its first iteration
    it, conv, _ = relax_device(t, 0.05, 0.05, 34.0, 5000, 1e-3)
    assert (it, conv) == (1, True)


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_relax_iteration_limits(precision):
    """max_iterations cuts the loop inside a two-iteration launch or at its
    end: the field after exactly that many iterations, not converged."""
    import torch
    from isaac_bulk_hip import relax_device

    dtype = torch.float64 if precision == "fp64" else torch.float32
    npdt = np.float64 if precision == "fp64" else np.float32
    crit = float(np.tan(np.deg2rad(30.0)))
    h = np.random.default_rng(11).uniform(0.0, 1.5, size=(61, 90)).astype(npdt)
    for limit in (1, 2, 3, 4, 5, 8):
        t = _device(h, dtype)
        it, conv, last = relax_device(t, 0.05, 0.05, 30.0, limit, 1e-12)
        o, oit, oconv, olast = oracle.relax_yx(h, 0.05, 0.05, crit, limit, 1e-12, dtype=npdt)
        assert (it, conv, last) == (oit, oconv, olast) == (limit, False, olast), limit
        np.testing.assert_array_equal(t.cpu().numpy(), o)


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_relax_marked_reports_the_excess_of_the_marked_iteration(precision):
    """mv2_relax_marked: the same field, iteration count and convergence as
    the plain relaxation, and the maximum excess of iteration `mark` -- what a
    call limited to `mark` iterations returns as its last excess (the oracle
    with that limit), also when the loop converges before or at the mark."""
    import ctypes

    import torch
    from isaac_bulk_hip import _lib
    from isaac_bulk_hip.relaxation import _runtime, legacy_critical_differences

    rt = _runtime("cuda:0")
    dtype = torch.float64 if precision == "fp64" else torch.float32
    npdt = np.float64 if precision == "fp64" else np.float32
    prec = _lib.MV2_FP64 if precision == "fp64" else _lib.MV2_FP32
    crit_t = float(np.tan(np.deg2rad(30.0)))
    crit = legacy_critical_differences(0.05, 0.05, 30.0)
    h = np.random.default_rng(12).uniform(0.0, 1.5, size=(61, 90)).astype(npdt)
    tol = 1e-3
    o_full, it_full, conv_full, last_full = oracle.relax_yx(h, 0.05, 0.05, crit_t, 5000, tol, dtype=npdt)
    assert conv_full and it_full > 6
    for mark in (1, 2, 3, 4, 5, it_full - 1, it_full, it_full + 3):
        t = _device(h, dtype)
        res = _lib.RelaxResult()
        marked = ctypes.c_double(-1.0)
        rt.bind_stream()
        rt.check(rt.lib.mv2_relax_marked(rt.ctx, prec, t.data_ptr(), 61, 90, crit.ctypes.data, 5000, tol, mark,
                                         ctypes.byref(res), ctypes.byref(marked)))
        assert (res.iterations, bool(res.converged), res.max_excess_last) == (it_full, conv_full, last_full), mark
        np.testing.assert_array_equal(t.cpu().numpy(), o_full)
        _, _, _, last_m = oracle.relax_yx(h, 0.05, 0.05, crit_t, mark, tol, dtype=npdt)
        assert marked.value == last_m, (mark, marked.value, last_m)


def test_relax_toe_701_iteration_count_and_field():
    import torch
    from isaac_bulk_hip import relax_device

    nx, dx = 701, 0.05
    drop = np.tan(np.deg2rad(30.0)) * dx
    base = np.repeat((1.0 + np.arange(nx) * 0.95 * drop)[None, :], nx, axis=0)
    toe = base.copy()
    toe[346:355, 0] -= 0.10
    t = _device(toe, torch.float64)
    it, conv, _ = relax_device(t, dx, dx, 30.0, 5000, 1e-8)
    assert (it, conv) == (183, True)
    o, oit, _, _ = oracle.relax_yx(toe, dx, dx, float(np.tan(np.deg2rad(30.0))), 5000, 1e-8)
    np.testing.assert_array_equal(t.cpu().numpy(), o)


def test_relax_host_api_legacy_layout():
    from isaac_bulk_hip import relax_critical_slope

    h = np.zeros((31, 31))
    h[15, 15] = 10.0
    out, stats = relax_critical_slope(h, (0.1, 0.1), 30.0)
    o, it, conv, _ = oracle.relax_legacy(h, 0.1, 0.1, float(np.tan(np.deg2rad(30.0))))
    assert stats.converged and stats.iterations == it
    np.testing.assert_array_equal(out, o)
    assert stats.max_slope <= np.tan(np.deg2rad(30.0)) + 1e-7


@pytest.mark.parametrize("density", [0.0, 0.3, 0.5, 0.6, 1.0])
@pytest.mark.parametrize("connectivity", [4, 8])
def test_label_mask_bit_exact(density, connectivity):
    from isaac_bulk_hip import label_mask

    rng = np.random.default_rng(int(density * 100) + connectivity)
    for shape in ((64, 64), (1, 17), (130, 257)):
        mask = (rng.random(shape) < density).astype(np.int32)
        got = label_mask(mask, connectivity=connectivity)
        np.testing.assert_array_equal(got, oracle.label(mask, connectivity))


def test_label_large_spiral_single_component():
    """A long serpentine path is one component whose root is its minimum index."""
    from isaac_bulk_hip import label_mask

    n = 301
    mask = np.zeros((n, n), np.int32)
    mask[::2, :] = 1
    for r in range(1, n, 2):
        mask[r, n - 1 if (r // 2) % 2 == 0 else 0] = 1
    got = label_mask(mask, connectivity=4)
    assert set(np.unique(got)) == {-1, 0}
    np.testing.assert_array_equal(got, oracle.label(mask, 4))


def _avalanche_state(seed=0, n=96, precision="fp64"):
    from isaac_bulk_hip import DeviceBulkState, TerrainGrid
    from isaac_bulk_hip.workloads import cut_face_terrain

    b, h, _ = cut_face_terrain(n, seed=seed, face_deg=45.0)
    grid = TerrainGrid(n, n, 0.05, 0.05)
    state = DeviceBulkState(grid, b, precision=precision)
    return state, grid, b


def test_avalanche_label_and_mobilize_match_oracle():
    from isaac_bulk_hip import (DeviceLargeAvalancheBridge, LargeAvalancheTransitionConfig,
                                MaterialScenario, TerrainVolumeIntegrator)

    state, grid, b = _avalanche_state()
    material = MaterialScenario("cohesionless", 2200.0, 34.0, 0.0, 0.45, 38.0, 30.0, 0.35)
    cfg = LargeAvalancheTransitionConfig(minimum_connected_cells=16, minimum_connected_area_m2=0.02,
                                         minimum_mobilizable_volume_m3=0.001, persistence_time_s=0.01,
                                         mobilization_depth_m=0.05)
    bridge = DeviceLargeAvalancheBridge(state, material, grid, TerrainVolumeIntegrator.from_grid(grid), cfg)
    res = bridge.observe_and_maybe_mobilize(0.02)
    n = grid.nx
    z = np.zeros((n, n))
    w = TerrainVolumeIntegrator.from_grid(grid).vertex_weights_m2
    o = oracle.avalanche_initialize(
        z, b, z, z, z, w, z, z, np.zeros((n, n), np.int32), dx=0.05, dy=0.05,
        tan_start=float(np.tan(np.deg2rad(38.0))), tan_stop=float(np.tan(np.deg2rad(30.0))),
        density=2200.0, gravity=9.81, cohesion=0.0, mobilization_depth=0.05, dry=1e-8,
        activity_speed=0.05, release=True)
    a = state.runtime.arrays
    np.testing.assert_array_equal(a["avalanche_unstable"].cpu().numpy().reshape(n, n), o["unstable"])
    lab = oracle.label(o["unstable"], 8)
    di, df = oracle.components(lab, z, b, np.zeros((n, n), np.int32), w, o["slope"], o["severity"],
                               np.zeros((n, n), np.int32), 0.05, float(np.deg2rad(38.0)),
                               float(np.deg2rad(30.0)))
    assert res.unstable_cell_count == di[3] > 0
    assert res.connected_region_count == di[0]
    assert res.largest_connected_cell_count == di[1]
    assert bridge.last_label.diag_i[2] == di[2]
    assert res.largest_connected_area_m2 == pytest.approx(df[4], rel=1e-12)
    assert res.largest_connected_mobilizable_volume_m3 == pytest.approx(df[5], rel=1e-12)
    assert res.classification == bridge.LARGE_AVALANCHE_MOBILE_PATH
    mob = oracle.mobilize(lab, di[2], z, b, z, w, np.zeros((n, n), np.int32), z, z, z, o["severity"],
                          np.zeros((n, n), np.int32), z, z, z, 0.0, 0.05, 1e-8)
    view = state.explicit_host_view(source="acceptance")
    np.testing.assert_array_equal(view.b_eff_m, mob["resting"])
    np.testing.assert_array_equal(view.H_mobile_m, mob["mobile"])
    assert res.transferred_volume_m3 == pytest.approx(mob["diag"][7], rel=1e-12)


def test_frontier_phase_matches_oracle():
    import torch
    from isaac_bulk_hip import CompactTileFrontier, HipCompactActiveEdgeOperator

    rng = np.random.default_rng(4)
    shape = (70, 90)
    h = rng.uniform(0.0, 1.0, size=shape)
    reached = (rng.random(shape) < 0.3).astype(np.int32)
    w = rng.uniform(0.5, 1.5, size=shape)
    op = HipCompactActiveEdgeOperator(shape, tile_size=32)
    op.initialize(h, reached, w)
    frontier = CompactTileFrontier(shape, 32)
    ho, ro = h.copy(), reached.copy()
    crit = float(np.tan(np.deg2rad(30.0))) * 0.05
    for d, (di, dj) in enumerate(((1, 0), (0, 1), (1, 1), (1, -1))):
        for phase in (0, 1):
            batch = frontier.edge_batch(np.arange(frontier.tile_count), direction_index=d, di=di,
                                        dj=dj, phase=phase)
            r = op.run_phase(batch, critical_difference_m=crit, tolerance_m=1e-8)
            ho, ro, moved, flags = oracle.frontier_relax(ho, w, ro, batch.a_i, batch.a_j, batch.b_i,
                                                         batch.b_j, crit, 32, frontier.tile_shape[1])
            uflags = oracle.frontier_scan(ho, ro, batch.tile_ids, batch.a_i, batch.a_j, batch.b_i,
                                          batch.b_j, crit, 1e-8, frontier.tile_count)
            assert r.moved_edge_count == moved
            assert r.triggered_tile_ids.tolist() == np.flatnonzero(flags).tolist()
            assert r.unstable_owner_tile_ids.tolist() == np.flatnonzero(uflags).tolist()
    hd, rd = op.download_state()
    np.testing.assert_array_equal(hd, ho)
    np.testing.assert_array_equal(rd, ro.astype(bool))


@pytest.mark.parametrize("ring", [0.0, 0.4])
def test_open_boundary_adapter_matches_oracle(ring):
    """Open boundary (minimum_slope_adapter.py:160-181, :185-212): per
    iteration the field is padded with a ring at the boundary height, one
    relaxation iteration runs, the interior is kept; the device loop equals
    the oracle's repetition of that bit for bit, and the outflow is the
    interior volume lost."""
    from isaac_bulk_hip import HipMinimumSlopeAdapter, SolverConfig, TerrainGrid

    rng = np.random.default_rng(11)
    ny, nx, dx = 30, 41, 0.05
    h = 0.5 + rng.uniform(0.0, 0.3, (ny, nx))
    h[:6, :8] += 1.2                                  # a steep block against the corner
    grid = TerrainGrid(nx, ny, dx, dx)
    cfg = SolverConfig(critical_angle_deg=33.0, max_iterations=4000, tolerance=1e-8, boundary_condition="open",
                       boundary_height_m=ring)
    res = HipMinimumSlopeAdapter(grid, cfg).solve(h)
    crit = float(np.tan(np.deg2rad(33.0)))
    cur, it, conv = h.copy(), 0, False
    while it < 4000 and not conv:
        it += 1
        padded = np.pad(cur, 1, mode="constant", constant_values=ring)
        out, _, conv, _ = oracle.relax_yx(padded, dx, dx, crit, 1, 1e-8)
        cur = out[1:-1, 1:-1]
    assert res.converged and conv and res.iteration_count == it
    np.testing.assert_array_equal(res.heightmap_stable, cur)
    lost = (h.sum() - cur.sum()) * dx * dx
    assert res.boundary_outflow_m3 == pytest.approx(max(lost, 0.0), rel=1e-9)
    assert res.boundary_outflow_m3 > 0.0
