"""GPU parity of the on-device residual projection rounds (mv2_frontier_rounds)
against the oracle's round loop and the host-driven per-phase loop."""

from __future__ import annotations

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

TAN30 = float(np.tan(np.deg2rad(30.0)))
DX = 0.05
NEIGHBORS = [(1, 0, DX), (0, 1, DX), (1, 1, float(np.hypot(DX, DX))), (1, -1, float(np.hypot(DX, DX)))]


def _operator(h, reached, w, ts):
    from isaac_bulk_hip import HipCompactActiveEdgeOperator

    op = HipCompactActiveEdgeOperator(h.shape, tile_size=ts)
    op.initialize(h, reached, w)
    return op


def _tiles_around(i, j, shape, ts):
    from isaac_bulk_hip import CompactTileFrontier

    return CompactTileFrontier(shape, ts).tiles_around_cells([i], [j])


def _spike(n, at):
    h = np.ones((n, n))
    h[at, at] += 0.30
    reached = np.zeros((n, n), dtype=np.int32)
    reached[at, at] = 1
    return h, reached


def _rough(ny, nx, seed):
    rng = np.random.default_rng(seed)
    h = 1.0 + np.cumsum(rng.normal(0.0, 0.03, (ny, nx)), axis=1) * 0.5
    return h, np.ones((ny, nx), dtype=np.int32)


def _compare(op, res, ref):
    h_ref, r_ref, rounds, active, moved, cleared = ref
    h, r = op.download_state()
    np.testing.assert_array_equal(h, h_ref)
    np.testing.assert_array_equal(r.astype(np.int32), r_ref)
    assert res.rounds == rounds
    assert res.moved_edge_count == moved
    assert res.cleared_reached == cleared
    np.testing.assert_array_equal(res.active_tile_ids, active)
    assert res.converged == (active.size == 0)


@pytest.mark.parametrize("case", ["spike129_tile16", "rough97x83_tile8", "rough64_tile64"])
def test_rounds_bit_exact_vs_oracle(case):
    print(f"[frontier rounds] {case}", flush=True)
    if case == "spike129_tile16":
        h, reached = _spike(129, 64)
        w, ts = np.ones(h.shape), 16
        seeds = _tiles_around(64, 64, h.shape, ts)
    else:
        ny, nx, ts = (97, 83, 8) if case == "rough97x83_tile8" else (64, 64, 64)
        h, reached = _rough(ny, nx, 7)
        w = oracle.vertex_weights(ny, nx, DX, DX)
        seeds = np.arange(((ny + ts - 1) // ts) * ((nx + ts - 1) // ts))
    ref = oracle.frontier_rounds(h, w, reached, seeds, tile_size=ts, neighbors=NEIGHBORS,
                                 critical_tangent=TAN30, tolerance=1e-8, max_rounds=4096)
    op = _operator(h, reached, w, ts)
    res = op.advance_rounds(seeds, neighbors=NEIGHBORS, critical_tangent=TAN30, tolerance_m=1e-8,
                            max_rounds=4096)
    assert res.converged and res.rounds > 1
    _compare(op, res, ref)


def test_rounds_match_host_driven_phase_loop():
    """The reference's host loop (v2_physics_core.py:1183-1254) driving the
    per-phase entry points gives the same bits as the device rounds."""
    from isaac_bulk_hip import CompactTileFrontier

    n, ts = 80, 16
    h, reached = _spike(n, 37)
    h[10:14, 60:70] += 0.2
    reached[10:14, 60:70] = 1
    w = oracle.vertex_weights(n, n, DX, DX)
    frontier = CompactTileFrontier((n, n), ts)
    seeds = frontier.tiles_around_cells(*np.nonzero(reached))
    host = _operator(h, reached, w, ts)
    active, rounds = seeds.astype(np.int32), 0
    while active.size:
        rounds += 1
        phase_tiles = active
        for d, (di, dj, dist) in enumerate(NEIGHBORS):
            for phase in (0, 1):
                batch = frontier.edge_batch(phase_tiles, direction_index=d, di=di, dj=dj, phase=phase)
                r = host.relax_phase(batch, critical_difference_m=TAN30 * dist)
                if r.triggered_tile_ids.size:
                    phase_tiles = np.union1d(phase_tiles, r.triggered_tile_ids).astype(np.int32)
        owners = []
        for d, (di, dj, dist) in enumerate(NEIGHBORS):
            for phase in (0, 1):
                batch = frontier.edge_batch(phase_tiles, direction_index=d, di=di, dj=dj, phase=phase)
                r = host.scan_phase(batch, critical_difference_m=TAN30 * dist, tolerance_m=1e-8)
                if r.unstable_owner_tile_ids.size:
                    owners.append(r.unstable_owner_tile_ids)
        active = np.unique(np.concatenate(owners)).astype(np.int32) if owners else np.empty(0, np.int32)
    h_host, _ = host.download_state()
    dev = _operator(h, reached, w, ts)
    res = dev.advance_rounds(seeds, neighbors=NEIGHBORS, critical_tangent=TAN30, tolerance_m=1e-8,
                             max_rounds=4096)
    h_dev, r_dev = dev.download_state()
    assert res.rounds == rounds and res.converged and res.cleared_reached
    np.testing.assert_array_equal(h_dev, h_host)
    assert not r_dev.any()


def test_round_budget_and_resume():
    """max_rounds=1 per call: each call runs one round, reports the tiles
    still active, and the calls end on the single call's field."""
    n, ts = 65, 8
    h, reached = _spike(n, 32)
    w = oracle.vertex_weights(n, n, DX, DX)
    seeds = _tiles_around(32, 32, (n, n), ts)
    kw = dict(neighbors=NEIGHBORS, critical_tangent=TAN30, tolerance_m=1e-8)
    complete = _operator(h, reached, w, ts)
    full = complete.advance_rounds(seeds, max_rounds=4096, **kw)
    step = _operator(h, reached, w, ts)
    ref = (h, reached, seeds)
    active, calls = seeds, 0
    while True:
        res = step.advance_rounds(active, max_rounds=1, **kw)
        o = oracle.frontier_rounds(ref[0], w, ref[1], ref[2], tile_size=ts, neighbors=NEIGHBORS,
                                   critical_tangent=TAN30, tolerance=1e-8, max_rounds=1)
        _compare(step, res, o)
        ref = (o[0], o[1], o[3])
        calls += 1
        active = res.active_tile_ids
        if res.converged:
            break
    assert calls == full.rounds
    np.testing.assert_array_equal(step.download_state()[0], complete.download_state()[0])


def test_residual_frontier_seeds_and_errors():
    import torch

    from isaac_bulk_hip import GpuResidualFrontier

    n, ts = 65, 8
    h = np.ones((n, n))
    h[32, 32] += 0.30
    op = _operator(h, np.zeros((n, n), dtype=np.int32), np.ones((n, n)), ts)
    res_frontier = GpuResidualFrontier(op, stop_angle_deg=30.0, dx=DX, dy=DX)
    assert res_frontier.advance().rounds == 0            # nothing seeded: no round
    res_frontier.seed_cells([32 * n + 32])
    assert res_frontier.active_tile_ids.size > 0
    result = res_frontier.advance()
    assert result.converged and result.rounds > 1 and res_frontier.iterations == result.rounds
    assert int(torch.count_nonzero(op.runtime.arrays["reached"]).item()) == 0
    assert res_frontier.active_tile_ids.size == 0

    op2 = _operator(h, np.zeros((n, n), dtype=np.int32), np.ones((n, n)), ts)
    bounded = GpuResidualFrontier(op2, stop_angle_deg=30.0, dx=DX, dy=DX, max_rounds=1)
    bounded.seed_cells([32 * n + 32])
    with pytest.raises(RuntimeError, match="exceeded 1 bounded numerical rounds"):
        bounded.advance()
    op3 = _operator(h, np.zeros((n, n), dtype=np.int32), np.ones((n, n)), ts)
    safety = GpuResidualFrontier(op3, stop_angle_deg=30.0, dx=DX, dy=DX,
                                 numerical_safety_max_iterations=1)
    safety.seed_cells([32 * n + 32])
    with pytest.raises(RuntimeError, match="numerical safety limit"):
        safety.advance()
    assert safety.iterations == 1
