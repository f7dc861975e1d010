"""The oracle's restatement of the residual projection rounds
(oracle.frontier_rounds, v2_physics_core.py:1172-1254) pinned against the
reference's compact-frontier tests and the relaxation oracle."""

from __future__ import annotations

import numpy as np
import pytest

import oracle

TAN30 = float(np.tan(np.deg2rad(30.0)))
DX = 0.05


def _neighbors(dx=DX, dy=DX):
    return [(1, 0, dx), (0, 1, dy), (1, 1, float(np.hypot(dx, dy))), (1, -1, float(np.hypot(dx, dy)))]


def _tiles_around(i, j, shape, ts):
    tiles_x = (shape[1] + ts - 1) // ts
    out = set()
    for oi in (-1, 0, 1):
        for oj in (-1, 0, 1):
            r, c = i + oi, j + oj
            if 0 <= r < shape[0] and 0 <= c < shape[1]:
                out.add((r // ts) * tiles_x + c // ts)
    return np.asarray(sorted(out), dtype=np.int32)


def test_edge_batches_cover_each_full_grid_edge_once():
    """tests/test_sparse_tile_slope.py:34-63 on the oracle's batch builder."""
    shape, ts = (17, 19), 8
    tiles = np.arange(((17 + 7) // 8) * ((19 + 7) // 8))
    for (di, dj), expected in zip(((1, 0), (0, 1), (1, 1), (1, -1)),
                                  (16 * 19, 17 * 18, 16 * 18, 16 * 18)):
        edges = []
        for phase in (0, 1):
            _, ai, aj, bi, bj = oracle.edge_batch(shape, ts, tiles, di, dj, phase)
            edges.extend(zip(ai.tolist(), aj.tolist(), bi.tolist(), bj.tolist()))
        assert len(edges) == expected and len(set(edges)) == expected


def _spike(n, at):
    h = np.ones((n, n))
    h[at, at] += 0.30
    reached = np.zeros((n, n), dtype=np.int32)
    reached[at, at] = 1
    return h, reached


def test_local_event_matches_the_full_relaxation():
    """tests/test_sparse_tile_slope.py:65-91: a 0.30 m spike on a flat
    129 x 129 field, tile 16 -- the compact rounds reach the full
    relaxation's field within 2e-8 m while touching few tiles."""
    n, ts = 129, 16
    h, reached = _spike(n, 64)
    w = np.ones((n, n))
    out, r_out, rounds, active, moved, cleared = oracle.frontier_rounds(
        h, w, reached, _tiles_around(64, 64, (n, n), ts), tile_size=ts, neighbors=_neighbors(),
        critical_tangent=TAN30, tolerance=1e-8, max_rounds=4096)
    # the rounds address H[axis0, axis1] with (1, 0) along axis 0: the same
    # edges and colour order as relax_critical_slope on the array as laid out
    full, it, conv, _ = oracle.relax_legacy(h, DX, DX, TAN30, 20000, 1e-8)
    assert conv and cleared and active.size == 0 and rounds > 1
    np.testing.assert_allclose(out, full, rtol=0.0, atol=2e-8)
    assert np.all(r_out == 0)
    assert abs(out.sum() - h.sum()) <= 1e-9
    assert np.count_nonzero(out != h) < 0.25 * n * n


def test_round_budget_resumes_to_the_complete_result():
    """tests/test_sparse_tile_slope.py:93-130: one round per call keeps the
    volume and ends on the complete solver's field (here bit for bit)."""
    n, ts = 65, 8
    h, reached = _spike(n, 32)
    w = oracle.vertex_weights(n, n, DX, DX)
    kw = dict(tile_size=ts, neighbors=_neighbors(), critical_tangent=TAN30, tolerance=1e-8)
    seeds = _tiles_around(32, 32, (n, n), ts)
    complete, r_c, rounds_c, _, _, _ = oracle.frontier_rounds(h, w, reached, seeds, max_rounds=4096, **kw)
    cur, r, active = h, reached, seeds
    calls, volume = 0, float(np.sum(h * w))
    while True:
        cur, r, rounds, active, _, cleared = oracle.frontier_rounds(cur, w, r, active, max_rounds=1, **kw)
        calls += 1
        assert rounds == 1
        assert abs(float(np.sum(cur * w)) - volume) <= 1e-12 * volume
        if calls == 1:
            assert active.size > 0
        if active.size == 0:
            break
    assert calls == rounds_c
    np.testing.assert_array_equal(cur, complete)


def test_converged_field_has_no_reached_edge_above_tolerance():
    """Every edge whose ends changed is owned by a tile the round scans (a
    transfer activates the 3x3 tile halo of both ends), so at convergence no
    edge anywhere exceeds the critical difference by more than the tolerance."""
    rng = np.random.default_rng(4)
    n, ts = 48, 8
    h = 1.0 + np.cumsum(rng.normal(0.0, 0.03, (n, n)), axis=1) * 0.5
    reached = np.ones((n, n), dtype=np.int32)
    w = oracle.vertex_weights(n, n, DX, DX)
    tiles = np.arange(((n + ts - 1) // ts) ** 2)
    out, _, rounds, active, moved, cleared = oracle.frontier_rounds(
        h, w, reached, tiles, tile_size=ts, neighbors=_neighbors(), critical_tangent=TAN30,
        tolerance=1e-8, max_rounds=4096)
    assert cleared and rounds > 1 and moved > 0
    for di, dj, dist in _neighbors():
        a = out[max(0, -di):n - max(0, di), max(0, -dj):n - max(0, dj)]
        b = out[max(0, di):n + min(0, di) or None, max(0, dj):n + min(0, dj) or None]
        assert np.max(np.abs(a - b)) - TAN30 * dist <= 1e-8 + 1e-15
    assert float(np.sum(out * w)) == pytest.approx(float(np.sum(h * w)), rel=1e-12)


def test_empty_active_set_runs_no_round():
    n, ts = 16, 4
    h, reached = _spike(n, 8)
    out, r, rounds, active, moved, cleared = oracle.frontier_rounds(
        h, np.ones((n, n)), reached, [], tile_size=ts, neighbors=_neighbors(), critical_tangent=TAN30,
        tolerance=1e-8, max_rounds=10)
    assert rounds == 0 and not cleared and moved == 0
    np.testing.assert_array_equal(out, h)
    np.testing.assert_array_equal(r, reached)
