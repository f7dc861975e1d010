"""GPU surface publication: dirty-tile tracking, tile gather and bilinear
queries (mv2_surface_*) against NumPy and the oracle."""

from __future__ import annotations

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

Q = 1.0 / 16777216.0


def _state(precision, shape=(53, 71), tile=16, seed=0):
    from isaac_bulk_hip import DeviceBulkState, TerrainGrid

    rng = np.random.default_rng(seed)
    ny, nx = shape
    grid = TerrainGrid(nx, ny, 0.05, 0.05, terrain_prim_path="/Terrain")
    state = DeviceBulkState(grid, 1.0 + rng.random(shape), tile_size=tile, precision=precision)
    state.set_fields(mobile=np.where(rng.random(shape) < 0.5, rng.random(shape) * 0.1, 0.0))
    return state


def _surface(state):
    f = state.download_fields(("b_eff", "mobile"))
    return f["b_eff"] + f["mobile"]


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_dirty_tiles_follow_the_surface_change(precision):
    """An entrainment moves material between b_eff and mobile (surface kept):
    no tile is dirty; a deposit-free change of b_eff at chosen cells flags
    exactly their tiles, in ascending order."""
    import torch

    state = _state(precision)
    state.capture_surface_for_dirty_tracking()
    state.consume_dirty_tile_ids()
    idx = np.array([5 * 71 + 7, 40 * 71 + 60])
    state.entrain_host_indices(idx, np.array([0.01, 0.02]), reason="test")
    state.consume_dirty_tile_ids()
    tol = 1e-10 if precision == "fp64" else 1e-6
    assert state.collect_surface_dirty_tiles(tolerance_m=tol).size == 0
    state.capture_surface_for_dirty_tracking()
    b = state.runtime.arrays["b_eff"]
    cells = [(0, 0), (17, 33), (52, 70), (31, 16)]
    for r, c in cells:
        b[r * 71 + c] += 0.05
    ids = state.collect_surface_dirty_tiles(tolerance_m=tol)
    tiles_x = state.tile_shape[1]
    expected = sorted({(r // 16) * tiles_x + c // 16 for r, c in cells})
    assert ids.tolist() == expected
    flags = state.runtime.arrays["dirty_tile_flags"].cpu().numpy()
    assert np.flatnonzero(flags).tolist() == expected
    assert sorted(state.consume_dirty_tile_ids().tolist()) == expected
    del torch


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_tile_gather_with_overlap(precision):
    state = _state(precision, shape=(50, 67), tile=16)
    surface = _surface(state)
    tiles = [0, 3, 7, 12, 15]
    got = state.download_dirty_tiles(tiles)
    tiles_x = state.tile_shape[1]
    for t in tiles:
        r0, c0 = (t // tiles_x) * 16, (t % tiles_x) * 16
        np.testing.assert_array_equal(got[t], surface[r0:min(r0 + 17, 50), c0:min(c0 + 17, 67)])
    state.mark_dirty_tiles([2, 9])
    assert sorted(state.download_dirty_tiles()) == [2, 9]
    with pytest.raises(ValueError):
        state.download_dirty_tiles([state.tile_count])


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_bilinear_queries_bit_exact(precision):
    state = _state(precision, seed=3)
    f = state.download_fields(("b_eff", "mobile"))
    rng = np.random.default_rng(7)
    q = np.column_stack((rng.uniform(-3, 56, 5000), rng.uniform(-3, 74, 5000)))
    q[:10] = np.floor(q[:10])
    got = state.sample_surface_bilinear(q, source="airborne")
    np.testing.assert_array_equal(got, oracle.surface_sample(f["b_eff"], f["mobile"], q))
    assert state.sample_surface_bilinear(np.empty((0, 2)), source="x").size == 0
    with pytest.raises(ValueError):
        state.sample_surface_bilinear(q, source="")
