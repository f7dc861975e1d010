"""Device checkpoint round trips (write_device_checkpoint / restore_device_checkpoint)."""

from __future__ import annotations

import json

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

Q = 1.0 / 16777216.0


def _state(precision, seed=0, shape=(37, 45), tile=8):
    import torch

    from isaac_bulk_hip import DeviceBulkState, TerrainGrid

    ny, nx = shape
    rng = np.random.default_rng(seed)
    grid = TerrainGrid(nx, ny, 0.05, 0.05, terrain_prim_path="/Terrain")
    state = DeviceBulkState(grid, 1.0 + rng.random(shape), tile_size=tile, precision=precision)
    state.set_fields(z_base=rng.random(shape) * 0.5, mobile=rng.random(shape) * 0.2,
                     momentum_x=rng.normal(0, 0.1, shape), momentum_y=rng.normal(0, 0.1, shape))
    rt = state.runtime
    for name in state._F64_FIELDS:
        rt.arrays[name].copy_(torch.from_numpy(rng.random(state.size)).to(rt.device))
    for name in state._INT_FIELDS:
        rt.arrays[name].copy_(torch.from_numpy(rng.integers(0, 3, state.size).astype(np.int32)).to(rt.device))
    state.advance_time(1.25)
    state.mark_dirty_tiles([1, 4])
    return state


def _all(state):
    from isaac_bulk_hip.checkpoint import FLOAT_FIELDS, INT_FIELDS

    return state.download_fields((*FLOAT_FIELDS, *INT_FIELDS))


def _bridge(state):
    from isaac_bulk_hip import (DeviceLargeAvalancheBridge, LargeAvalancheTransitionConfig, MaterialScenario,
                                TerrainVolumeIntegrator)

    material = MaterialScenario("ckpt", 1370.0, 29.8, 0.0, 0.5, 38.0, 30.0, 0.35)
    return DeviceLargeAvalancheBridge(state, material, state.grid, TerrainVolumeIntegrator.from_grid(state.grid),
                                      LargeAvalancheTransitionConfig())


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_round_trip_is_bit_exact(precision, tmp_path):
    from isaac_bulk_hip import restore_device_checkpoint, write_device_checkpoint
    from isaac_bulk_hip.checkpoint import sha256_file

    src = _state(precision)
    bridge = _bridge(src)
    bridge._persistence_s, bridge._has_previous = 0.75, True
    before = _all(src)
    dirty = sorted(src._dirty_tile_ids)
    rec = write_device_checkpoint(src, tmp_path / "ckpt", provenance={"case": precision},
                                  large_avalanche=bridge)
    assert rec["schema"] == "390F_PRODUCTION_DEVICE_CHECKPOINT/v2"
    assert rec["checkpoint_hash_sha256"] == sha256_file(rec["checkpoint_path"])
    dst = _state(precision, seed=9)
    dst_bridge = _bridge(dst)
    out = restore_device_checkpoint(dst, rec["checkpoint_path"], large_avalanche=dst_bridge)
    assert out["legacy_checkpoint_migration"] == "NONE_V2_NATIVE"
    after = _all(dst)
    for name, value in before.items():
        np.testing.assert_array_equal(after[name], value, err_msg=name)
    assert dst.timestamp_device_s == 1.25
    assert sorted(dst._dirty_tile_ids) == dirty
    assert (dst_bridge._persistence_s, dst_bridge._has_previous) == (0.75, True)
    if precision == "fp32":
        np.testing.assert_array_equal(np.rint(after["mobile"] / Q), np.rint(before["mobile"] / Q))


def test_cross_precision_and_replay(tmp_path):
    """An FP64 checkpoint restores into an FP32 state as the rounded fields,
    and a restored FP64 state continues the Mobile V2 step bit for bit."""
    from isaac_bulk_hip import (MobileV2Config, WarpProductionMobileV2Solver, restore_device_checkpoint,
                                write_device_checkpoint)

    src = _state("fp64", seed=2)
    before = _all(src)
    rec = write_device_checkpoint(src, tmp_path / "c64.npz", provenance={}, compress=False)
    f32 = _state("fp32", seed=4)
    restore_device_checkpoint(f32, rec["checkpoint_path"])
    got = _all(f32)
    np.testing.assert_array_equal(got["b_eff"], before["b_eff"].astype(np.float32).astype(np.float64))
    np.testing.assert_array_equal(got["mobile"], np.rint(before["mobile"] / Q) * Q)

    def run(state):
        solver = WarpProductionMobileV2Solver(runtime=state.runtime, config=MobileV2Config())
        solver.bind_device_state(state, _bridge(state).material, state.grid, None)
        for _ in range(3):
            solver.step_resident(1.0 / 60.0)
        return state.download_fields(("b_eff", "mobile", "momentum_x", "momentum_y"))

    copy = _state("fp64", seed=7)
    restore_device_checkpoint(copy, rec["checkpoint_path"])
    a, b = run(src), run(copy)
    for name in a:
        np.testing.assert_array_equal(a[name], b[name], err_msg=name)


def test_restore_checks(tmp_path):
    from isaac_bulk_hip import restore_device_checkpoint, write_device_checkpoint

    src = _state("fp64")
    rec = write_device_checkpoint(src, tmp_path / "ok.npz", provenance={})
    with pytest.raises(RuntimeError, match="DEVICE_CHECKPOINT_SHAPE_MISMATCH"):
        restore_device_checkpoint(_state("fp64", shape=(36, 45)), rec["checkpoint_path"])
    with np.load(rec["checkpoint_path"], allow_pickle=False) as archive:
        arrays = {k: archive[k] for k in archive.files}
    record = json.loads(str(arrays["checkpoint_record_json"]))
    record["terrain_hash_sha256"] = "0" * 64
    arrays["checkpoint_record_json"] = np.asarray(json.dumps(record))
    np.savez(tmp_path / "bad.npz", **arrays)
    with pytest.raises(RuntimeError, match="DEVICE_CHECKPOINT_TERRAIN_HASH_MISMATCH"):
        restore_device_checkpoint(_state("fp64"), tmp_path / "bad.npz")
    record["terrain_hash_sha256"] = json.loads(str(np.load(rec["checkpoint_path"])["checkpoint_record_json"]))[
        "terrain_hash_sha256"]
    record["resolution_xy_m"] = [0.1, 0.05]
    arrays["checkpoint_record_json"] = np.asarray(json.dumps(record))
    np.savez(tmp_path / "res.npz", **arrays)
    with pytest.raises(RuntimeError, match="DEVICE_CHECKPOINT_RESOLUTION_MISMATCH"):
        restore_device_checkpoint(_state("fp64"), tmp_path / "res.npz")


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_restore_independently_written_fixture(precision):
    """A checkpoint written by plain NumPy in the reference layout
    (tests/golden/make_checkpoint_fixture.py) restores onto the device: FP64
    fields bit for bit; FP32 fields as their float32 rounding, mobile as whole
    quanta (the fixture's depths are whole quanta, so exactly); record state
    (timestamp, persistence, dirty tiles) carried over."""
    from pathlib import Path

    from isaac_bulk_hip import DeviceBulkState, TerrainGrid, restore_device_checkpoint
    from isaac_bulk_hip.checkpoint import FLOAT_FIELDS, INT_FIELDS

    path = Path(__file__).parent / "golden" / "device_checkpoint_13x17.npz"
    grid = TerrainGrid(17, 13, 0.05, 0.04)
    state = DeviceBulkState(grid, np.ones((13, 17)), tile_size=8, precision=precision)
    bridge = _bridge(state)
    rec = restore_device_checkpoint(state, path, large_avalanche=bridge)
    assert rec["legacy_checkpoint_migration"] == "NONE_V2_NATIVE"
    assert state.timestamp_device_s == 3.5 and bridge._persistence_s == 0.25 and bridge._has_previous
    got = state.download_fields((*FLOAT_FIELDS, *INT_FIELDS))
    with np.load(path, allow_pickle=False) as archive:
        for name in (*FLOAT_FIELDS, *INT_FIELDS):
            want = np.asarray(archive[name])
            if precision == "fp32" and name in ("z_base", "b_eff", "momentum_x", "momentum_y"):
                want = want.astype(np.float32).astype(np.float64)
            np.testing.assert_array_equal(got[name], want, err_msg=name)
    assert state.consume_dirty_tile_ids().tolist() == [1, 2]
