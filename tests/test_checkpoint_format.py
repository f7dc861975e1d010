"""Host side of the device checkpoint format (device_checkpoint.py:21-212):
hash definition, legacy migrations and the frozen-boundary adapter."""

from __future__ import annotations

import hashlib

import numpy as np
import pytest

from isaac_bulk_hip import checkpoint as ck


def test_array_hash_covers_dtype_shape_and_bytes():
    a = np.arange(6, dtype=np.float64).reshape(2, 3)
    digest = hashlib.sha256(b"float64" + np.asarray([2, 3], np.int64).tobytes() + a.tobytes()).hexdigest()
    assert ck.sha256_array(a) == digest
    assert ck.sha256_array(a.reshape(3, 2)) != digest          # same bytes, other shape
    assert ck.sha256_array(a.astype(np.float32)) != digest
    assert ck.sha256_array(np.asfortranarray(a)) == digest      # hashed in C order


def _fields(shape=(4, 5)):
    rng = np.random.default_rng(0)
    return {"z_base": rng.random(shape), "b_eff": 1 + rng.random(shape), "mobile": rng.random(shape) * 0.1,
            "momentum_x": rng.random(shape), "momentum_y": rng.random(shape)}


def test_native_fields_pass_through():
    f = _fields()
    f["avalanche_owned_surface"] = np.ones((4, 5))
    out, migration = ck.checkpoint_fields(f)
    assert migration == "NONE_V2_NATIVE"
    np.testing.assert_array_equal(out["avalanche_owned_surface"], 1.0)
    for name in ("avalanche_owned_export_baseline", "mobile_export_cumulative"):
        np.testing.assert_array_equal(out[name], 0.0)


def test_legacy_resting_and_missing_z_base():
    f = _fields()
    legacy = {k: v for k, v in f.items() if k not in ("b_eff", "z_base")}
    legacy["resting"] = f["b_eff"]
    out, migration = ck.checkpoint_fields(legacy)
    assert migration.startswith("UNCALIBRATED_SNAPSHOT_COMPATIBILITY_ADAPTER:B_EFF=H_FREE_OLD")
    np.testing.assert_array_equal(out["b_eff"], f["b_eff"])
    np.testing.assert_array_equal(out["z_base"], 0.0)
    np.testing.assert_array_equal(out["avalanche_owned_surface"], f["b_eff"] + f["mobile"])
    no_z = {k: v for k, v in f.items() if k != "z_base"}
    out, migration = ck.checkpoint_fields(no_z)
    assert migration.endswith("B_EFF_DIRECT;IMPLICIT_Z_BASE=0")
    with pytest.raises(RuntimeError, match="DEVICE_CHECKPOINT_MISSING_STATIC_BED"):
        ck.checkpoint_fields({k: v for k, v in f.items() if k != "b_eff"})


def test_frozen_boundary_adapter(tmp_path):
    """tools/run_mobile_v2_frozen_5s.py:99-106: b_eff = (resting + h) - h,
    which may differ from the stored resting label in the last bit."""
    rng = np.random.default_rng(3)
    resting = 1.0 + rng.random((6, 7))
    h = rng.random((6, 7)) * 0.3
    qx, qy = rng.random((6, 7)), rng.random((6, 7))
    path = tmp_path / "frozen.npz"
    np.savez(path, PRE_resting=resting, PRE_mobile=h, PRE_momentum_x=qx, PRE_momentum_y=qy)
    b, hh, q, free = ck.load_frozen_boundary(path)
    np.testing.assert_array_equal(free, resting + h)
    np.testing.assert_array_equal(b, (resting + h) - h)
    np.testing.assert_array_equal(q[..., 0], qx)
    np.testing.assert_array_equal(hh, h)
    assert np.max(np.abs(b - resting)) <= 4e-16


def test_independently_written_fixture_loads_as_native_v2():
    """tests/golden/device_checkpoint_13x17.npz was written by plain NumPy in
    the reference layout (tests/golden/make_checkpoint_fixture.py): the loader
    reads every field unchanged, the record's terrain hash matches, and the
    field sums equal the writer's independent summary."""
    import json
    from pathlib import Path

    golden = Path(__file__).parent / "golden"
    summary = json.loads((golden / "device_checkpoint_13x17.json").read_text())
    with np.load(golden / "device_checkpoint_13x17.npz", allow_pickle=False) as archive:
        record = json.loads(str(archive["checkpoint_record_json"]))
        fields, migration = ck.checkpoint_fields(archive)
    assert migration == "NONE_V2_NATIVE"
    assert record["schema"] == "390F_PRODUCTION_DEVICE_CHECKPOINT/v2" and record["shape_yx"] == [13, 17]
    assert ck.sha256_array(fields["b_eff"]) == record["terrain_hash_sha256"]
    for name in ck.FLOAT_FIELDS:
        assert fields[name].dtype == np.float64 and float(np.sum(fields[name])) == summary[name]
    for name in ck.INT_FIELDS:
        assert fields[name].dtype == np.int32 and int(np.sum(fields[name])) == summary[name]
