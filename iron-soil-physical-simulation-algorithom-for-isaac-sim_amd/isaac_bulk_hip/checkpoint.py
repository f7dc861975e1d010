"""Explicit device checkpoint boundaries (the 390F production checkpoint format).

Mirrors pkg/runtime/device_checkpoint.py:21-212: the same ``.npz`` layout
(one float64/int32 array per named field plus ``checkpoint_record_json``), the
schema ``390F_PRODUCTION_DEVICE_CHECKPOINT/v2``, the same record keys and
SHA-256 definitions, and the same restore checks and legacy migrations, so
files written by either side load on the other.  The reference's
``EarthmovingPhysicsCore`` is outside this package: the caller passes the
DeviceBulkState and, when it has them, the LargeAvalanche bridge, the residual
frontier and the runtime-metadata record.

Full-field traffic is packed (DeviceBulkState.download_fields and
restore_checkpoint_fields): every float field is widened to float64 and
concatenated on the device, then copied to host memory in one transfer (one
more for the int32 fields); a restore uploads one packed buffer per dtype and
splits it on the device.  FP32 states are written in the file's float64
metres (mobile quanta times 2^-24, exact) and read back exactly.

``load_frozen_boundary`` reads the frozen-boundary keys ``PRE_resting``,
``PRE_mobile``, ``PRE_momentum_x/y`` (tools/run_mobile_v2_frozen_5s.py:99-106)
with the reference's explicit compatibility adapter b_eff = (resting + h) - h.
"""

from __future__ import annotations

import hashlib
import json
from pathlib import Path
from typing import Any, Mapping

import numpy as np

from .bulk_state import DeviceBulkState

SCHEMA = "390F_PRODUCTION_DEVICE_CHECKPOINT/v2"
FLOAT_FIELDS = (
    "z_base", "b_eff", "mobile", "momentum_x", "momentum_y", "track_rut",
    "avalanche_first_activation_time", "avalanche_last_activation_time",
    "avalanche_owned_surface", "avalanche_owned_export_baseline", "mobile_export_cumulative",
    "mobile_flux_export_cumulative", "avalanche_r2m_cumulative", "avalanche_m2r_cumulative",
)
INT_FIELDS = (
    "active_mask", "material_mask", "frontier_reached", "avalanche_latch",
    "avalanche_previous_component", "avalanche_activation_count", "deposition_mask",
)


def sha256_file(path: str | Path) -> str:
    digest = hashlib.sha256()
    with Path(path).open("rb") as stream:
        for block in iter(lambda: stream.read(1024 * 1024), b""):
            digest.update(block)
    return digest.hexdigest()


def sha256_array(value: np.ndarray) -> str:
    """Hash of dtype name, int64 shape and C-order bytes (device_checkpoint.py:56-62)."""
    array = np.ascontiguousarray(np.asarray(value))
    digest = hashlib.sha256()
    digest.update(str(array.dtype).encode("ascii"))
    digest.update(np.asarray(array.shape, dtype=np.int64).tobytes())
    digest.update(array.tobytes())
    return digest.hexdigest()


def write_device_checkpoint(state: DeviceBulkState, path: str | Path, *, provenance: Mapping[str, Any],
                            large_avalanche=None, residual_frontier=None,
                            metadata_record: Mapping[str, Any] | None = None,
                            static_relaxation_pending: bool = False, compress: bool = True) -> dict[str, Any]:
    """write_device_checkpoint (device_checkpoint.py:65-124) for this package's state."""
    arrays = state.download_fields((*FLOAT_FIELDS, *INT_FIELDS))
    record = {
        "schema": SCHEMA,
        "terrain_state_semantics": {
            "authoritative": ["z_base", "b_eff", "mobile", "momentum_x", "momentum_y"],
            "H_free": "b_eff + mobile",
            "h_resting": "b_eff - z_base",
        },
        "runtime_backend": "GPU_RUNTIME",
        "state_authority": state.authority.value,
        "shape_yx": list(state.shape),
        "resolution_xy_m": [state.grid.dx, state.grid.dy],
        "device_timestamp_s": state.timestamp_device_s,
        "metadata": dict(metadata_record or {}),
        "large_avalanche": {
            "persistence_s": 0.0 if large_avalanche is None else large_avalanche._persistence_s,
            "has_previous": False if large_avalanche is None else large_avalanche._has_previous,
        },
        "minislope": {
            "active_tiles": [] if residual_frontier is None else residual_frontier.active_tile_ids.tolist(),
            "iterations": 0 if residual_frontier is None else int(residual_frontier.iterations),
            "pending": bool(static_relaxation_pending),
        },
        "dirty_tiles": sorted(state._dirty_tile_ids),
        "terrain_hash_sha256": sha256_array(arrays["b_eff"]),
        "provenance": dict(provenance),
    }
    target = Path(path)
    target.parent.mkdir(parents=True, exist_ok=True)
    save = np.savez_compressed if compress else np.savez
    save(target, **arrays,
         checkpoint_record_json=np.asarray(json.dumps(record, sort_keys=True, separators=(",", ":"))))
    if target.suffix != ".npz":
        target = target.with_name(target.name + ".npz")
    result = dict(record)
    result["checkpoint_path"] = str(target)
    result["checkpoint_hash_sha256"] = sha256_file(target)
    return result


def checkpoint_fields(archive) -> tuple[dict[str, np.ndarray], str]:
    """The field set of a checkpoint with the reference's legacy migrations
    (device_checkpoint.py:146-180): returns (fields, migration label)."""
    fields = {name: np.asarray(archive[name]).copy() for name in (*FLOAT_FIELDS, *INT_FIELDS)
              if name in archive}
    migration = "NONE_V2_NATIVE"
    if "b_eff" not in fields:
        if "resting" not in archive:
            raise RuntimeError("DEVICE_CHECKPOINT_MISSING_STATIC_BED")
        fields["b_eff"] = np.asarray(archive["resting"]).copy()
        fields["z_base"] = np.zeros_like(fields["b_eff"])
        migration = "UNCALIBRATED_SNAPSHOT_COMPATIBILITY_ADAPTER:B_EFF=H_FREE_OLD-H_MOBILE_OLD;Z_BASE=0"
    elif "z_base" not in fields:
        fields["z_base"] = np.zeros_like(fields["b_eff"])
        migration = "UNCALIBRATED_SNAPSHOT_COMPATIBILITY_ADAPTER:B_EFF_DIRECT;IMPLICIT_Z_BASE=0"
    if "avalanche_owned_surface" not in fields:
        fields["avalanche_owned_surface"] = fields["b_eff"] + fields["mobile"]
    for name in ("avalanche_owned_export_baseline", "mobile_export_cumulative"):
        if name not in fields:
            fields[name] = np.zeros_like(fields["b_eff"])
    return fields, migration


def restore_device_checkpoint(state: DeviceBulkState, path: str | Path, *, large_avalanche=None,
                              residual_frontier=None) -> dict[str, Any]:
    """restore_device_checkpoint (device_checkpoint.py:127-212) into this package's state."""
    source = Path(path)
    with np.load(source, allow_pickle=False) as archive:
        record = json.loads(str(archive["checkpoint_record_json"]))
        if tuple(record["shape_yx"]) != state.shape:
            raise RuntimeError("DEVICE_CHECKPOINT_SHAPE_MISMATCH")
        if list(record["resolution_xy_m"]) != [state.grid.dx, state.grid.dy]:
            raise RuntimeError("DEVICE_CHECKPOINT_RESOLUTION_MISMATCH")
        fields, migration = checkpoint_fields(archive)
    state.restore_checkpoint_fields(fields, timestamp_s=float(record["device_timestamp_s"]),
                                    source="checkpoint")
    if large_avalanche is not None:
        large_avalanche._persistence_s = float(record["large_avalanche"]["persistence_s"])
        large_avalanche._has_previous = bool(record["large_avalanche"]["has_previous"])
    if residual_frontier is not None:
        residual_frontier._flags.zero_()
        residual_frontier.add_tiles(record["minislope"]["active_tiles"])
        residual_frontier.iterations = int(record["minislope"]["iterations"])
    state._dirty_tile_ids.clear()
    state.mark_dirty_tiles(record.get("dirty_tiles", []))
    if sha256_array(fields["b_eff"]) != record["terrain_hash_sha256"]:
        raise RuntimeError("DEVICE_CHECKPOINT_TERRAIN_HASH_MISMATCH")
    result = dict(record)
    result["legacy_checkpoint_migration"] = migration
    result["checkpoint_path"] = str(source)
    result["checkpoint_hash_sha256"] = sha256_file(source)
    return result


def load_frozen_boundary(path: str | Path):
    """Frozen-boundary snapshot (tools/run_mobile_v2_frozen_5s.py:99-106):
    returns (b_eff, h, q[..., 2], H_free) with b_eff = H_free - h."""
    with np.load(path, allow_pickle=False) as data:
        old_resting = np.asarray(data["PRE_resting"], dtype=np.float64)
        h = np.asarray(data["PRE_mobile"], dtype=np.float64)
        q = np.stack((data["PRE_momentum_x"], data["PRE_momentum_y"]), axis=-1).astype(np.float64)
    free = old_resting + h
    return free - h, h, q, free
