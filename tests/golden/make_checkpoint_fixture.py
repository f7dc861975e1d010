"""Write a device checkpoint fixture in the reference's file layout with
plain NumPy (no import of the package under test), so the restore path is
checked against a file it did not write.

Layout restated from the reference writer (pkg/runtime/device_checkpoint.py:21-45,
:86-124): np.savez_compressed with the 14 float64 and 7 int32 named fields
of shape (ny, nx) and `checkpoint_record_json` (sorted keys, compact
separators; schema 390F_PRODUCTION_DEVICE_CHECKPOINT/v2; terrain hash =
SHA-256 over the dtype name, the int64 shape and the bytes of b_eff).

    python tests/golden/make_checkpoint_fixture.py
"""

from __future__ import annotations

import hashlib
import json
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
FLOAT_FIELDS = ("z_base", "b_eff", "mobile", "momentum_x", "momentum_y", "track_rut",
                "avalanche_first_activation_time", "avalanche_last_activation_time",
                "avalanche_owned_surface", "avalanche_owned_export_baseline", "mobile_export_cumulative",
                "mobile_flux_export_cumulative", "avalanche_r2m_cumulative", "avalanche_m2r_cumulative")
INT_FIELDS = ("active_mask", "material_mask", "frontier_reached", "avalanche_latch",
              "avalanche_previous_component", "avalanche_activation_count", "deposition_mask")


def terrain_hash(a: np.ndarray) -> str:
    a = np.ascontiguousarray(a)
    d = hashlib.sha256()
    d.update(str(a.dtype).encode("ascii"))
    d.update(np.asarray(a.shape, dtype=np.int64).tobytes())
    d.update(a.tobytes())
    return d.hexdigest()


def main() -> None:
    ny, nx, dx, dy = 13, 17, 0.05, 0.04
    rng = np.random.default_rng(2024)
    f = {}
    f["z_base"] = rng.uniform(0.0, 0.4, (ny, nx))
    f["b_eff"] = f["z_base"] + rng.uniform(0.2, 1.5, (ny, nx))
    # mobile depth on whole quanta of 2^-24 m, so an FP32 restore is exact too
    f["mobile"] = np.rint(rng.uniform(0.0, 0.2, (ny, nx)) * 2.0 ** 24) * 2.0 ** -24
    f["momentum_x"] = rng.normal(0.0, 0.05, (ny, nx))
    f["momentum_y"] = rng.normal(0.0, 0.05, (ny, nx))
    for name in FLOAT_FIELDS[5:]:
        f[name] = rng.uniform(0.0, 2.0, (ny, nx))
    for name in INT_FIELDS:
        f[name] = rng.integers(0, 3, (ny, nx)).astype(np.int32)
    record = {
        "schema": "390F_PRODUCTION_DEVICE_CHECKPOINT/v2",
        "terrain_state_semantics": {"authoritative": ["z_base", "b_eff", "mobile", "momentum_x", "momentum_y"],
                                    "H_free": "b_eff + mobile", "h_resting": "b_eff - z_base"},
        "runtime_backend": "GPU_RUNTIME",
        "state_authority": "DEVICE",
        "shape_yx": [ny, nx],
        "resolution_xy_m": [dx, dy],
        "device_timestamp_s": 3.5,
        "metadata": {"payload_m3": 0.125},
        "large_avalanche": {"persistence_s": 0.25, "has_previous": True},
        "minislope": {"active_tiles": [0, 3], "iterations": 7, "pending": False},
        "dirty_tiles": [1, 2],
        "terrain_hash_sha256": terrain_hash(f["b_eff"]),
        "provenance": {"writer": "tests/golden/make_checkpoint_fixture.py (NumPy only)"},
    }
    np.savez_compressed(HERE / "device_checkpoint_13x17.npz", **f,
                        checkpoint_record_json=np.asarray(json.dumps(record, sort_keys=True, separators=(",", ":"))))
    # independent summary of the fixture's contents for the CPU test
    summary = {name: float(np.sum(f[name], dtype=np.float64)) for name in FLOAT_FIELDS}
    summary.update({name: int(np.sum(f[name])) for name in INT_FIELDS})
    (HERE / "device_checkpoint_13x17.json").write_text(json.dumps(summary, indent=1, sort_keys=True))


if __name__ == "__main__":
    main()
