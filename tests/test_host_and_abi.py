"""CPU-side checks: host logic of the plugin surface and the C-ABI library exports."""

from __future__ import annotations

import ctypes
import re
from pathlib import Path

import numpy as np
import pytest

from isaac_bulk_hip import _lib
from isaac_bulk_hip.config import (
    LargeAvalancheTransitionConfig,
    MaterialScenario,
    MobileV2Config,
    TerrainGrid,
    TerrainVolumeIntegrator,
    triangle_vertex_weights,
)
from isaac_bulk_hip.relaxation import CompactTileFrontier, legacy_critical_differences

ROOT = Path(__file__).resolve().parents[1]


def declared_symbols() -> list[str]:
    text = (ROOT / "include" / "mv2.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(mv2_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    declared = declared_symbols()
    assert len(declared) >= 20
    missing = [name for name in declared if not hasattr(lib, name)]
    assert not missing, missing
    assert set(declared) <= set(_lib.SIGNATURES)


def test_abi_constants():
    lib = _lib.load()
    assert lib.mv2_abi_version() == 1
    assert lib.mv2_mobile_quantum() == 2.0 ** -24


STRUCTS = {
    "StepConfig": "mv2_step_config",
    "StepRecord": "mv2_step_record",
    "RelaxResult": "mv2_relax_result",
    "AvalancheConfig": "mv2_avalanche_config",
    "AvalancheFields": "mv2_avalanche_fields",
    "LabelResult": "mv2_label_result",
    "BatchRecord": "mv2_batch_record",
}


def test_struct_layouts_match_header(tmp_path):
    """Every ctypes record has the C compiler's size and field offsets for include/mv2.h."""
    import shutil
    import subprocess

    if shutil.which("gcc") is None:
        pytest.skip("gcc not available")
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "mv2.h"', "int main(void) {"]
    expected = []
    for py_name, c_name in STRUCTS.items():
        cls = getattr(_lib, py_name)
        lines.append(f'printf("%zu\\n", sizeof({c_name}));')
        expected.append(ctypes.sizeof(cls))
        for field, _ in cls._fields_:
            lines.append(f'printf("%zu\\n", offsetof({c_name}, {field}));')
            expected.append(getattr(cls, field).offset)
    lines += ["return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", str(ROOT / "include"), str(src), "-o", str(exe)], check=True)
    got = [int(v) for v in subprocess.run([str(exe)], check=True, capture_output=True,
                                          text=True).stdout.split()]
    assert got == expected


def test_create_without_gpu_reports_backend_unavailable():
    torch = pytest.importorskip("torch")
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    lib = _lib.load()
    handle = ctypes.c_void_p()
    rc = lib.mv2_create(0, ctypes.byref(handle))
    assert rc == _lib.MV2_EDEVICE
    assert "unavailable" in _lib.last_error(None) or "HIP" in _lib.last_error(None)
    from isaac_bulk_hip import HipBackendUnavailable, HipRuntime

    with pytest.raises(HipBackendUnavailable):
        HipRuntime("cuda:0")


def test_mobile_v2_config_guards():
    with pytest.raises(ValueError, match="CFL must be <= 0.25"):
        MobileV2Config(cfl=0.4)
    assert MobileV2Config(cfl=0.4, allow_cfl_above_quarter=True).cfl == 0.4
    with pytest.raises(ValueError, match="finite/positive"):
        MobileV2Config(dx_m=0.0)


def test_large_avalanche_config_mapping_and_sensitivity():
    with pytest.raises(ValueError, match="unknown configuration fields"):
        LargeAvalancheTransitionConfig.from_mapping({"hidden_magic_threshold": 3})
    nominal = LargeAvalancheTransitionConfig.from_mapping({"persistence_time_s": 0.1})
    cases = nominal.sensitivity_ensemble(0.20)
    assert [c.sensitivity_case for c in cases] == [
        "more_mobile_uncalibrated", "nominal_uncalibrated", "less_mobile_uncalibrated"]
    assert cases[0].persistence_time_s < nominal.persistence_time_s < cases[2].persistence_time_s
    with pytest.raises(ValueError):
        LargeAvalancheTransitionConfig(connectivity=6)


def test_material_scenario_validation():
    m = MaterialScenario("iron_ore_fines_frozen", 1370.0, 29.8, 800.0, 0.5, 38.0, 30.0, 0.35)
    assert m.assumed_bulk_density_kg_m3 == 1370.0
    with pytest.raises(ValueError, match="hysteresis"):
        MaterialScenario("x", 1370.0, 29.8, 800.0, 0.5, 30.0, 38.0, 0.35)


def test_triangle_weights_integrate_domain_area():
    w = triangle_vertex_weights(7, 9, 0.05, 0.05)
    assert abs(w.sum() - 6 * 8 * 0.05 * 0.05) < 1e-15
    integ = TerrainVolumeIntegrator.from_grid(TerrainGrid(9, 7, 0.05, 0.05))
    np.testing.assert_array_equal(integ.vertex_weights_m2, w)


def test_edge_batch_counts_match_reference_test():
    """tests/test_sparse_tile_slope.py:34-63 (17 x 19 grid, one tile covering it)."""
    frontier = CompactTileFrontier((17, 19), 32)
    counts = []
    for d, (di, dj) in enumerate(((1, 0), (0, 1), (1, 1), (1, -1))):
        total = 0
        for phase in (0, 1):
            total += frontier.edge_batch([0], direction_index=d, di=di, dj=dj, phase=phase).edge_count
        counts.append(total)
    assert counts == [16 * 19, 17 * 18, 16 * 18, 16 * 18]


def test_tiles_around_cells_stencil():
    frontier = CompactTileFrontier((70, 70), 32)
    tiles = frontier.tiles_around_cells(np.asarray([31]), np.asarray([31]))
    assert tiles.tolist() == [0, 1, 3, 4]


def test_legacy_critical_differences():
    crit = legacy_critical_differences(0.05, 0.04, 30.0)
    s = float(np.tan(np.deg2rad(30.0)))
    assert crit.tolist() == [s * 0.05, s * 0.04, s * float(np.hypot(0.05, 0.04)),
                             s * float(np.hypot(0.05, 0.04))]


def test_workload_generators_are_seeded():
    from isaac_bulk_hip.workloads import cut_face_terrain, pile_with_patch

    a = pile_with_patch(128, seed=3)
    b = pile_with_patch(128, seed=3)
    for x, y in zip(a, b):
        np.testing.assert_array_equal(x, y)
    bb, hh, face = cut_face_terrain(101, seed=1)
    assert bb.min() >= 0.0 and hh.min() >= 0.0 and face.any()
    assert np.all(hh[~face] == 0.0)


def test_committed_warm_start_weights_load():
    from isaac_bulk_hip.warm_start_training import DEFAULT_WEIGHTS, load_weights

    params, meta = load_weights(DEFAULT_WEIGHTS)
    assert params["w1"].shape == (9, 32, 3) and params["w5"].shape == (9, 1, 32)
    for k in (2, 3, 4):
        assert params[f"w{k}"].shape == (9, 32, 32)
    assert meta["critical_deg"] == 34.0 and 0.0 <= meta["margin_deg"] <= 1.0
    # zero-sum taps on the b_eff channel (constant-shift invariance)
    assert np.abs(params["w1"][:, :, 0].sum(axis=0)).max() < 1e-5
    assert all(np.isfinite(v).all() for v in params.values())


def test_batched_tile_shapes_match_library():
    from isaac_bulk_hip.batched import TILE_SHAPES

    lib = _lib.load()
    for shape, (th, tw) in TILE_SHAPES.items():
        a, b = ctypes.c_int(), ctypes.c_int()
        assert lib.mv2_batch_tile_shape(shape, ctypes.byref(a), ctypes.byref(b)) == 0
        assert (a.value, b.value) == (th, tw)
    assert lib.mv2_batch_tile_shape(len(TILE_SHAPES), ctypes.byref(a), ctypes.byref(b)) != 0


def test_batch_records_view_matches_ctypes_layout():
    from isaac_bulk_hip.batched import _RECORD_DTYPE, BatchRecords

    raw = (_lib.BatchRecord * 2)()
    raw[1].substeps = 7
    raw[1].static_exit = 1
    raw[1].mass_quanta_after = 12345.0
    raw[1].momentum_after[1] = -2.5
    raw[1].kinetic_before = 0.25
    table = np.frombuffer(bytes(raw), dtype=_RECORD_DTYPE)
    recs = BatchRecords(table)
    assert len(recs) == 2 and recs.substeps.tolist() == [0, 7]
    r = recs[1]
    assert r.static_exit and r.mass_quanta_after == 12345 and r.momentum_after == (0.0, -2.5)
    assert r.kinetic_before == 0.25 and [x.substeps for x in recs] == [0, 7]
