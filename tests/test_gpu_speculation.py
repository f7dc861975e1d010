"""Speculative FP32 steps against steps with a separate begin launch.

A step whose predecessor left a CFL prediction runs the begin pass and
substep 0 in one launch with the predicted maximum; the next launch keeps
that substep only when the true maximum of the input equals the prediction
and the input is not static.  Every case here runs the same frames on two
states, one with speculation and one without, and requires identical fields
and records bit for bit: steady frames with relaxation between them (the
prediction holds), a write to the mobile layer that bypasses the library
(the prediction fails and the step restarts from its input), a library
entrainment (no speculation on the next step) and an input made static
behind the library's back (the speculative substep is discarded)."""

from __future__ import annotations

import ctypes
import dataclasses

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

N = 130
DT = 1.0 / 60.0


def _runs():
    from isaac_bulk_hip import (DeviceBulkState, MaterialScenario, MobileV2Config, TerrainGrid,
                                TerrainVolumeIntegrator, WarpProductionMobileV2Solver)
    from isaac_bulk_hip.workloads import cut_face_terrain

    b, h, _ = cut_face_terrain(N, 0.05, seed=5)
    grid = TerrainGrid(N, N, 0.05, 0.05)
    material = MaterialScenario("iron_ore_fines_uncalibrated", 1370.0, 29.8, 0.0, 0.5, 34.0, 30.0, 0.55)
    runs = []
    for spec in (True, False):
        state = DeviceBulkState(grid, b, precision="fp32")
        state.set_fields(mobile=h)
        solver = WarpProductionMobileV2Solver(runtime=state.runtime, config=MobileV2Config())
        solver.bind_device_state(state, material, grid, TerrainVolumeIntegrator.from_grid(grid))
        rt = state.runtime
        rt.check(rt.lib.mv2_set_speculation(rt.ctx, int(spec)))
        runs.append((state, solver))
    return runs


def _stats(state):
    rt = state.runtime
    out = (ctypes.c_longlong * 2)()
    rt.check(rt.lib.mv2_speculation_stats(rt.ctx, out))
    return int(out[0]), int(out[1])


def test_solver_switch_and_counters():
    runs = _runs()
    for _ in range(3):
        _frame(runs)
    assert runs[0][1].speculation_stats() == _stats(runs[0][0]) == (2, 0)
    runs[0][1].set_speculation(False)
    _frame(runs)
    assert runs[0][1].speculation_stats() == (2, 0)
    runs[0][1].set_speculation(True)
    _frame(runs)                                 # the step after a plain one left a prediction again
    assert runs[0][1].speculation_stats() == (3, 0)


def _frame(runs, relax=True):
    """One frame on both states; fields and records must agree bit for bit."""
    import torch
    from isaac_bulk_hip import relax_device

    recs = []
    for state, solver in runs:
        recs.append(solver.step_resident(DT))
        if relax:
            b = state.runtime.arrays["b_eff"].view(N, N)
            relax_device(b, 0.05, 0.05, 34.0, 1000, 1e-6, runtime=state.runtime)
    torch.cuda.synchronize()
    a, b = (dataclasses.asdict(r) for r in recs)
    for key in a:
        va, vb = np.asarray(a[key]), np.asarray(b[key])
        assert np.array_equal(va, vb), key
    raw = [np.asarray(s.last_record.diag[:]) for _, s in runs]
    assert np.array_equal(raw[0], raw[1])
    fields = [state.runtime.arrays for state, _ in runs]
    for name in ("mobile", "momentum_x", "momentum_y", "b_eff"):
        assert torch.equal(fields[0][name], fields[1][name]), name
    return recs[0]


def test_speculative_frames_match_plain_frames():
    runs = _runs()
    substeps = [_frame(runs).substeps for _ in range(10)]
    assert min(substeps) > 0
    spec, restarts = _stats(runs[0][0])
    assert spec == 9 and restarts == 0          # every step after the first speculated, none restarted
    assert _stats(runs[1][0]) == (0, 0)


def test_write_outside_the_library_restarts_the_step():
    import torch

    runs = _runs()
    for _ in range(3):
        _frame(runs)
    # add mobile material to a patch by writing the tensor directly: the
    # library does not see it, so the next step speculates on a stale maximum
    for state, _ in runs:
        m = state.runtime.arrays["mobile"].view(N, N)
        m[60:70, 60:70] += int(0.4 * 2 ** 24)
    torch.cuda.synchronize()
    before = _stats(runs[0][0])
    _frame(runs)
    after = _stats(runs[0][0])
    assert after[0] == before[0] + 1 and after[1] == before[1] + 1
    for _ in range(3):
        _frame(runs)
    assert _stats(runs[0][0])[1] == after[1]    # the following steps' predictions hold again


def test_library_entrainment_skips_speculation():
    runs = _runs()
    for _ in range(3):
        _frame(runs)
    idx = np.arange(40 * N + 40, 40 * N + 60, dtype=np.int32)
    for state, _ in runs:
        state.entrain_host_indices(idx, np.full(idx.shape, 0.01), reason="speculation test")
    before = _stats(runs[0][0])
    _frame(runs)
    assert _stats(runs[0][0]) == before          # no prediction after the library wrote mobile
    _frame(runs)
    assert _stats(runs[0][0])[0] == before[0] + 1


def test_static_input_after_a_speculative_substep():
    import torch

    runs = _runs()
    for _ in range(3):
        _frame(runs)
    # a flat, resting state written behind the library's back: the next step
    # speculates, and its second launch finds the input static
    for state, _ in runs:
        arr = state.runtime.arrays
        arr["b_eff"].fill_(1.0)
        arr["mobile"].fill_(int(0.05 * 2 ** 24))
        arr["momentum_x"].zero_()
        arr["momentum_y"].zero_()
    torch.cuda.synchronize()
    before = _stats(runs[0][0])
    rec = _frame(runs, relax=False)
    assert rec.substeps == 0
    assert _stats(runs[0][0])[0] == before[0] + 1
    rec = _frame(runs, relax=False)              # no prediction after a static exit: a plain step
    assert rec.substeps == 0 and _stats(runs[0][0])[0] == before[0] + 1
