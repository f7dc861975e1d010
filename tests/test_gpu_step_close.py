"""The FP32 step's closing ledger, stored by its last substep.

The last substep of an FP32 step forms the ledger after the step from the new
state in its registers.  It also forms the CFL maximum that the next step's
first launch uses as its prediction. The fold pass copies the state back when
the substep count is odd.  Without relaxation between frames, a step's closing
ledger and the next step's opening ledger (the begin pass over the same state)
must be the same bits.  That pins the fused ledger to the begin pass, which the
oracle tests pin.  A failure in the last substep rolls the step back to its
input."""

from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

DT = 1.0 / 60.0


def _production(n, b, h, speculation=True, qx=None):
    from isaac_bulk_hip import (DeviceBulkState, MaterialScenario, MobileV2Config, TerrainGrid,
                                TerrainVolumeIntegrator, WarpProductionMobileV2Solver)

    grid = TerrainGrid(n, n, 0.05, 0.05)
    state = DeviceBulkState(grid, b, precision="fp32")
    state.set_fields(mobile=h, momentum_x=qx)
    material = MaterialScenario("iron_ore_fines_uncalibrated", 1370.0, 29.8, 0.0, 0.5, 34.0, 30.0, 0.55)
    solver = WarpProductionMobileV2Solver(runtime=state.runtime, config=MobileV2Config())
    solver.bind_device_state(state, material, grid, TerrainVolumeIntegrator.from_grid(grid))
    solver.set_speculation(speculation)
    return state, solver


@pytest.mark.parametrize("speculation", [True, False])
def test_closing_ledger_equals_next_opening_ledger(speculation):
    from isaac_bulk_hip.workloads import cut_face_terrain

    n = 150
    b, h, _ = cut_face_terrain(n, 0.05, seed=3)
    state, solver = _production(n, b, h, speculation)
    raw, substeps = [], []
    # frame lengths that give 3, 2 and 1 substeps: a step of one substep
    # closes in the launch after it (a speculative substep cannot close)
    for f in range(15):
        rec = solver.step_resident(DT * (1.0, 0.6, 0.3)[f % 3])
        substeps.append(rec.substeps)
        r = solver.last_record
        raw.append((np.array(r.summary_before[:]), np.array(r.summary_after[:]),
                    float(r.mass_quanta_before), float(r.mass_quanta_after)))
    # both parities of the substep count occur (even: the last substep wrote
    # buffer 0; odd: the fold copied buffer 1 back)
    assert {s % 2 for s in substeps if s > 0} == {0, 1} and 1 in substeps, substeps
    for f in range(len(raw) - 1):
        np.testing.assert_array_equal(raw[f][1], raw[f + 1][0], err_msg=f"frame {f}")
        assert raw[f][3] == raw[f + 1][2], f
        assert raw[f][2] == raw[f][3], f          # the closed domain keeps every quantum
    if speculation:
        spec, restarts = solver.speculation_stats()
        assert spec == len(raw) - 1 and restarts == 0


def test_fp32_range_failure_in_the_last_substep_rolls_back():
    """A step of one substep whose substep overflows the FP32 thickness range
    (a pit surrounded by 127.9 m of material): the step raises, and the fields
    are the step's input.  The previous step also took one substep, so the
    first round holds exactly that launch, and the failure is seen by the fold
    pass; the launch of the second round rolls back.  A normal step follows."""
    n, dt = 16, 1.0e-4
    b = np.zeros((n, n))
    h = np.ones((n, n))
    state, solver = _production(n, b, h, speculation=False, qx=np.full((n, n), 0.1))
    assert solver.step_resident(dt).substeps == 1
    b2 = np.zeros((n, n))
    b2[8, 8] = -20.0
    h2 = np.full((n, n), 127.9)
    state.set_fields(b_eff=b2, mobile=h2, momentum_x=np.zeros((n, n)), momentum_y=np.zeros((n, n)))
    before = state.explicit_host_view(source="acceptance")
    h_in = before.H_mobile_m.copy()
    with pytest.raises(RuntimeError, match="storage range"):
        solver.step_resident(dt)
    after = state.explicit_host_view(source="acceptance")
    np.testing.assert_array_equal(after.H_mobile_m, h_in)
    assert not np.any(after.mobile_momentum_m2_s)
    # the library recovers: a safe state steps normally
    state.set_fields(b_eff=b, mobile=h, momentum_x=np.full((n, n), 0.1), momentum_y=np.zeros((n, n)))
    rec = solver.step_resident(dt)
    assert rec.substeps == 1 and rec.volume_after_m3 == pytest.approx(rec.volume_before_m3, rel=1e-12)
