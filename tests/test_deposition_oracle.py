"""The oracle's deposition step (oracle/mv2_oracle.c orc_deposition*) pinned
against the reference's own test case, an independent NumPy statement of the
decision rule, and the transfer invariants."""

from __future__ import annotations

import numpy as np
import pytest

import deposition_cases as dc
import oracle

TAN_STOP = float(np.tan(np.deg2rad(dc.STOP_DEG)))


def _run(b, h, qx, qy, cum, *, dx, dy, dt, settle=False, forcing=None, fp32=False, **kw):
    w = oracle.vertex_weights(*b.shape, dx, dy)
    kw.setdefault("cohesion", dc.COHESION)
    return oracle.deposition(b, h, qx, qy, cum, w, dx=dx, dy=dy, dt=dt, density=dc.DENSITY,
                             gravity=dc.GRAVITY, tan_stop=TAN_STOP, settle_subcell_tail=settle,
                             forcing=forcing, fp32=fp32, **kw), w


def test_reference_subcell_tail_case():
    """tests/test_device_bulk_state.py:195-225: a 0.01 m mobile cell beside a
    0.20 m resting cell settles completely; the deposited volume is the
    integrated mobile volume and the resting surface gains exactly it."""
    n, dx = 16, 0.05
    resting = np.zeros((n, n))
    resting[8, 8] = 0.20
    mobile = np.zeros((n, n))
    mobile[8, 9] = 0.01
    qx = np.zeros((n, n))
    qx[8, 9] = 0.02
    (b, m, qx1, qy1, dep, rx, ry, cum, out), w = _run(
        resting, mobile, qx, np.zeros((n, n)), np.zeros((n, n)), dx=dx, dy=dx, dt=1.0 / 60.0,
        settle=True)
    assert out[0] == pytest.approx(float(np.sum(mobile * w)), abs=1e-13)
    np.testing.assert_allclose(b, resting + mobile, atol=1e-13)
    np.testing.assert_allclose(m, 0.0, atol=1e-13)
    assert qx1[8, 9] == 0.0 and rx[8, 9] == 0.02
    assert out[1] == pytest.approx(dc.DENSITY * 0.02 * w[8, 9], rel=1e-15)
    assert cum[8, 9] == 0.01 * w[8, 9]


@pytest.mark.parametrize("cohesion", [0.0, 50.0, 800.0])
def test_decisions_match_numpy_statement(cohesion):
    """Anisotropic spacing (dy != dx) catches swapped axes; the edges exercise
    the one-sided differences; cohesion 0/50/800 Pa moves the yield line."""
    dx, dy, dt = 0.05, 0.07, 1.0 / 60.0
    b, h, qx, qy, forcing, cum0 = dc.fields(37, 53, 11, dx=dx, dy=dy)
    (b1, m1, qx1, qy1, dep, rx, ry, cum, out), w = _run(
        b, h, qx, qy, cum0, dx=dx, dy=dy, dt=dt, forcing=forcing, cohesion=cohesion)
    amount, knife = dc.numpy_rule(b, h, qx, qy, forcing, dx=dx, dy=dy, dt=dt, cohesion=cohesion)
    ok = ~knife
    assert knife.sum() <= 2
    np.testing.assert_array_equal(dep[ok], amount[ok])
    decided = amount > 0.0
    assert 0 < decided.sum() < (h > 0).sum()          # both outcomes occur
    np.testing.assert_array_equal(b1[ok], (b + amount)[ok])
    np.testing.assert_array_equal(m1[ok], (h - amount)[ok])
    np.testing.assert_array_equal(dep[forcing != 0], 0.0)


def test_transfer_invariants():
    """deposition.py:144-159: free surface unchanged per cell, removed plus
    remaining momentum equals the old momentum, the exact transfer balance,
    and the cumulative field grows by the weighted amount."""
    dx = dy = 0.05
    b, h, qx, qy, forcing, cum0 = dc.fields(41, 29, 5)
    (b1, m1, qx1, qy1, dep, rx, ry, cum, out), w = _run(
        b, h, qx, qy, cum0, dx=dx, dy=dy, dt=1.0 / 30.0, forcing=forcing, cohesion=50.0)
    np.testing.assert_allclose(b1 + m1, b + h, rtol=0, atol=4e-16)
    np.testing.assert_allclose(qx1 + rx, qx, rtol=0, atol=1e-18)
    np.testing.assert_allclose(qy1 + ry, qy, rtol=0, atol=1e-18)
    before, after = float(np.sum(h * w)), float(np.sum(m1 * w))
    assert abs(before - after - out[0]) <= 1e-15
    assert out[1] == pytest.approx(float(np.sum(dc.DENSITY * rx * w)), rel=1e-12)
    np.testing.assert_allclose(cum - cum0, dep * w, rtol=1e-12, atol=1e-20)
    assert np.all(m1 >= 0.0) and np.all(dep <= np.minimum(h, 0.30 / 30.0))


def test_settle_mode_takes_every_unforced_cell():
    b, h, qx, qy, forcing, cum0 = dc.fields(19, 23, 3)
    (b1, m1, qx1, qy1, dep, rx, ry, cum, out), w = _run(
        b, h, qx, qy, cum0, dx=0.05, dy=0.05, dt=1e-3, settle=True, forcing=forcing)
    free = forcing == 0
    np.testing.assert_array_equal(dep[free], h[free])
    np.testing.assert_array_equal(m1[free], 0.0)
    np.testing.assert_array_equal(qx1[free], 0.0)
    np.testing.assert_array_equal(rx[free], qx[free])
    np.testing.assert_array_equal(m1[~free], h[~free])
    np.testing.assert_array_equal(qx1[~free], qx[~free])


def test_fp32_storage_matches_fp64_decisions_in_whole_quanta():
    """FP32 storage: the same decisions as the FP64 step on the widened
    fields, amounts min(h, floor(rate dt / q)) in whole quanta."""
    q = oracle_quantum = 1.0 / 16777216.0
    dx = dy = 0.05
    dt = 1.0 / 60.0
    b, h, qx, qy, forcing, cum0 = dc.fields(31, 45, 8)
    b32 = b.astype(np.float32)
    hq = np.rint(h / oracle_quantum).astype(np.int32)
    qx32, qy32 = qx.astype(np.float32), qy.astype(np.float32)
    (b1, m1, qx1, qy1, dep, rx, ry, cum, out), w = _run(
        b32, hq, qx32, qy32, cum0, dx=dx, dy=dy, dt=dt, forcing=forcing, cohesion=50.0, fp32=True)
    (_, _, _, _, dep64, _, _, _, _), _ = _run(
        b32.astype(np.float64), hq * q, qx32.astype(np.float64), qy32.astype(np.float64), cum0,
        dx=dx, dy=dy, dt=dt, forcing=forcing, cohesion=50.0)
    cap = int(np.floor(0.30 * dt / q))
    np.testing.assert_array_equal(dep > 0, dep64 > 0)
    np.testing.assert_array_equal(dep[dep > 0], np.minimum(hq, cap)[dep > 0])
    np.testing.assert_array_equal(m1, hq - dep)
    np.testing.assert_array_equal(b1, (b32.astype(np.float64) + dep * q).astype(np.float32))
    assert out[0] == pytest.approx(float(np.sum(dep * q * w)), rel=1e-13)
