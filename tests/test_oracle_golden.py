"""Pin the CPU oracle against the reference's own evidence and closed forms.

Golden data: tests/golden/mobile_v2_cpu_contracts.json (copied from the
reference's committed outputs by tests/golden/make_contract_fixture.py).
"""

from __future__ import annotations

import json
from pathlib import Path

import numpy as np
import pytest
import scipy.ndimage

import contract_cases
import oracle

GOLD = json.loads((Path(__file__).parent / "golden" / "mobile_v2_cpu_contracts.json").read_text())


@pytest.mark.parametrize("name", sorted(contract_cases.cases()))
def test_contract_case_scalars_bit_exact(name):
    """substeps, max CFL, energies (16 digits) and mass residual match the reference run."""
    b, h, q, dt = contract_cases.cases()[name]
    g = GOLD["contract_cases"][name]
    h1, qx1, qy1, rec = oracle.v2_step(b, h, q[..., 0], q[..., 1], dt)
    assert rec["status"] == 0
    assert rec["substeps"] == g["substeps"]
    assert rec["maximum_cfl"] == g["maximum_CFL"]
    e = g["energy"]
    k0, g0, i0, e0 = oracle.energy_terms(h, q[..., 0], q[..., 1], b, 0.05, 0.05)
    k1, g1, i1, e1 = oracle.energy_terms(h1, qx1, qy1, b, 0.05, 0.05)
    assert (k0, g0, i0, e0) == (e["kinetic_before_j_per_density"], e["gravitational_before_j_per_density"],
                                e["internal_before_j_per_density"], e["before_j_per_density"])
    assert (k1, g1, i1, e1) == (e["kinetic_after_j_per_density"], e["gravitational_after_j_per_density"],
                                e["internal_after_j_per_density"], e["after_j_per_density"])
    assert rec["friction_dissipation"] == e["friction_dissipation_j_per_density"]
    assert rec["external_work"] == e["external_work_j_per_density"]
    m0 = float(np.sum(h, dtype=np.float64) * 0.05 * 0.05)
    m1 = float(np.sum(h1, dtype=np.float64) * 0.05 * 0.05)
    assert m1 - m0 == g["mass_residual_m3"]
    assert float(np.min(h1, initial=0.0)) == g["minimum_h_m"]
    # momentum accounting residuals are ~1e-17 sums whose last bits depend on
    # the NumPy build's reduction order; they agree to 1e-16.
    p0 = np.sum(q, axis=(0, 1)) * 0.05 * 0.05
    p1 = np.sum(np.stack((qx1, qy1), -1), axis=(0, 1)) * 0.05 * 0.05
    res = p1 - p0 - (rec["conservative_impulse"] + rec["friction_impulse"] + rec["external_impulse"])
    assert np.max(np.abs(res - np.asarray(g["momentum_accounting_residual_m4_s"]))) <= 1.0e-16


def test_valley_single_face_response_exact():
    """One shared face (columns 20|21) reproduces the reference's uphill-impulse check."""
    b, h, q, dt = contract_cases.cases()["VALLEY"]
    vals = [0.0 - oracle.v2_face(b[r, 20], b[r, 21], h[r, 20], h[r, 21], q[r, 20, 0], q[r, 20, 1],
                                 q[r, 21, 0], q[r, 21, 1], 0, dt / 0.05)[1] for r in range(41)]
    assert float(np.mean(vals)) == GOLD["contract_cases"]["VALLEY"]["low_to_forward_face_delta_q_m2_s"]


def test_lake_at_rest_is_balanced_in_the_interior():
    """b + K h constant with q = 0: zero mass flux, and interior momentum updates cancel.

    Closed walls have no boundary face (mobile_v2_reference.py:378-431), so only
    interior cells are in exact hydrostatic balance.
    """
    K = 0.45
    n = 17
    y, x = np.indices((n, n))
    b = 0.2 + 0.01 * np.sin(0.7 * x) * np.cos(0.5 * y)
    h = (0.4 - b) / K
    h1, qx1, qy1, rec = oracle.v2_step(b, h, np.zeros_like(h), np.zeros_like(h), 1.0e-3,
                                       static_exit=False)
    assert rec["substeps"] == 1
    assert np.max(np.abs(h1 - h)) < 1e-15
    assert np.max(np.abs(qx1[1:-1, 1:-1])) < 1e-15 and np.max(np.abs(qy1[1:-1, 1:-1])) < 1e-15


def test_friction_decrement_closed_form():
    """A uniform sheet on a flat bed loses mu*g*dt of speed per substep (no flux)."""
    n = 12
    b = np.zeros((n, n))
    h = np.full((n, n), 0.1)
    v = 0.5
    qx = h * v
    dt = 0.01
    h1, qx1, qy1, rec = oracle.v2_step(b, h, qx, np.zeros_like(h), dt, static_exit=False)
    # interior cells see no net flux (uniform state); speed drops by mu*g*sum(dt)
    expect = v - 0.55 * 9.81 * dt
    got = qx1[n // 2, n // 2] / h1[n // 2, n // 2]
    assert abs(got - expect) < 1e-12


def test_mass_invariance_and_mirror_symmetry():
    n = 33
    y, x = np.indices((n, n))
    b = 0.1 + 0.002 * ((x - 16) ** 2)
    h = 0.05 + 0.2 * np.exp(-((x - 16) ** 2 + (y - 16) ** 2) / 12.0)
    h1, qx1, qy1, rec = oracle.v2_step(b, h, np.zeros_like(h), np.zeros_like(h), 0.05)
    assert rec["substeps"] > 1
    assert abs(np.sum(h1) - np.sum(h)) <= 1e-12 * np.sum(h)
    np.testing.assert_allclose(h1, h1[:, ::-1], atol=1e-13)
    np.testing.assert_allclose(h1, h1[::-1, :], atol=1e-13)
    np.testing.assert_allclose(qx1, -qx1[:, ::-1], atol=1e-13)


def test_static_cases_take_zero_substeps():
    for name in ("STATIC_FLAT", "STATIC_SLOPE"):
        b, h, q, dt = contract_cases.cases()[name]
        h1, qx1, qy1, rec = oracle.v2_step(b, h, q[..., 0], q[..., 1], dt)
        assert rec["static_exit"] and rec["substeps"] == 0
        np.testing.assert_array_equal(h1, h)


def test_entrain_deposit_gate_a_closed_forms():
    """Gate A: 16x16 ones, entrain [0.03, 2.0] then deposit [0.01, 0.25] at (7,7),(7,8)."""
    import sys
    sys.path.insert(0, str(Path(__file__).resolve().parents[1] /
                           "iron-soil-physical-simulation-algorithom-for-isaac-sim_amd"))
    from isaac_bulk_hip.config import triangle_vertex_weights

    n = 16
    w = triangle_vertex_weights(n, n, 0.05, 0.05)
    z = np.zeros((n, n))
    b = np.ones((n, n))
    m = np.zeros((n, n))
    idx = np.asarray([7 * n + 7, 7 * n + 8])
    b1, m1, vol = oracle.entrain(z.ravel(), b.ravel(), m.ravel(), w.ravel(), idx, [0.03, 2.0])
    assert vol == GOLD["gate_a"]["entrained_volume_m3"]
    assert np.max(np.abs((b1 + m1) - (b.ravel() + m.ravel()))) == 0.0
    b2, m2, qx2, qy2, out = oracle.deposit(b1, m1, np.zeros(n * n), np.zeros(n * n), w.ravel(), idx,
                                           [0.01, 0.25], 1370.0)
    assert out[0] == GOLD["gate_a"]["deposited_volume_m3"]
    assert list(out[1:]) == GOLD["gate_a"]["deposition_removed_momentum_kg_m_s"][:2]
    assert np.max(np.abs((b2 + m2) - (b1 + m1))) == 0.0


def _toe_and_whole(nx=701, dx=0.05):
    critical_drop = np.tan(np.deg2rad(30.0)) * dx
    baseline = np.repeat((1.0 + np.arange(nx) * 0.95 * critical_drop)[None, :], nx, axis=0)
    toe = baseline.copy()
    toe[346:355, 0] -= 0.10
    rows, cols = np.indices((nx, nx))
    whole = np.ones((nx, nx)) + 0.030 * ((rows + cols) & 1)
    return toe, whole


def test_relaxation_published_iteration_counts_701():
    """docs/MINISLOPE_SPARSE_TILE_PERFORMANCE.md: toe 183, whole pile 2 (30 deg, tol 1e-8)."""
    crit = float(np.tan(np.deg2rad(30.0)))
    toe, whole = _toe_and_whole()
    _, it, conv, _ = oracle.relax_yx(toe, 0.05, 0.05, crit, 5000, 1e-8)
    assert (it, conv) == (183, True)
    _, it2, conv2, _ = oracle.relax_yx(whole, 0.05, 0.05, crit, 5000, 1e-8)
    assert (it2, conv2) == (2, True)


def test_relaxation_invariants_spike_and_plane():
    """test_slope_model.py:18-34: conservation, slope <= tan+1e-7, stable plane unchanged."""
    import isaac_bulk_hip.relaxation as R

    h = np.zeros((31, 31))
    h[15, 15] = 10.0
    crit = float(np.tan(np.deg2rad(30.0)))
    out, it, conv, _ = oracle.relax_legacy(h, 0.1, 0.1, crit)
    assert conv
    assert abs(out.sum() - h.sum()) < 1e-10
    assert R.max_neighbor_slope(out, (0.1, 0.1)) <= crit + 1e-7
    assert out.min() >= 0.0
    x = np.arange(20)[:, None] * 0.02
    plane = np.broadcast_to(x, (20, 17)).copy()
    out2, _, _, _ = oracle.relax_legacy(plane, 0.1, 0.1, crit)
    np.testing.assert_array_equal(out2, plane)


def test_relaxation_two_cell_pair_brute_force():
    """A 2x2 block: the first sub-phase moves exactly half the excess of the (1,0) edges."""
    crit = float(np.tan(np.deg2rad(30.0)))
    h = np.array([[1.0, 0.2], [0.3, 0.2]])
    out, it, conv, _ = oracle.relax_legacy(h, 0.1, 0.1, crit, max_iterations=1)
    d = 1.0 - 0.3
    ex = d - crit * 0.1
    a = 1.0 - 0.5 * ex
    # later sub-phases also move mass; replay them by hand in numpy
    hh = h.copy()
    dist = [0.1, 0.1, float(np.hypot(0.1, 0.1)), float(np.hypot(0.1, 0.1))]
    views = [lambda z: (z[:-1, :], z[1:, :]), lambda z: (z[:, :-1], z[:, 1:]),
             lambda z: (z[:-1, :-1], z[1:, 1:]), lambda z: (z[:-1, 1:], z[1:, :-1])]
    for di, view in enumerate(views):
        for phase in (0, 1):
            A, B = view(hh)
            coord = np.arange(A.shape[0])[:, None] if di != 1 else np.arange(A.shape[1])[None, :]
            active = (coord & 1) == phase
            diff = A - B
            exc = np.maximum(np.abs(diff) - crit * dist[di], 0)
            t = 0.5 * exc * np.sign(diff) * active
            A -= t
            B += t
            if di == 0 and phase == 0:
                assert hh[0, 0] == a
    np.testing.assert_array_equal(out, hh)


def test_float_relaxation_tracks_double():
    rng = np.random.default_rng(5)
    h = rng.uniform(0.0, 2.0, size=(40, 37))
    crit = float(np.tan(np.deg2rad(34.0)))
    o64, it64, _, _ = oracle.relax_legacy(h, 0.05, 0.05, crit, 2000, 1e-6)
    o32, it32, _, _ = oracle.relax_legacy(h, 0.05, 0.05, crit, 2000, 1e-6, dtype=np.float32)
    assert np.max(np.abs(o64 - o32)) < 1e-4
    assert abs(it64 - it32) <= max(3, it64 // 10)


def _canonical_scipy(mask, connectivity):
    structure = np.ones((3, 3)) if connectivity == 8 else None
    lab, n = scipy.ndimage.label(mask, structure=structure)
    out = np.full(mask.shape, -1, dtype=np.int64)
    flat = lab.ravel()
    for k in range(1, n + 1):
        idx = np.flatnonzero(flat == k)
        out.ravel()[idx] = idx.min()
    return out


@pytest.mark.parametrize("density", [0.3, 0.5, 0.6])
@pytest.mark.parametrize("connectivity", [4, 8])
def test_oracle_labels_match_scipy(density, connectivity):
    rng = np.random.default_rng(int(density * 10) + connectivity)
    mask = rng.random((64, 57)) < density
    lab = oracle.label(mask.astype(np.int32), connectivity)
    np.testing.assert_array_equal(lab, _canonical_scipy(mask, connectivity))


def test_yield_analytic_numbers_and_cohesive_plane_is_stable():
    """test_large_avalanche_transition.py:387-410: drive 763.0389..., resistance 2096.1513..."""
    rho, g, d = 2200.0, 9.81, 0.05
    slope = np.deg2rad(45.0)
    mag = np.tan(slope)
    norm = np.sqrt(1.0 + mag * mag)
    drive = rho * g * d * mag / norm
    normal = rho * g * d / norm
    resist = 1500.0 + normal * np.tan(np.deg2rad(38.0))
    assert abs(drive - 763.0389275784034) < 1e-6
    assert abs(resist - 2096.1513465821067) < 1e-6
    n = 21
    x = np.arange(n) * 0.05
    plane = np.repeat((2.0 - np.tan(slope) * x)[None, :], n, axis=0)
    z = np.zeros((n, n))
    out = oracle.avalanche_initialize(
        z, plane, z, z, z, np.ones((n, n)), z, z, np.zeros((n, n), np.int32), dx=0.05, dy=0.05,
        tan_start=float(np.tan(np.deg2rad(38.0))), tan_stop=float(np.tan(np.deg2rad(30.0))),
        density=rho, gravity=g, cohesion=1500.0, mobilization_depth=d, dry=1e-8,
        activity_speed=0.05, release=True)
    assert int(out["unstable"].sum()) == 0
    out0 = oracle.avalanche_initialize(
        z, plane, z, z, z, np.ones((n, n)), z, z, np.zeros((n, n), np.int32), dx=0.05, dy=0.05,
        tan_start=float(np.tan(np.deg2rad(38.0))), tan_stop=float(np.tan(np.deg2rad(30.0))),
        density=rho, gravity=g, cohesion=0.0, mobilization_depth=d, dry=1e-8,
        activity_speed=0.05, release=True)
    assert int(out0["unstable"].sum()) == n * n


def test_disconnected_spikes_give_four_regions():
    """test_large_avalanche_transition.py:307-328: four isolated spikes -> >= 4 regions."""
    n = 21
    resting = np.ones((n, n))
    for r, c in ((4, 4), (4, 16), (16, 4), (16, 16)):
        resting[r, c] += 0.2
    z = np.zeros((n, n))
    out = oracle.avalanche_initialize(
        z, resting, z, z, z, np.ones((n, n)), z, z, np.zeros((n, n), np.int32), dx=0.05, dy=0.05,
        tan_start=float(np.tan(np.deg2rad(38.0))), tan_stop=float(np.tan(np.deg2rad(30.0))),
        density=2200.0, gravity=9.81, cohesion=0.0, mobilization_depth=0.05, dry=1e-8,
        activity_speed=0.05, release=True)
    lab = oracle.label(out["unstable"], 8)
    di, df = oracle.components(lab, z, resting, np.zeros((n, n), np.int32), np.ones((n, n)),
                               out["slope"], out["severity"], np.zeros((n, n), np.int32), 0.05,
                               float(np.deg2rad(38.0)), float(np.deg2rad(30.0)))
    assert di[0] >= 4
    assert di[3] == int(out["unstable"].sum())


def test_threaded_oracle_gives_the_same_fields():
    """The OpenMP build of the restatement (CPU baseline on every core) gives
    bit-identical fields and the same substep count."""
    import os

    from isaac_bulk_hip.workloads import pile_with_patch

    b, h, qx, qy = pile_with_patch(96, seed=3)
    os.environ.setdefault("OMP_NUM_THREADS", "4")
    h1, qx1, qy1, r1 = oracle.v2_step(b, h, qx, qy, 0.5)
    h2, qx2, qy2, r2 = oracle.v2_step(b, h, qx, qy, 0.5, threads=True)
    assert r1["substeps"] == r2["substeps"] > 3
    np.testing.assert_array_equal(h1, h2)
    np.testing.assert_array_equal(qx1, qx2)
    np.testing.assert_array_equal(qy1, qy2)


def test_persistence_reduction_hand_case():
    """reduce_persistence (gpu_large_avalanche.py:378-451) on a 6 x 8 grid
    whose every scalar is known by hand: one dry cell activated twice at a
    slope below the stop angle (a retired candidate), one wet cell moving at
    0.5 m/s, one unstable cell, and a mass-conserving move of 0.02 m."""
    ny, nx, ts = 6, 8, 4
    w = np.full((ny, nx), 0.0025)
    initial = np.ones((ny, nx))
    resting = initial.copy()
    resting[0, 0] += 0.02
    resting[5, 7] -= 0.02
    mobile = np.zeros((ny, nx))
    qx = np.zeros((ny, nx))
    qy = np.zeros((ny, nx))
    mobile[1, 6] = 0.1
    qx[1, 6], qy[1, 6] = 0.03, 0.04           # |v| = 0.5 m/s
    resting[1, 6] -= 0.1                       # the mobile layer came from below
    slope = np.zeros((ny, nx))
    count = np.zeros((ny, nx), dtype=np.int32)
    count[4, 2], slope[4, 2] = 2, 0.4
    unstable = np.zeros((ny, nx), dtype=np.int32)
    unstable[4, 5] = 1
    latch = np.zeros((ny, nx), dtype=np.int32)
    r2m = np.zeros((ny, nx))
    r2m[4, 2] = 3e-4
    m2r = np.zeros((ny, nx))
    m2r[1, 6] = 1e-4
    stop, start = np.deg2rad(30.0), np.deg2rad(38.0)
    di, df, flags = oracle.persistence(initial, resting, mobile, qx, qy, w, slope, unstable, latch,
                                       count, r2m, m2r, tile_size=ts, start_rad=start, stop_rad=stop,
                                       dry=1e-9, activity_speed=0.05)
    assert list(di) == [1, 1, 1, 1]
    assert df[0] == 3e-4 and df[1] == 1e-4
    assert df[2] == 0.0025 and df[3] == 0.0025 and df[12] == 0.0025
    assert df[4] == pytest.approx(0.1 * 0.0025) and df[13] == df[4]
    assert df[5] == pytest.approx(0.5, rel=1e-15)
    assert df[6] == pytest.approx(0.0, abs=1e-18)          # the moves conserve the surface
    assert df[7] == pytest.approx(0.5 * 0.04 * 0.0025, rel=1e-12)
    assert df[8] == 0.4 and df[9] == 0.4
    assert df[10] == pytest.approx(0.4 - stop) and df[11] == pytest.approx(start - 0.4)
    # tiles (2 x 2 of 4 x 4 cells): the moving cell (1, 6) -> tile 1; the unstable (4, 5) -> tile 3
    assert list(flags) == [0, 1, 0, 1]


def test_surface_sample_reproduces_planes_and_grid_values():
    """bilinear_surface (bulk_state_authority.py:196-228): a bilinear
    interpolant reproduces a plane everywhere inside the grid, returns the
    grid value at integer points, and clamps queries to the grid."""
    ny, nx = 9, 13
    rr, cc = np.meshgrid(np.arange(ny, dtype=float), np.arange(nx, dtype=float), indexing="ij")
    plane = 0.7 + 0.03 * rr - 0.011 * cc
    m = np.zeros((ny, nx))
    m[2, 3] = 0.25
    rng = np.random.default_rng(2)
    q = np.column_stack((rng.uniform(0, ny - 1, 200), rng.uniform(0, nx - 1, 200)))
    got = oracle.surface_sample(plane, np.zeros((ny, nx)), q)
    np.testing.assert_allclose(got, 0.7 + 0.03 * q[:, 0] - 0.011 * q[:, 1], rtol=0, atol=1e-15)
    pts = np.array([[2.0, 3.0], [0.0, 0.0], [ny - 1.0, nx - 1.0], [4.0, 12.0]])
    np.testing.assert_array_equal(oracle.surface_sample(plane, m, pts),
                                  [plane[2, 3] + 0.25, plane[0, 0], plane[-1, -1], plane[4, 12]])
    far = np.array([[-5.0, -3.0], [100.0, 200.0], [2.0, -1.0]])
    np.testing.assert_array_equal(oracle.surface_sample(plane, m, far),
                                  [plane[0, 0], plane[-1, -1], plane[2, 0]])
    # halfway between the bump cell and its neighbour: the mean of the two
    assert oracle.surface_sample(plane, m, [[2.0, 3.5]])[0] == pytest.approx(
        0.5 * (plane[2, 3] + 0.25 + plane[2, 4]), rel=1e-15)


def _components_by_scipy(mask, connectivity):
    """Components of a mask with scipy.ndimage.label: (sizes, min linear index) per component."""
    structure = np.ones((3, 3)) if connectivity == 8 else None
    lab, nfeat = scipy.ndimage.label(mask, structure=structure)
    flat = lab.ravel()
    sizes = np.bincount(flat)[1:]
    first = np.full(nfeat, flat.size)
    idx = np.arange(flat.size)
    np.minimum.at(first, flat[flat > 0] - 1, idx[flat > 0])
    return sizes, first


@pytest.mark.parametrize("connectivity", [4, 8])
def test_component_statistics_against_scipy_and_closed_forms(connectivity):
    """reduce_components / select_winner / winner_diagnostics
    (gpu_large_avalanche.py:201-281): region count, unstable cells and the
    largest size from scipy's components; the winner is the largest component
    whose first cell comes first (ties included: two equal largest blocks);
    on the winner every cell has slope start + 0.1, stop + 0.3, transfer
    min(resting - z, depth x severity) = 0.2 unless latched, so every sum is a
    closed form."""
    ny, nx = 20, 24
    mask = np.zeros((ny, nx), np.int32)
    mask[2:6, 3:9] = 1                   # 24 cells, first index 2*24+3
    mask[12:16, 10:16] = 1               # 24 cells (tie), later first index
    mask[8, 20] = mask[9, 21] = 1        # diagonal pair: one region under 8-connectivity only
    mask[18, 1:4] = 1
    rng = np.random.default_rng(1)
    start, stop = np.deg2rad(38.0), np.deg2rad(30.0)
    slope = np.where(mask > 0, start + 0.1, rng.uniform(0.0, 0.3, (ny, nx)))
    z = np.zeros((ny, nx))
    resting = np.full((ny, nx), 0.5)
    severity = np.ones((ny, nx))
    latch = np.zeros((ny, nx), np.int32)
    latch[2, 3] = 1                      # the winner's first cell is latched: no transfer there
    previous = np.zeros((ny, nx), np.int32)
    previous[3, 4] = previous[13, 11] = 1
    weights = oracle.vertex_weights(ny, nx, 0.05, 0.05)
    lab = oracle.label(mask, connectivity)
    di, df = oracle.components(lab, z, resting, latch, weights, slope, severity, previous, 0.2, start, stop)
    sizes, first = _components_by_scipy(mask, connectivity)
    assert di[0] == sizes.size
    assert di[1] == sizes.max() == 24
    assert di[3] == int(mask.sum())
    winners = np.flatnonzero(sizes == sizes.max())
    assert di[2] == first[winners].min() == 2 * nx + 3
    win = np.zeros((ny, nx), bool)
    win[2:6, 3:9] = True
    assert di[4] == 1                                            # the winner's previous overlap
    assert df[0] == pytest.approx(24 * 0.1, rel=1e-12) and df[6] == df[0]
    assert df[1] == pytest.approx(0.1, rel=1e-12)
    assert df[2] == pytest.approx(24 * (start + 0.1 - stop), rel=1e-12)
    assert df[3] == pytest.approx(start + 0.1 - stop, rel=1e-12)
    assert df[4] == pytest.approx(float(weights[win].sum()), rel=1e-14)
    moved = win.copy()
    moved[2, 3] = False
    assert df[5] == pytest.approx(0.2 * float(weights[moved].sum()), rel=1e-12)


def test_mobilize_closed_forms():
    """mobilize (gpu_large_avalanche.py:309-376) on a constructed winner:
    t = min(max(b - z, 0), depth x severity) moves from b to h on every
    unlatched winner cell above the dry tolerance, the free surface is kept,
    the cell latches, its activation count and times advance, and the four
    weighted sums split first activations from re-activations."""
    ny, nx = 10, 12
    labels = np.full((ny, nx), -1, np.int32)
    labels[2:5, 2:7] = 2 * nx + 2                     # the winner (root = first cell)
    labels[7:9, 1:4] = 7 * nx + 1                     # another component
    z = np.zeros((ny, nx))
    resting = np.full((ny, nx), 0.8)
    resting[2, 2:7] = 0.05                            # shallow: depth-limited transfer
    resting[3, 5] = 1e-12                             # below the dry tolerance: no move
    mobile = np.full((ny, nx), 0.01)
    severity = np.full((ny, nx), 0.5)
    latch = np.zeros((ny, nx), np.int32)
    latch[4, 6] = 1
    count = np.zeros((ny, nx), np.int32)
    count[4, 2:6] = 2                                 # re-activations
    weights = oracle.vertex_weights(ny, nx, 0.05, 0.05)
    zeros = np.zeros((ny, nx))
    out = oracle.mobilize(labels, 2 * nx + 2, z, resting, mobile, weights, latch, zeros, zeros, zeros,
                          severity, count, zeros, zeros, zeros, 7.5, 0.3, 1e-9)
    expect_t = np.zeros((ny, nx))
    win = labels == 2 * nx + 2
    expect_t[win] = np.minimum(resting[win], 0.3 * 0.5)
    expect_t[4, 6] = 0.0
    expect_t[3, 5] = 0.0
    moved = expect_t > 0
    np.testing.assert_array_equal(out["resting"][moved], resting[moved] - expect_t[moved])
    np.testing.assert_array_equal(out["mobile"][moved], mobile[moved] + expect_t[moved])
    np.testing.assert_array_equal(out["resting"][~moved], resting[~moved])
    np.testing.assert_allclose(out["resting"] + out["mobile"], resting + mobile, rtol=0, atol=2e-16)
    np.testing.assert_array_equal(out["latch"], np.where(moved, 1, latch))
    np.testing.assert_array_equal(out["activation_count"], count + moved)
    first = moved & (count == 0)
    np.testing.assert_array_equal(out["first_time"], np.where(first, 7.5, 0.0))
    np.testing.assert_array_equal(out["last_time"], np.where(moved, 7.5, 0.0))
    d = out["diag"]
    assert d[7] == pytest.approx(float(np.sum(expect_t * weights)), rel=1e-13)
    assert d[10] == pytest.approx(float(np.sum((expect_t * weights)[first])), rel=1e-13)
    assert d[11] == pytest.approx(float(np.sum((expect_t * weights)[moved & ~first])), rel=1e-13)
    assert d[12] == pytest.approx(float(weights[first].sum()), rel=1e-14)
    assert d[13] == pytest.approx(float(weights[moved & ~first].sum()), rel=1e-14)


def test_vertex_weights_fast_equals_the_per_cell_restatement():
    for ny, nx, dx, dy in ((7, 9, 0.05, 0.05), (12, 5, 0.1, 0.04), (2, 2, 0.05, 0.07)):
        np.testing.assert_array_equal(oracle.vertex_weights_fast(ny, nx, dx, dy),
                                      oracle.vertex_weights(ny, nx, dx, dy))
        # the weights tile the rectangle: their sum is its area (ny-1)(nx-1) dx dy
        assert float(oracle.vertex_weights_fast(ny, nx, dx, dy).sum()) == pytest.approx(
            (ny - 1) * (nx - 1) * dx * dy, rel=1e-14)


def test_frame_sequence_keeps_mass_and_ends_each_frame_relaxed():
    """The FP64 frame of the full-size workload at a small size: the V2 layer
    keeps its mass, b_eff + h is kept by the relaxation, and every frame ends
    with no legacy edge above the critical difference by more than the
    tolerance (or at the iteration cap)."""
    from frame_sequence import OracleFrameSequence
    from isaac_bulk_hip.workloads import cut_face_terrain

    b, h, _ = cut_face_terrain(64, 0.05, seed=2)
    seq = OracleFrameSequence(b, h, dx=0.05, dt=1.0 / 60.0, v2={"limit_mode": 1}, relax_angle_deg=34.0,
                              relax_max_iterations=1000, relax_tolerance=1.0e-6,
                              avalanche={"tan_start": float(np.tan(np.deg2rad(34.0))),
                                         "tan_stop": float(np.tan(np.deg2rad(30.0))), "density": 1370.0,
                                         "gravity": 9.81, "cohesion": 0.0, "mobilization_depth": 0.08,
                                         "dry": 1.0e-8, "activity_speed": 0.05, "release": True})
    crit = float(np.tan(np.deg2rad(34.0)))
    diag = float(np.hypot(0.05, 0.05))
    for _ in range(4):
        rec = seq.step()
        assert rec["substeps"] > 0
        assert seq.mass_error() <= 1e-14
        assert seq.total_error() <= 1e-13
        if rec["relax_converged"]:
            bb = seq.b
            for a, c, d in ((bb[:-1, :], bb[1:, :], 0.05), (bb[:, :-1], bb[:, 1:], 0.05),
                            (bb[:-1, :-1], bb[1:, 1:], diag), (bb[:-1, 1:], bb[1:, :-1], diag)):
                assert float(np.max(np.abs(a - c) - crit * d)) <= 1e-6 + 1e-12
        np.testing.assert_array_equal(rec["labels"], oracle.label(rec["unstable"], 8))


def _exchange_samples(rng, n):
    """float32 heights across binades (0.01 .. 3.9 m, some exactly at binade
    edges and one float below them) and requests of 0 .. 4000 quanta."""
    b = np.concatenate([rng.uniform(0.01, 3.9, n).astype(np.float32),
                        np.float32([0.25, 0.5, 1.0, 2.0]),
                        np.nextafter(np.float32([0.25, 0.5, 1.0, 2.0]), np.float32(0))])
    q = rng.integers(0, 4000, b.size)
    return b, q


def test_fp32_entrain_amount_exact_bounded_and_near_maximal():
    """H_free = b_eff + mobile stays bit-invariant: b - moved is a float32 and
    moved is whole quanta; moved never exceeds the donor-limited depth, and it
    is within one step max(quantum, spacing(b)) of it; brute force confirms no
    exact amount lies above it within that step... except via lower binades."""
    rng = np.random.default_rng(5)
    b, q = _exchange_samples(rng, 400)
    Q = oracle.QUANTUM
    for bi, qi in zip(b, q):
        depth = qi * Q * 1.37                      # not a whole number of quanta
        dq = oracle.entrain_quanta_f32(bi, depth)
        moved = dq * Q
        assert moved <= depth
        assert float(np.float32(float(bi) - moved)) == float(bi) - moved       # exactly on the float grid
        g = max(Q, oracle.f32_spacing(bi))
        assert moved > depth - g - 1e-18
        # brute force: every larger whole-quanta amount up to the depth either
        # leaves the float grid or stays within one step of the chosen one
        for k in range(dq + 1, int(np.floor(depth / Q)) + 1):
            if float(np.float32(float(bi) - k * Q)) == float(bi) - k * Q:
                assert (k - dq) * Q < g


def test_fp32_deposit_amount_exact_unless_crossing_off_grid():
    rng = np.random.default_rng(6)
    b, q = _exchange_samples(rng, 400)
    Q = oracle.QUANTUM
    for bi, qi in zip(b, q):
        dq, b_new = oracle.deposit_quanta_f32(bi, int(qi))
        assert 0 <= dq <= qi
        on_grid = np.floor(float(bi) / Q) == float(bi) / Q
        exact = float(b_new) == float(bi) + dq * Q
        if on_grid:
            # on the quantum grid: always exact, and within one float spacing
            # of the request (brute force finds nothing exact in between)
            assert exact
            for k in range(dq + 1, int(qi) + 1):
                assert float(np.float32(float(bi) + k * Q)) != float(bi) + k * Q
        elif not exact:
            # the rounded crossing: everything moved, b_eff left its binade,
            # the error is below half the new spacing, and b_eff is on the grid after
            assert dq == qi
            err = abs(float(b_new) - (float(bi) + dq * Q))
            assert err <= 0.5 * oracle.f32_spacing(b_new)
            assert oracle.f32_spacing(b_new) > oracle.f32_spacing(bi)
            assert np.floor(float(b_new) / Q) == float(b_new) / Q or oracle.f32_spacing(b_new) < Q
    # planted mistakes: the pre-existing rounding rule breaks exactness
    bad = [float(np.float32(float(bi) + int(qi) * Q)) == float(bi) + int(qi) * Q for bi, qi in zip(b, q) if qi]
    assert not all(bad)


def test_failure_activation_hand_case():
    """A 5 x 6 patch at (2, 3) of a 10 x 12 field: activation 0.01 m on a 3 x 3
    block, one affected cell in its centre and one affected dry cell on the
    patch edge.  Forced: the centre and its four neighbours (all wet), not the
    diagonals, nothing around the dry affected cell; the flat indices come in
    row-major order; the moved volume is 9 x 0.01 x w; H_free is unchanged."""
    ny, nx = 10, 12
    z = np.zeros((ny, nx))
    b = np.full((ny, nx), 1.0)
    m = np.zeros((ny, nx))
    w = oracle.vertex_weights_fast(ny, nx, 0.05, 0.05)
    bbox = (2, 7, 3, 9)
    act = np.zeros((5, 6))
    act[1:4, 1:4] = 0.01
    aff = np.zeros((5, 6), bool)
    aff[2, 2] = True
    aff[0, 5] = True                                   # dry: forces nothing
    b1, m1, vol, forcing = oracle.failure_activation(z, b, m, w, bbox, act, aff)
    centre = (2 + 2) * nx + (3 + 2)
    assert forcing.tolist() == [centre - nx, centre - 1, centre, centre + 1, centre + nx]
    assert vol == pytest.approx(9 * 0.01 * 0.05 * 0.05, rel=1e-14)
    np.testing.assert_array_equal(b1 + m1, b + m)
    # FP32 storage: the same forcing set, whole quanta, H_free bit-invariant
    b32, m32, vol32, forcing32 = oracle.failure_activation(z, b.astype(np.float32), np.zeros((ny, nx), np.int32),
                                                          w, bbox, act, aff, fp32=True)
    assert forcing32.tolist() == forcing.tolist()
    np.testing.assert_array_equal(b32.astype(np.float64) + m32 * oracle.QUANTUM, b)
    assert vol32 == pytest.approx(vol, rel=1e-6)
    # planted mistake: 8-neighbourhood forcing would add the diagonals
    assert (centre - nx - 1) not in forcing.tolist()


def test_per_cell_external_acceleration_closed_form():
    """A uniform 0.1 m layer at rest on a flat bed, pushed for one substep by
    a per-cell acceleration (20 m/s^2 along x in the left half, none
    elsewhere): the faces carry nothing, so each pushed cell ends with
    q = h (a dt)(1 - mu g / a) and every other cell stays at rest; the static
    exit is skipped because max |a| > 1e-14 (mobile_v2_reference.py:226-229)."""
    ny, nx, dt, mu = 6, 10, 1.0e-3, 0.55
    b = np.zeros((ny, nx))
    h = np.full((ny, nx), 0.1)
    z = np.zeros((ny, nx))
    ext = np.zeros((ny, nx, 2))
    ext[:, : nx // 2, 0] = 20.0
    h1, qx1, qy1, rec = oracle.v2_step(b, h, z, z, dt, mu=mu, external=ext)
    assert rec["substeps"] == 1 and not rec["static_exit"]
    # interior cells: both faces equal, so the flux terms cancel exactly (the
    # wall cells have no boundary face and take a one-sided pressure push)
    expect = 0.1 * (20.0 * dt) * (1.0 - mu * 9.81 * dt / (20.0 * dt))
    np.testing.assert_allclose(qx1[1:-1, 1: nx // 2], expect, rtol=1e-12)
    assert np.all(qx1[1:-1, nx // 2:-1] == 0.0) and np.all(qy1[1:-1, 1:-1] == 0.0)
    np.testing.assert_array_equal(h1, h)
    # the impulse of the push over the pushed cells, per the reference's ledger
    assert rec["external_impulse"][0] == pytest.approx(ny * (nx // 2) * 0.1 * 20.0 * dt * 0.05 * 0.05, rel=1e-12)
    # planted mistake: the uniform field of the first cell's value pushes every cell
    _, qx_u, _, _ = oracle.v2_step(b, h, z, z, dt, mu=mu, external=(20.0, 0.0))
    assert np.all(qx_u[1:-1, 1:-1] > 0.0)
