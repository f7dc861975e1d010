"""ctypes front-end of the CPU oracle (``liboracle.so``).

TEST INFRASTRUCTURE ONLY: imported by ``tests/``, ``__graft_entry__.smoke()``
and the ``cpu_baseline`` leg of ``bench.py`` -- never by the product package.
The C restatement in ``mv2_oracle.c`` cites the reference routine each entry
follows.
"""

from __future__ import annotations

import ctypes
import os
import subprocess
from pathlib import Path

import numpy as np

_HERE = Path(__file__).resolve().parent
_LIB_PATH = _HERE / "liboracle.so"
_lib: ctypes.CDLL | None = None

_D = ctypes.c_double
_I = ctypes.c_int
_P = ctypes.c_void_p


class V2Params(ctypes.Structure):
    _fields_ = [
        ("dx", _D), ("dy", _D), ("gravity", _D), ("earth_pressure", _D),
        ("cfl", _D), ("dry_tolerance", _D), ("basal_friction", _D),
        ("tan_start", _D), ("external_ax", _D), ("external_ay", _D),
        ("maximum_substeps", _I), ("limit_mode", _I), ("static_exit", _I),
    ]


class DepositionParams(ctypes.Structure):
    _fields_ = [
        ("dx", _D), ("dy", _D), ("dt", _D), ("settling_rate", _D), ("speed_threshold", _D),
        ("density", _D), ("gravity", _D), ("cohesion", _D), ("tan_stop", _D),
        ("settle_subcell_tail", _I),
    ]


class V2Result(ctypes.Structure):
    _fields_ = [
        ("status", _I), ("substeps", _I), ("cfl_limited", _I), ("static_exit", _I),
        ("maximum_cfl", _D), ("remaining", _D),
        ("conservative_impulse", _D * 2), ("friction_impulse", _D * 2),
        ("external_impulse", _D * 2), ("friction_dissipation", _D),
        ("external_work", _D), ("diag", _D * 8),
    ]


class AvalancheParams(ctypes.Structure):
    _fields_ = [
        ("dx", _D), ("dy", _D), ("tan_start", _D), ("tan_stop", _D),
        ("density", _D), ("gravity", _D), ("cohesion", _D),
        ("mobilization_depth", _D), ("dry_tolerance", _D), ("activity_speed", _D),
        ("release_settled_latches", _I),
    ]


def build() -> Path:
    """Compile the oracle with its Makefile (idempotent)."""
    subprocess.run(["make", "-s", "-C", str(_HERE)], check=True)
    return _LIB_PATH


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not _LIB_PATH.exists():
            build()
        _lib = ctypes.CDLL(str(_LIB_PATH))
        _lib.orc_pairwise_sum.restype = _D
        _lib.orc_pairwise_sum.argtypes = [_P, ctypes.c_int64, ctypes.c_int64]
        _lib.orc_v2_step.argtypes = [_I, _I, ctypes.POINTER(V2Params), _P, _P, _P, _P, _D,
                                     ctypes.POINTER(V2Result)]
        _lib.orc_v2_step.restype = _I
        _lib.orc_v2_step_ext.argtypes = [_I, _I, ctypes.POINTER(V2Params), _P, _P, _P, _P, _D, _P, _P,
                                         ctypes.POINTER(V2Result)]
        _lib.orc_v2_step_ext.restype = _I
        _lib.orc_v2_face.argtypes = [_D] * 8 + [_I, _D, _D, _D, _D, _P]
        _lib.orc_v2_static.argtypes = [_I, _I, _D, _D, _P, _P, _P, _P, _P]
        _lib.orc_vertex_weight.restype = _D
        _lib.orc_vertex_weight.argtypes = [_I, _I, _D, _D, _I, _I]
        _lib.orc_v2_summary.argtypes = [_I, _I, _D, _D, _I, _D, _D, _D, _D, _P, _P, _P, _P, _P]
        _lib.orc_entrain.restype = _D
        _lib.orc_entrain.argtypes = [_I, _P, _P, _P, _P, _P, _P, _I]
        _lib.orc_deposit.argtypes = [_I, _P, _P, _P, _P, _P, _P, _P, _I, _D, _P]
        _lib.orc_reservoirs.argtypes = [_I, _P, _P, _P, _P, _P, _P, _D, _P]
        for name in ("orc_deposition", "orc_deposition32"):
            getattr(_lib, name).argtypes = [_I, _I, ctypes.POINTER(DepositionParams), _P] + [_P] * 10
        for name in ("orc_relax_legacy_f64", "orc_relax_legacy_f32"):
            fn = getattr(_lib, name)
            fn.restype = _I
            fn.argtypes = [_I, _I, _D, _D, _D, _I, _D, _P, ctypes.POINTER(_I),
                           ctypes.POINTER(_D)]
        _lib.orc_frontier_relax.restype = _I
        _lib.orc_frontier_relax.argtypes = [_I, _I, _I, _I, _P, _P, _P, _P, _P, _P, _P, _I,
                                            _D, _P]
        _lib.orc_frontier_scan.argtypes = [_I, _P, _P, _P, _P, _P, _P, _P, _I, _D, _D, _P]
        _lib.orc_avalanche_initialize.argtypes = [
            _I, _I, ctypes.POINTER(AvalancheParams), _P, _P, _P, _P, _P, _P, _P, _P,
            _P, _P, _P, _P, _P, _P]
        _lib.orc_label.argtypes = [_I, _I, _I, _P, _P]
        _lib.orc_surface_sample.argtypes = [_I, _I, _P, _P, _P, _I, _P]
        _lib.orc_persistence.argtypes = [_I, _I, _I, _I] + [_P] * 12 + [_D] * 4 + [_P] * 3
        _lib.orc_components.argtypes = [_I, _I, _P, _P, _P, _P, _P, _P, _P, _P, _D, _D, _D,
                                        _P, _P]
        _lib.orc_mobilize.argtypes = [_I, _P, ctypes.c_int32, _P, _P, _P, _P, _P, _P, _P, _P,
                                      _P, _P, _P, _P, _P, _D, _D, _D, _P]
    return _lib


_lib_omp: ctypes.CDLL | None = None


def lib_threads() -> ctypes.CDLL:
    """liboracle_omp.so: the same restatement with its per-cell loops over
    OpenMP threads (OMP_NUM_THREADS); fields identical to lib()'s."""
    global _lib_omp
    if _lib_omp is None:
        path = _HERE / "liboracle_omp.so"
        if not path.exists():
            build()
        _lib_omp = ctypes.CDLL(str(path))
        _lib_omp.orc_v2_step.argtypes = [_I, _I, ctypes.POINTER(V2Params), _P, _P, _P, _P, _D,
                                         ctypes.POINTER(V2Result)]
        _lib_omp.orc_v2_step.restype = _I
    return _lib_omp


def _ptr(a: np.ndarray) -> int:
    return a.ctypes.data


def _f64(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64)


def _i32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.int32)


def pairwise_sum(a: np.ndarray) -> float:
    a = _f64(a).ravel()
    return float(lib().orc_pairwise_sum(_ptr(a), a.size, 1))


def v2_step(b, h, qx, qy, dt, *, dx=0.05, dy=0.05, gravity=9.81, K=0.45, cfl=0.18,
            dry=1.0e-10, mu=0.55, tan_start=None, start_angle_deg=34.0,
            external=(0.0, 0.0), maximum_substeps=4096, limit_mode=0, static_exit=True, threads=False):
    """Restated ``MobileV2ReferenceSolver.step`` (FP64).  Returns new fields + record.
    ``external`` is a uniform (ax, ay) pair or any array broadcastable to
    (ny, nx, 2), per cell as the reference takes it.  threads=True runs the
    OpenMP build of the same restatement."""
    b = _f64(b)
    h = _f64(h).copy()
    qx = _f64(qx).copy()
    qy = _f64(qy).copy()
    ny, nx = b.shape
    if tan_start is None:
        tan_start = float(np.tan(np.deg2rad(start_angle_deg)))
    ext = np.asarray(external, dtype=np.float64)
    per_cell = ext.shape != (2,)
    ux, uy = (0.0, 0.0) if per_cell else (float(ext[0]), float(ext[1]))
    p = V2Params(dx, dy, gravity, K, cfl, dry, mu, tan_start, ux, uy,
                 int(maximum_substeps), int(limit_mode), int(bool(static_exit)))
    r = V2Result()
    if per_cell:
        # mobile_v2_reference.py:204-210: broadcast to q.shape = (ny, nx, 2)
        e = np.ascontiguousarray(np.broadcast_to(ext, (ny, nx, 2)))
        ex, ey = np.ascontiguousarray(e[..., 0]), np.ascontiguousarray(e[..., 1])
        lib().orc_v2_step_ext(ny, nx, ctypes.byref(p), _ptr(b), _ptr(h), _ptr(qx), _ptr(qy), float(dt),
                              _ptr(ex), _ptr(ey), ctypes.byref(r))
    else:
        (lib_threads() if threads else lib()).orc_v2_step(ny, nx, ctypes.byref(p), _ptr(b), _ptr(h), _ptr(qx),
                                                          _ptr(qy), float(dt), ctypes.byref(r))
    rec = {
        "status": r.status, "substeps": r.substeps, "cfl_limited": bool(r.cfl_limited),
        "static_exit": bool(r.static_exit), "maximum_cfl": r.maximum_cfl,
        "remaining": r.remaining,
        "conservative_impulse": np.array(r.conservative_impulse[:]),
        "friction_impulse": np.array(r.friction_impulse[:]),
        "external_impulse": np.array(r.external_impulse[:]),
        "friction_dissipation": r.friction_dissipation,
        "external_work": r.external_work,
        "diag": np.array(r.diag[:]),
    }
    return h, qx, qy, rec


def v2_face(bl, br, hl, hr, qxl, qyl, qxr, qyr, normal, factor, K=0.45, g=9.81, dry=1.0e-10):
    """One shared face (already scaled by factor): [F_h, qx_left, qy_left, qx_right, qy_right]."""
    out = np.zeros(5)
    lib().orc_v2_face(bl, br, hl, hr, qxl, qyl, qxr, qyr, int(normal), K, g, dry, factor,
                      _ptr(out))
    return out


def v2_static(b, h, qx, qy, dx=0.05, dy=0.05):
    b, h, qx, qy = (_f64(v) for v in (b, h, qx, qy))
    out = np.zeros(2)
    lib().orc_v2_static(b.shape[0], b.shape[1], dx, dy, _ptr(b), _ptr(h), _ptr(qx), _ptr(qy),
                        _ptr(out))
    return out


def vertex_weights(ny, nx, dx, dy) -> np.ndarray:
    w = np.empty((ny, nx))
    for r in range(ny):
        for c in range(nx):
            w[r, c] = lib().orc_vertex_weight(ny, nx, dx, dy, r, c)
    return w


def vertex_weights_fast(ny, nx, dx, dy) -> np.ndarray:
    """Triangle-A-C vertex weights as whole-array sums
    (DeviceBulkState._vertex_weights, bulk_state_authority.py:624-634);
    equal to ``vertex_weights`` cell by cell."""
    w = np.zeros((ny, nx))
    f = dx * dy / 6.0
    w[:-1, :-1] += 2.0 * f
    w[:-1, 1:] += f
    w[1:, 1:] += 2.0 * f
    w[1:, :-1] += f
    return w


def v2_summary(b, h, qx, qy, *, dx, dy, weight_mode, density, K=0.45, g=9.81, dry=1.0e-10):
    b, h, qx, qy = (_f64(v) for v in (b, h, qx, qy))
    out = np.zeros(7)
    lib().orc_v2_summary(b.shape[0], b.shape[1], dx, dy, int(weight_mode), density, K, g,
                         dry, _ptr(b), _ptr(h), _ptr(qx), _ptr(qy), _ptr(out))
    return out


def entrain(z, b, m, w, idx, req):
    z, w, req = _f64(z), _f64(w), _f64(req)
    b, m = _f64(b).copy(), _f64(m).copy()
    idx = _i32(idx)
    v = lib().orc_entrain(b.size, _ptr(z), _ptr(b), _ptr(m), _ptr(w), _ptr(idx), _ptr(req),
                          idx.size)
    return b, m, float(v)


def deposit(b, m, qx, qy, w, idx, req, density):
    b, m, qx, qy = (_f64(v).copy() for v in (b, m, qx, qy))
    w, req, idx = _f64(w), _f64(req), _i32(idx)
    out = np.zeros(3)
    lib().orc_deposit(b.size, _ptr(b), _ptr(m), _ptr(qx), _ptr(qy), _ptr(w), _ptr(idx),
                      _ptr(req), idx.size, density, _ptr(out))
    return b, m, qx, qy, out


def deposition(b, m, qx, qy, cumulative, weights, *, dx, dy, dt, settling_rate=0.30,
               speed_threshold=0.18, density=1800.0, gravity=9.81, cohesion=0.0, tan_stop=0.0,
               settle_subcell_tail=False, forcing=None, fp32=False):
    """One WarpDepositionOperator.step_resident call on copies of the fields.

    fp32=False: float64 fields.  fp32=True: b/qx/qy float32, m int32 quanta.
    Returns (b, m, qx, qy, deposited, removed_x, removed_y, cumulative, sums[3]).
    """
    ny, nx = np.shape(b)
    if fp32:
        fd, md = np.float32, np.int32
    else:
        fd, md = np.float64, np.float64
    b, qx, qy = (np.ascontiguousarray(v, dtype=fd).copy() for v in (b, qx, qy))
    m = np.ascontiguousarray(m, dtype=md).copy()
    cum = _f64(cumulative).copy()
    w = _f64(weights)
    dep = np.zeros((ny, nx), dtype=md)
    rx = np.zeros((ny, nx), dtype=fd)
    ry = np.zeros((ny, nx), dtype=fd)
    f = None if forcing is None else _i32(forcing)
    p = DepositionParams(dx, dy, dt, settling_rate, speed_threshold, density, gravity, cohesion,
                         tan_stop, int(bool(settle_subcell_tail)))
    out = np.zeros(3)
    fn = lib().orc_deposition32 if fp32 else lib().orc_deposition
    fn(ny, nx, ctypes.byref(p), None if f is None else _ptr(f), _ptr(b), _ptr(m), _ptr(qx),
       _ptr(qy), _ptr(dep), _ptr(rx), _ptr(ry), _ptr(cum), _ptr(w), _ptr(out))
    return b, m, qx, qy, dep, rx, ry, cum, out


def surface_sample(b, m, rows_cols):
    """bilinear_surface (bulk_state_authority.py:196-228) of b + m at (row, col) pairs."""
    b, m = _f64(b), _f64(m)
    q = _f64(rows_cols).reshape(-1, 2)
    out = np.zeros(q.shape[0])
    lib().orc_surface_sample(b.shape[0], b.shape[1], _ptr(b), _ptr(m), _ptr(q), q.shape[0], _ptr(out))
    return out


def reservoirs(z, b, m, qx, qy, w, density):
    z, b, m, qx, qy, w = (_f64(v) for v in (z, b, m, qx, qy, w))
    out = np.zeros(4)
    lib().orc_reservoirs(b.size, _ptr(z), _ptr(b), _ptr(m), _ptr(qx), _ptr(qy), _ptr(w),
                         density, _ptr(out))
    return out


def relax_legacy(h, d0, d1, crit_slope, max_iterations=20000, tolerance=1.0e-8,
                 dtype=np.float64):
    """``relax_critical_slope`` on a legacy-layout array; returns (h, iterations, converged, last)."""
    h = np.ascontiguousarray(h, dtype=dtype).copy()
    n0, n1 = h.shape
    conv = _I(0)
    last = _D(0.0)
    fn = lib().orc_relax_legacy_f64 if dtype == np.float64 else lib().orc_relax_legacy_f32
    it = fn(n0, n1, d0, d1, float(crit_slope), int(max_iterations), float(tolerance),
            _ptr(h), ctypes.byref(conv), ctypes.byref(last))
    return h, int(it), bool(conv.value), float(last.value)


def relax_yx(h_yx, dx, dy, crit_slope, max_iterations=20000, tolerance=1.0e-8,
             dtype=np.float64):
    """``MinimumSlopeAdapter`` closed-boundary semantics on canonical ``H[y,x]``."""
    legacy = np.ascontiguousarray(np.asarray(h_yx).T)
    out, it, conv, last = relax_legacy(legacy, dx, dy, crit_slope, max_iterations, tolerance,
                                       dtype)
    return np.ascontiguousarray(out.T), it, conv, last


def frontier_relax(height, weights, reached, a_i, a_j, b_i, b_j, critical_difference,
                   tile_size, tiles_x):
    height = _f64(height).copy()
    reached = _i32(reached).copy()
    weights = _f64(weights)
    rows, cols = height.shape
    a_i, a_j, b_i, b_j = (_i32(v) for v in (a_i, a_j, b_i, b_j))
    tiles_y = (rows + tile_size - 1) // tile_size
    flags = np.zeros(tiles_y * tiles_x, dtype=np.int32)
    moved = lib().orc_frontier_relax(rows, cols, tile_size, tiles_x, _ptr(height), _ptr(weights),
                                     _ptr(reached), _ptr(a_i), _ptr(a_j), _ptr(b_i), _ptr(b_j),
                                     a_i.size, critical_difference, _ptr(flags))
    return height, reached, int(moved), flags


def frontier_scan(height, reached, tile_ids, a_i, a_j, b_i, b_j, critical_difference, tolerance,
                  tile_count):
    height, reached = _f64(height), _i32(reached)
    tile_ids, a_i, a_j, b_i, b_j = (_i32(v) for v in (tile_ids, a_i, a_j, b_i, b_j))
    flags = np.zeros(tile_count, dtype=np.int32)
    lib().orc_frontier_scan(height.shape[1], _ptr(height), _ptr(reached), _ptr(tile_ids),
                            _ptr(a_i), _ptr(a_j), _ptr(b_i), _ptr(b_j), a_i.size,
                            critical_difference, tolerance, _ptr(flags))
    return flags


def edge_batch(shape, tile_size, active_tiles, di, dj, phase):
    """CompactTileFrontier.edge_batch (sparse_tile_slope.py:116-230): the edges
    (a, a + (di, dj)) owned by every cell a of the active tiles whose
    coordinate along the direction (row if di else column) has parity
    `phase`.  Returns (owner tile, a_i, a_j, b_i, b_j)."""
    rows, cols = shape
    tiles_x = (cols + tile_size - 1) // tile_size
    tiles = np.unique(np.asarray(active_tiles, dtype=np.int64))
    li, lj = np.indices((tile_size, tile_size))
    li, lj = li.ravel(), lj.ravel()
    a_i = ((tiles // tiles_x)[:, None] * tile_size + li[None, :]).ravel()
    a_j = ((tiles % tiles_x)[:, None] * tile_size + lj[None, :]).ravel()
    owner = np.repeat(tiles, li.size)
    b_i, b_j = a_i + di, a_j + dj
    coord = a_i if di else a_j
    keep = ((a_i < rows) & (a_j < cols) & (b_i >= 0) & (b_i < rows) & (b_j >= 0) & (b_j < cols)
            & ((coord & 1) == phase))
    return tuple(v[keep].astype(np.int32) for v in (owner, a_i, a_j, b_i, b_j))


def frontier_rounds(height, weights, reached, active_tiles, *, tile_size, neighbors, critical_tangent,
                    tolerance, max_rounds):
    """The residual projection loop of V2PhysicsCore._advance_gpu_residual_frontier
    (pkg/runtime/v2_physics_core.py:1172-1254) over the oracle's phase
    functions: per round the relax phases of every direction and parity over
    the active tiles (tiles a phase moves join for the later phases), then
    the scan of the round's tiles; its unstable owner tiles are the next
    active set.  `neighbors` is a list of (di, dj, distance).

    Returns (height, reached, rounds, active_tiles, moved_edges, cleared)."""
    height = _f64(height).copy()
    reached = _i32(reached).copy()
    rows, cols = height.shape
    tiles_x = (cols + tile_size - 1) // tile_size
    tile_count = ((rows + tile_size - 1) // tile_size) * tiles_x
    active = np.unique(np.asarray(active_tiles, dtype=np.int32))
    rounds, moved_total, cleared = 0, 0, False
    for _ in range(max_rounds):
        if active.size == 0:
            break
        rounds += 1
        phase_tiles = active
        for di, dj, distance in neighbors:
            for phase in (0, 1):
                _, a_i, a_j, b_i, b_j = edge_batch((rows, cols), tile_size, phase_tiles, di, dj, phase)
                if a_i.size == 0:
                    continue
                height, reached, moved, flags = frontier_relax(
                    height, weights, reached, a_i, a_j, b_i, b_j, critical_tangent * distance,
                    tile_size, tiles_x)
                moved_total += moved
                triggered = np.flatnonzero(flags).astype(np.int32)
                if triggered.size:
                    phase_tiles = np.union1d(phase_tiles, triggered).astype(np.int32)
        owners = np.zeros(tile_count, dtype=np.int32)
        for di, dj, distance in neighbors:
            for phase in (0, 1):
                owner, a_i, a_j, b_i, b_j = edge_batch((rows, cols), tile_size, phase_tiles, di, dj, phase)
                if a_i.size == 0:
                    continue
                owners |= frontier_scan(height, reached, owner, a_i, a_j, b_i, b_j,
                                        critical_tangent * distance, tolerance, tile_count)
        active = np.flatnonzero(owners).astype(np.int32)
        if active.size == 0:
            reached[:] = 0
            cleared = True
            break
    return height, reached, rounds, active, moved_total, cleared


def persistence(initial, resting, mobile, qx, qy, weights, slope, unstable, latch, activation_count,
                r2m, m2r, *, tile_size, start_rad, stop_rad, dry, activity_speed):
    """reduce_persistence (gpu_large_avalanche.py:378-451): (diag_i[4], diag_f[14], tile_flags)."""
    ny, nx = np.shape(resting)
    f = [_f64(v) for v in (initial, resting, mobile, qx, qy, weights, slope)]
    ints = [_i32(v) for v in (unstable, latch, activation_count)]
    r2m, m2r = _f64(r2m), _f64(m2r)
    tiles_x = (nx + tile_size - 1) // tile_size
    flags = np.zeros(((ny + tile_size - 1) // tile_size) * tiles_x, dtype=np.int32)
    di = np.zeros(4, dtype=np.int32)
    df = np.zeros(14)
    lib().orc_persistence(ny, nx, tile_size, tiles_x, *(_ptr(v) for v in f), *(_ptr(v) for v in ints),
                          _ptr(r2m), _ptr(m2r), start_rad, stop_rad, dry, activity_speed, _ptr(di),
                          _ptr(df), _ptr(flags))
    return di, df, flags


def avalanche_initialize(z, resting, mobile, qx, qy, weights, owned_baseline, export_cum, latch,
                         *, dx, dy, tan_start, tan_stop, density, gravity, cohesion,
                         mobilization_depth, dry, activity_speed, release):
    z, resting, mobile, qx, qy, weights, owned_baseline, export_cum = (
        _f64(v) for v in (z, resting, mobile, qx, qy, weights, owned_baseline, export_cum))
    ny, nx = resting.shape
    latch = _i32(latch).copy()
    unstable = np.zeros((ny, nx), np.int32)
    slope, gx, gy, sev = (np.zeros((ny, nx)) for _ in range(4))
    p = AvalancheParams(dx, dy, tan_start, tan_stop, density, gravity, cohesion,
                        mobilization_depth, dry, activity_speed, int(bool(release)))
    lib().orc_avalanche_initialize(ny, nx, ctypes.byref(p), _ptr(z), _ptr(resting), _ptr(mobile),
                                   _ptr(qx), _ptr(qy), _ptr(weights), _ptr(owned_baseline),
                                   _ptr(export_cum), _ptr(unstable), _ptr(latch), _ptr(slope),
                                   _ptr(gx), _ptr(gy), _ptr(sev))
    return {"unstable": unstable, "latch": latch, "slope": slope, "grad_x": gx, "grad_y": gy,
            "severity": sev}


def label(mask, connectivity=8) -> np.ndarray:
    mask = _i32(mask)
    ny, nx = mask.shape
    out = np.empty((ny, nx), np.int32)
    lib().orc_label(ny, nx, int(connectivity), _ptr(mask), _ptr(out))
    return out


def components(labels, z, resting, latch, weights, slope, severity, previous, mobilization_depth,
               start_rad, stop_rad):
    labels, latch, previous = _i32(labels), _i32(latch), _i32(previous)
    z, resting, weights, slope, severity = (_f64(v) for v in (z, resting, weights, slope, severity))
    ny, nx = labels.shape
    di = np.zeros(8, np.int32)
    df = np.zeros(8)
    lib().orc_components(ny, nx, _ptr(labels), _ptr(z), _ptr(resting), _ptr(latch), _ptr(weights),
                         _ptr(slope), _ptr(severity), _ptr(previous), mobilization_depth,
                         start_rad, stop_rad, _ptr(di), _ptr(df))
    return di, df


def mobilize(labels, winner, z, resting, mobile, weights, latch, owned_surface, owned_baseline,
             export_cum, severity, activation_count, first_time, last_time, r2m, activation_time,
             mobilization_depth, dry):
    labels = _i32(labels)
    z, weights, export_cum, severity = (_f64(v) for v in (z, weights, export_cum, severity))
    resting, mobile, owned_surface, owned_baseline, first_time, last_time, r2m = (
        _f64(v).copy() for v in (resting, mobile, owned_surface, owned_baseline, first_time,
                                 last_time, r2m))
    latch, activation_count = _i32(latch).copy(), _i32(activation_count).copy()
    diag = np.zeros(16)
    lib().orc_mobilize(labels.size, _ptr(labels), int(winner), _ptr(z), _ptr(resting),
                       _ptr(mobile), _ptr(weights), _ptr(latch), _ptr(owned_surface),
                       _ptr(owned_baseline), _ptr(export_cum), _ptr(severity),
                       _ptr(activation_count), _ptr(first_time), _ptr(last_time), _ptr(r2m),
                       activation_time, mobilization_depth, dry, _ptr(diag))
    return {"resting": resting, "mobile": mobile, "latch": latch, "owned_surface": owned_surface,
            "owned_baseline": owned_baseline, "activation_count": activation_count,
            "first_time": first_time, "last_time": last_time, "r2m": r2m, "diag": diag}


def energy_terms(h, qx, qy, b, dx, dy, g=9.81, K=0.45, dry=1.0e-10):
    """``MobileV2ReferenceSolver.energy_terms`` (mobile_v2_reference.py:110-127)."""
    h = np.asarray(h)
    q = np.stack((qx, qy), axis=-1)
    area = dx * dy
    speed2 = np.divide(np.sum(q * q, axis=-1), h * h, out=np.zeros_like(h), where=h > dry)
    kinetic = float(0.5 * np.sum(h * speed2, dtype=np.float64) * area)
    grav = float(g * np.sum(h * b, dtype=np.float64) * area)
    internal = float(0.5 * K * g * np.sum(h * h, dtype=np.float64) * area)
    return kinetic, grav, internal, kinetic + grav + internal


if os.environ.get("MV2_ORACLE_AUTOBUILD", "1") == "1" and not _LIB_PATH.exists():
    try:
        build()
    except Exception:  # pragma: no cover - reported on first lib() call
        pass


# ---- FP32-storage exchange amounts (restating mv2_internal.hpp
#      entrain_quanta_f32 / deposit_quanta_f32; the contract they serve is
#      bulk_state_authority.py:104-150 with H_free = b_eff + mobile invariant
#      per cell, here with b_eff float32 and mobile whole quanta of 2^-24 m) ----
QUANTUM = 2.0 ** -24


def f32_spacing(b) -> float:
    """Spacing of float32 numbers in the binade of b."""
    e = (int(np.float32(b).view(np.uint32)) >> 23) & 0xFF
    return 2.0 ** -149 if e == 0 else 2.0 ** (e - 150)


def _f32_exact(r: float) -> bool:
    return float(np.float32(r)) == r


def entrain_quanta_f32(b, depth, room=2 ** 31 - 1) -> int:
    """Quanta moved Resting -> Mobile for a donor-limited depth (metres)."""
    b = np.float32(b)
    g = max(QUANTUM, f32_spacing(b))
    d = np.floor(depth / g) * g
    cap = np.floor(room * QUANTUM / g) * g
    d = min(d, cap)
    if not d > 0.0 or not _f32_exact(float(b) - d):
        return 0
    return int(d / QUANTUM)


def deposit_quanta_f32(b, quanta):
    """(quanta moved Mobile -> Resting, new float32 b_eff) for a request of
    `quanta` whole quanta (already limited by the mobile layer)."""
    b = np.float32(b)
    if quanta <= 0:
        return 0, b
    d0 = quanta * QUANTUM
    r = float(b) + d0
    if _f32_exact(r):
        return int(quanta), np.float32(r)
    rf = np.float32(r)
    if float(rf) > r:
        rf = np.nextafter(rf, np.float32(-np.inf))
    d = float(rf) - float(b)
    if d > 0.0 and np.floor(d / QUANTUM) == d / QUANTUM:
        return int(d / QUANTUM), rf
    if np.floor(float(b) / QUANTUM) == float(b) / QUANTUM:
        return 0, b
    sb = f32_spacing(b)
    e = (int(b.view(np.uint32)) >> 23) & 0xFF
    top = 2.0 ** (e - 126) - sb
    k = np.floor((top - float(b)) / QUANTUM)
    if k >= 1.0 and _f32_exact(float(b) + k * QUANTUM):
        return int(k), np.float32(float(b) + k * QUANTUM)
    return int(quanta), np.float32(r)


def failure_activation(z, b, m, weights, bbox_yx, activated, affected, *, fp32=False):
    """FailureZone activation commit (DeviceFailureZoneBridge.execute after the
    CAD geometry, gpu_failure_bridge.py:114-136) on full fields: the
    entrainment contract over the patch (FP64: orc_entrain; FP32 storage:
    whole quanta by entrain_quanta_f32, m given in quanta), then the forcing
    cells -- wet (patch mobile + activated > 0) cells that are, or are a
    4-neighbour inside the patch of, an affected wet cell -- as ascending flat
    indices.  Returns (b, m, volume, forcing)."""
    row0, row1, col0, col1 = (int(v) for v in bbox_yx)
    ny, nx = np.shape(b)
    act = np.asarray(activated, dtype=np.float64)
    aff = np.asarray(affected, dtype=bool)
    rows, cols = np.indices(act.shape)
    flat = ((rows + row0) * nx + (cols + col0)).ravel()
    if not fp32:
        mob_before = np.asarray(m, dtype=np.float64)[row0:row1, col0:col1]
        b, m, volume = entrain(z, b, m, weights, flat, act.ravel())
        b, m = b.reshape(ny, nx), m.reshape(ny, nx)
    else:
        b = np.asarray(b, dtype=np.float32).copy().ravel()
        m = np.asarray(m, dtype=np.int32).copy().ravel()
        zf = np.asarray(z, dtype=np.float32).ravel()
        w = np.asarray(weights, dtype=np.float64).ravel()
        mob_before = m.reshape(ny, nx)[row0:row1, col0:col1].astype(np.float64) * QUANTUM
        volume = 0.0
        for i, req in zip(flat, act.ravel()):
            depth = min(max(req, 0.0), max(float(b[i]) - float(zf[i]), 0.0))
            dq = entrain_quanta_f32(b[i], depth, 2 ** 31 - 1 - int(m[i]))
            b[i] = np.float32(float(b[i]) - dq * QUANTUM)
            m[i] += dq
            volume += dq * QUANTUM * w[i]
        b, m = b.reshape(ny, nx), m.reshape(ny, nx)
    wet = mob_before + act > 0.0
    contact = aff & wet
    near = contact.copy()
    near[1:, :] |= contact[:-1, :]
    near[:-1, :] |= contact[1:, :]
    near[:, 1:] |= contact[:, :-1]
    near[:, :-1] |= contact[:, 1:]
    near &= wet
    r, c = np.nonzero(near)
    forcing = ((r + row0) * nx + (c + col0)).astype(np.int32)
    return b, m, float(volume), forcing
