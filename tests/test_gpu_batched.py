"""Batched Mobile V2: every instance of an uneven batch must equal the same
terrain stepped alone by the single-terrain FP32 path, bit for bit, for every
tile shape, with and without the captured HIP graph."""

from __future__ import annotations

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

QUANTUM = 2.0 ** -24
SIZES = [(40, 50), (64, 64), (33, 97), (130, 70), (96, 128), (17, 23)]


def _terrains():
    from isaac_bulk_hip.workloads import gaussian_pile, mobile_patch

    out = []
    for k, (ny, nx) in enumerate(SIZES):
        rng = np.random.default_rng(100 + k)
        b = gaussian_pile(ny, nx, 0.05, rng, bumps=4, sigma=(0.3, 0.8))
        h = mobile_patch(ny, nx, rng, size=min(16, ny, nx), hmax=0.2)
        out.append((b.astype(np.float32), h))
    # a flat, dry terrain takes the static exit
    out.append((np.full((24, 30), 0.5, np.float32), np.zeros((24, 30))))
    return out


def _single_reference(terrains, dt, frames, config):
    from isaac_bulk_hip import HipMobileV2ReferenceSolver, MobileV2State

    solver = HipMobileV2ReferenceSolver(config, precision="fp32")
    results = []
    for b, h in terrains:
        q = np.rint(h / QUANTUM) * QUANTUM
        state = MobileV2State(b.astype(np.float64), q, np.zeros(h.shape + (2,)))
        subs = []
        for _ in range(frames):
            step = solver.step(state, dt)
            state = step.state
            subs.append(step.substeps)
        results.append((state, subs))
    return results


@pytest.mark.parametrize("shape", [0, 1, 2, 3])
@pytest.mark.parametrize("use_graph", [True, False])
def test_batch_matches_single_terrain_bit_exact(shape, use_graph):
    from isaac_bulk_hip import MobileV2Config
    from isaac_bulk_hip.batched import BatchedMobileV2
    from isaac_bulk_hip.relaxation import _runtime

    config = MobileV2Config(cfl=0.18)
    terrains = _terrains()
    dt, frames = 0.25, 2
    ref = _single_reference(terrains, dt, frames, config)
    batch = BatchedMobileV2(terrains, 0.05, runtime=_runtime("cuda:0"), config=config, use_graph=use_graph,
                            tile_shape=shape)
    recs = [batch.step(dt) for _ in range(frames)]
    assert max(r.substeps for r in recs[0]) > 8           # more than one graph replay
    for i, (state, subs) in enumerate(ref):
        b, h, qx, qy = batch.fields(i)
        assert [recs[f][i].substeps for f in range(frames)] == subs, i
        np.testing.assert_array_equal(np.rint(h / QUANTUM), np.rint(state.h_m / QUANTUM))
        np.testing.assert_array_equal(qx.astype(np.float64), state.q_m2_s[..., 0])
        np.testing.assert_array_equal(qy.astype(np.float64), state.q_m2_s[..., 1])
        for f in range(frames):
            assert recs[f][i].mass_quanta_before == recs[f][i].mass_quanta_after
    assert recs[0][-1].static_exit and recs[0][-1].substeps == 0


def test_graph_replays_agree_with_eager_launches():
    """Replays of the captured graph, with eager work and syncs between them,
    give the same fields as plain launches."""
    import torch
    from isaac_bulk_hip import MobileV2Config
    from isaac_bulk_hip.batched import BatchedMobileV2
    from isaac_bulk_hip.relaxation import _runtime

    config = MobileV2Config(cfl=0.18)
    terrains = _terrains()[:4]
    a = BatchedMobileV2(terrains, 0.05, runtime=_runtime("cuda:0"), config=config, use_graph=True)
    for frame in range(4):
        a.step(0.1)
        torch.cuda.synchronize()
        _ = (a.h.sum() + 1).item()             # eager work between replays
    b = BatchedMobileV2(terrains, 0.05, runtime=_runtime("cuda:0"), config=config, use_graph=False)
    for frame in range(4):
        b.step(0.1)
    for i in range(len(terrains)):
        for x, y in zip(a.fields(i), b.fields(i)):
            np.testing.assert_array_equal(x, y)


def test_autotune_picks_a_shape_and_keeps_the_state():
    from isaac_bulk_hip import MobileV2Config
    from isaac_bulk_hip.batched import TILE_SHAPES, BatchedMobileV2
    from isaac_bulk_hip.relaxation import _runtime

    terrains = _terrains()[:3]
    batch = BatchedMobileV2(terrains, 0.05, runtime=_runtime("cuda:0"), config=MobileV2Config(cfl=0.18))
    before = [f[1].copy() for f in (batch.fields(i) for i in range(3))]
    result = batch.autotune(0.05, frames=1)
    assert result["best"] in TILE_SHAPES and set(result["ms"]) == set(TILE_SHAPES)
    for i in range(3):
        np.testing.assert_array_equal(batch.fields(i)[1], before[i])


def _relax_terrains():
    rng = np.random.default_rng(7)
    out = []
    for ny, nx in [(40, 37), (65, 130), (97, 33), (2, 2), (128, 96), (17, 50)]:
        b = rng.uniform(0.0, 1.5, size=(ny, nx)).astype(np.float32)
        out.append((b, np.zeros((ny, nx))))
    return out


def test_batched_relax_matches_single_field_relaxation():
    """Every instance relaxed in the batch equals the same field relaxed alone
    (FP32), bit for bit, with the same iteration count."""
    import torch
    from isaac_bulk_hip import relax_device
    from isaac_bulk_hip.batched import BatchedMobileV2
    from isaac_bulk_hip.relaxation import _runtime

    terrains = _relax_terrains()
    batch = BatchedMobileV2(terrains, 0.05, runtime=_runtime("cuda:0"))
    res = batch.relax(34.0, 1e-6, max_iterations=5000)
    for i, (b, _) in enumerate(terrains):
        t = torch.from_numpy(b.copy()).to("cuda:0")
        it, conv, _ = relax_device(t, 0.05, 0.05, 34.0, 5000, 1e-6)
        assert (int(res["iterations"][i]), bool(res["converged"][i])) == (it, conv), i
        np.testing.assert_array_equal(batch.fields(i)[0], t.cpu().numpy())
    # a second call starts from rest: every instance converges at once
    again = batch.relax(34.0, 1e-6)
    assert again["converged"].all() and (again["iterations"] <= 1).all()


def test_batched_entrain_contract():
    """entrain moves whole quanta, limited by the donor layer, per instance,
    keeping b_eff + mobile bit-invariant (the oracle's FP32 amount rule);
    duplicate (instance, cell) pairs are rejected."""
    from isaac_bulk_hip.batched import BatchedMobileV2
    from isaac_bulk_hip.relaxation import _runtime

    terrains = _relax_terrains()[:3]
    batch = BatchedMobileV2(terrains, 0.05, runtime=_runtime("cuda:0"))
    before = [batch.fields(i) for i in range(3)]
    inst = np.array([0, 0, 1, 2, 2], np.int32)
    cell = np.array([0, 5, 17, 3, 40], np.int32)
    depth = np.array([0.01, 5.0, 0.0123456, -1.0, 0.2])
    vol = batch.entrain(inst, cell, depth)
    expect = np.zeros(3)
    for k in range(len(inst)):
        i, c = int(inst[k]), int(cell[k])
        b0 = float(before[i][0].ravel()[c])
        d = min(max(depth[k], 0.0), max(b0, 0.0))
        h0 = int(np.rint(before[i][1].ravel()[c] / QUANTUM))
        dq = oracle.entrain_quanta_f32(np.float32(b0), d, 2 ** 31 - 1 - h0)
        b1, h1 = batch.fields(i)[0].ravel()[c], batch.fields(i)[1].ravel()[c]
        assert float(b1) == b0 - dq * QUANTUM                      # exact: H_free unchanged
        assert np.rint(h1 / QUANTUM) - h0 == dq
        expect[i] += dq * QUANTUM * 0.05 * 0.05
    np.testing.assert_allclose(vol, expect, rtol=0, atol=1e-18)
    # cells outside the lists are untouched
    assert np.array_equal(batch.fields(1)[0].ravel()[:17], before[1][0].ravel()[:17])
    with pytest.raises(ValueError, match="duplicate"):
        batch.entrain(np.array([0, 1, 0], np.int32), np.array([3, 3, 3], np.int32), np.full(3, 0.01))


@pytest.mark.parametrize("connectivity", [4, 8])
def test_batched_label_matches_oracle(connectivity):
    """Failure masks, labels and component statistics of every instance equal
    the oracle's on that terrain alone (FP32 storage read into FP64)."""
    import oracle
    from isaac_bulk_hip.batched import BatchedMobileV2
    from isaac_bulk_hip.relaxation import _runtime
    from isaac_bulk_hip.workloads import cut_face_terrain

    terrains = []
    for k, n in enumerate((96, 64, 50, 33)):
        b, h, _ = cut_face_terrain(n, seed=k, face_deg=45.0)
        terrains.append((b.astype(np.float32), h))
    batch = BatchedMobileV2(terrains, 0.05, runtime=_runtime("cuda:0"))
    labels, stats = batch.label(start_angle_deg=38.0, stop_angle_deg=30.0, density=2200.0, cohesion=0.0,
                                mobilization_depth=0.05, dry_tolerance=1e-8, connectivity=connectivity)
    labels = labels.cpu().numpy()
    some = 0
    for i, (b, h) in enumerate(terrains):
        ny, nx = b.shape
        bf, hf = batch.fields(i)[0].astype(np.float64), batch.fields(i)[1]
        z = np.zeros((ny, nx))
        o = oracle.avalanche_initialize(
            z, bf, hf, z, z, np.ones((ny, nx)), z, z, np.zeros((ny, nx), np.int32), dx=0.05, dy=0.05,
            tan_start=float(np.tan(np.deg2rad(38.0))), tan_stop=float(np.tan(np.deg2rad(30.0))),
            density=2200.0, gravity=9.81, cohesion=0.0, mobilization_depth=0.05, dry=1e-8,
            activity_speed=0.0, release=False)
        lab = oracle.label(o["unstable"], connectivity)
        off = int(batch.offset[i])
        np.testing.assert_array_equal(labels[off:off + ny * nx].reshape(ny, nx), lab)
        roots, counts = np.unique(lab[lab >= 0], return_counts=True)
        regions = len(roots)
        largest = int(counts.max()) if regions else 0
        winner = int(roots[counts == largest].min()) if regions else -1
        assert stats[i].tolist() == [regions, largest, winner, int((lab >= 0).sum())], i
        some += regions
    assert some > 0


@pytest.mark.parametrize("settle_rounds", [1, 3])
def test_batched_warm_start_matches_single_field(settle_rounds):
    """The learned warm start of a batch of equal-size terrains equals the
    single-field warm start of each terrain, bit for bit (same settle rounds)."""
    import torch
    from isaac_bulk_hip.batched import BatchedMobileV2
    from isaac_bulk_hip.relaxation import _runtime
    from isaac_bulk_hip.surrogate import warm_start_relax
    from isaac_bulk_hip.warm_start_training import load_net
    from isaac_bulk_hip.workloads import cut_face_terrain

    rt = _runtime("cuda:0")
    net, _ = load_net(rt)
    terrains = []
    for k in range(4):
        b, h, _ = cut_face_terrain(96, seed=10 + k, face_deg=45.0)
        terrains.append((b.astype(np.float32), h))
    batch = BatchedMobileV2(terrains, 0.05, runtime=rt)
    out = batch.warm_start_relax(net, 34.0, 1e-6, max_iterations=5000, settle_rounds=settle_rounds)
    assert out["relax"]["converged"].all()
    for i, (b, _) in enumerate(terrains):
        bt = torch.from_numpy(b.copy()).to("cuda:0")
        ht = torch.from_numpy(batch.fields(i)[1].astype(np.float32)).to("cuda:0")
        single = warm_start_relax(net, bt, ht, 0.05, 34.0, 1e-6, max_iterations=5000, settle_rounds=settle_rounds)
        assert single.converged
        np.testing.assert_array_equal(batch.fields(i)[0], bt.cpu().numpy())


def test_two_batches_on_one_runtime():
    """A runtime's context holds one bound batch: two batches used
    alternately on the same runtime each give exactly the fields of that batch
    used alone on its own runtime."""
    from isaac_bulk_hip import HipRuntime, MobileV2Config
    from isaac_bulk_hip.batched import BatchedMobileV2

    config = MobileV2Config(cfl=0.18)
    sets = [_terrains()[:3], _terrains()[3:6]]
    shared = HipRuntime("cuda:0")
    together = [BatchedMobileV2(t, 0.05, runtime=shared, config=config) for t in sets]
    alone = [BatchedMobileV2(t, 0.05, runtime=HipRuntime("cuda:0"), config=config) for t in sets]
    for frame in range(3):
        for group in (together, alone):
            for batch in group:
                batch.step(0.1)
                batch.relax(34.0, 1e-6, max_iterations=200)
                batch.entrain(np.array([0], np.int32), np.array([5], np.int32), np.array([0.001]))
    for a, b in zip(together, alone):
        for i in range(a.count):
            for x, y in zip(a.fields(i), b.fields(i)):
                np.testing.assert_array_equal(x, y)
    for rt in {id(b.rt): b.rt for b in together + alone}.values():
        rt.close()
