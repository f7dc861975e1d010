"""Batched Mobile V2: every instance of an uneven batch must equal the same
terrain stepped alone by the single-terrain FP32 path, bit for bit, for every
tile shape, with and without the captured HIP graph."""

from __future__ import annotations

import numpy as np
import pytest

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
