"""The batched, learned and digging configurations at their stated sizes.

C3: 256 terrains with sizes uniform over {64, 96, 128, 192, 256, 384, 512}^2,
    each with its own Gaussian pile and mobile patch, FP32, 1/60 s frames for
    2 s simulated: exact mass (quanta) per instance and frame; sampled
    instances against the FP64 oracle of that terrain alone within the
    full-size frame tolerances (L-inf <= 1e-3 m, RMSE <= 1e-5 m).
C4: the learned warm start on the 701^2 cut-face terrain: the plain
    relaxation's judge accepts it, mass is kept, and it needs at most a third
    of the plain relaxation's exact iterations.
C5: 64 terrains of 256^2 dug by a 1.5 x 0.8 m bucket at 0.02 m per frame
    (entrainment contract), learned-warm-start relaxation at 34 deg and
    failure labelling every frame: b_eff + mobile volume per terrain kept,
    every terrain at rest by the plain judge after each frame, labels of
    sampled terrains equal to the oracle's.
"""

from __future__ import annotations

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

DT = 1.0 / 60.0
Q = 2.0 ** -24


def test_c3_256_uneven_terrains_two_seconds():
    from isaac_bulk_hip import MobileV2Config
    from isaac_bulk_hip.batched import BatchedMobileV2
    from isaac_bulk_hip.workloads import batched_piles

    terrains = [(b.astype(np.float32), h) for b, h in batched_piles(256, 0.05, seed=0)]
    sizes = np.array([t[0].shape[0] for t in terrains])
    assert set(sizes.tolist()) == {64, 96, 128, 192, 256, 384, 512}
    batch = BatchedMobileV2(terrains, 0.05, config=MobileV2Config(cfl=0.18))
    # sampled instances: the first terrain of the smallest, a middle and the largest size
    sample = [int(np.flatnonzero(sizes == s)[0]) for s in (64, 192, 512)]
    seq = {}
    for i in sample:
        b, h = terrains[i]
        seq[i] = [b.astype(np.float64), np.rint(h / Q) * Q, np.zeros(h.shape), np.zeros(h.shape)]
    oracle_frames = 30
    worst = 0.0
    for f in range(120):                                   # 2 s simulated
        recs = batch.step(DT)
        assert np.all(recs.numeric_error == 0) and np.all(recs.complete == 1)
        # mass moves in whole quanta: every instance keeps its total exactly
        np.testing.assert_array_equal(recs.mass_quanta_before, recs.mass_quanta_after)
        if f < oracle_frames:
            for i in sample:
                b, h, qx, qy = seq[i]
                h, qx, qy, rec = oracle.v2_step(b, h, qx, qy, DT)
                seq[i][1:] = [h, qx, qy]
                assert abs(int(recs.substeps[i]) - rec["substeps"]) <= 1, (f, i)
                _, hd, qxd, qyd = batch.fields(i)
                for got, ref in ((hd, h), (qxd, qx), (qyd, qy)):
                    err = np.asarray(got, np.float64) - ref
                    worst = max(worst, float(np.max(np.abs(err))))
                    assert np.max(np.abs(err)) <= 1e-3, (f, i)
                    assert np.sqrt(np.mean(err * err)) <= 1e-5, (f, i)
    print(f"C3: worst L-inf over {oracle_frames} frames of 3 sampled instances {worst:.3e} m")


def test_c4_learned_warm_start_701_cut_face():
    import torch
    from isaac_bulk_hip.relaxation import _runtime, relax_device
    from isaac_bulk_hip.surrogate import warm_start_relax
    from isaac_bulk_hip.warm_start_training import load_net
    from isaac_bulk_hip.workloads import cut_face_terrain

    rt = _runtime("cuda:0")
    net, _ = load_net(rt)
    b0, _, _ = cut_face_terrain(701, 0.05, seed=0)
    base = torch.from_numpy(b0.astype(np.float32)).to("cuda:0")
    tol = 1e-6
    plain = base.clone()
    it_plain, conv_plain, _ = relax_device(plain, 0.05, 0.05, 34.0, 20000, tol, runtime=rt)
    learned = base.clone()
    res = warm_start_relax(net, learned, None, 0.05, 34.0, tol)
    exact = res.iterations_exact + res.fallback_iterations
    assert conv_plain and res.converged and exact * 3 <= it_plain, (exact, it_plain)
    it, conv, last = relax_device(learned.clone(), 0.05, 0.05, 34.0, 1, tol, runtime=rt)
    assert conv and last <= tol
    m0 = float(base.double().sum())
    assert abs(float(learned.double().sum()) - m0) <= 1e-6 * m0
    assert float((learned - plain).abs().max()) < 0.1
    print(f"C4: plain {it_plain} iterations, learned {res.iterations_exact} + {res.fallback_iterations} "
          f"fallback, envelope {res.envelope_launches} launches")


def test_c5_digging_64_terrains():
    import torch
    from isaac_bulk_hip import MobileV2Config
    from isaac_bulk_hip.batched import BatchedMobileV2
    from isaac_bulk_hip.warm_start_training import load_net
    from isaac_bulk_hip.workloads import bucket_cells, digging_sites

    n, count, frames = 256, 64, 12
    terrains, lines = digging_sites(count, n, 0.05, seed=0)
    batch = BatchedMobileV2([(b.astype(np.float32), h) for b, h in terrains], 0.05, config=MobileV2Config(cfl=0.18))
    rest = batch.relax(34.0, 1e-6, max_iterations=20000)
    assert np.all(rest["converged"])
    net, _ = load_net(batch.rt)

    def volumes():
        b = batch.b.view(count, n, n).double().sum(dim=(1, 2))
        h = batch.h.view(count, n, n).double().sum(dim=(1, 2)) * Q
        return (b + h).cpu().numpy()

    v0 = volumes()
    for f in range(frames):
        cells = [bucket_cells(line, f * 0.02, n) for line in lines]
        inst = np.concatenate([np.full(len(c), i, np.int32) for i, c in enumerate(cells)])
        batch.entrain(inst, np.concatenate(cells), np.full(inst.size, 0.01))
        batch.step(DT)
        ws = batch.warm_start_relax(net, 34.0, 1e-6, max_iterations=20000)
        assert np.all(ws["relax"]["converged"])
        labels, stats = batch.label(start_angle_deg=38.0, stop_angle_deg=30.0, density=1370.0, cohesion=0.0,
                                    mobilization_depth=0.05, dry_tolerance=1e-8, connectivity=8)
    # the plain judge finds every terrain at rest
    judge = batch.relax(34.0, 1e-6, max_iterations=1, field=batch.b.clone())
    assert np.all(judge["converged"]), judge["max_excess_last"].max()
    # b_eff + mobile per terrain: exact transfers except the relaxation's float32 rounding
    v1 = volumes()
    assert np.max(np.abs(v1 - v0) / v0) <= 1e-6
    # labels of sampled terrains against the oracle's mask and labels of that terrain
    lab = labels.view(count, n, n).cpu().numpy() if hasattr(labels, "view") else np.asarray(labels).reshape(count, n, n)
    for i in (0, 21, 63):
        b, h, qx, qy = batch.fields(i)
        z = np.zeros((n, n))
        m = oracle.avalanche_initialize(
            z, b.astype(np.float64), h, qx.astype(np.float64), qy.astype(np.float64),
            oracle.vertex_weights_fast(n, n, 0.05, 0.05), z, z, np.zeros((n, n), np.int32), dx=0.05, dy=0.05,
            tan_start=float(np.tan(np.deg2rad(38.0))), tan_stop=float(np.tan(np.deg2rad(30.0))), density=1370.0,
            gravity=9.81, cohesion=0.0, mobilization_depth=0.05, dry=1e-8, activity_speed=0.05, release=True)
        np.testing.assert_array_equal(lab[i], oracle.label(m["unstable"], 8), err_msg=str(i))
        assert int(stats[i][3]) == int(m["unstable"].sum())
    torch.cuda.synchronize()
