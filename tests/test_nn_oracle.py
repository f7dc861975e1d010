"""Pin the NumPy network oracle: finite differences of its float64 backward
pass and a library convolution for its forward layers."""

from __future__ import annotations

import numpy as np
import pytest

import nn_oracle


def test_oracle_backward_matches_finite_differences():
    rng = np.random.default_rng(0)
    params = {k: v.astype(np.float64) for k, v in nn_oracle.init_params(1).items()}
    params["w5"] *= 10.0
    b = rng.uniform(0.5, 1.5, size=(1, 20, 20))
    h = np.where(rng.random((1, 20, 20)) < 0.2, rng.uniform(0, 0.1, (1, 20, 20)), 0.0)
    feat = nn_oracle.features(b, h, 0.05, 0.05).astype(np.float64)
    go = rng.standard_normal((1, 20, 20))
    out, acts = nn_oracle.forward(params, feat, bf16_mode=False)
    grads = nn_oracle.backward(params, acts, go, bf16_mode=False)
    eps = 1e-6
    for name, idx in (("w1", (4, 3, 1)), ("w1", (0, 5, 0)), ("w2", (2, 7, 9)), ("w3", (8, 1, 30)),
                      ("w4", (5, 31, 2)), ("w5", (4, 0, 11)), ("b2", (3,)), ("b5", (0,))):
        p2 = {k: v.copy() for k, v in params.items()}
        p2[name][idx] += eps
        o2, _ = nn_oracle.forward(p2, feat, bf16_mode=False)
        p3 = {k: v.copy() for k, v in params.items()}
        p3[name][idx] -= eps
        o3, _ = nn_oracle.forward(p3, feat, bf16_mode=False)
        fd = float(np.sum(go * (o2 - o3)) / (2 * eps))
        assert abs(fd - grads[name][idx]) <= 1e-5 * max(1.0, abs(fd)), (name, idx, fd, grads[name][idx])


def test_zero_sum_b_taps_make_the_net_shift_invariant():
    rng = np.random.default_rng(3)
    params = nn_oracle.init_params(2)
    b = rng.uniform(0.5, 1.5, size=(1, 24, 24))
    f1 = nn_oracle.features(b, None, 0.05, 0.05)
    f2 = nn_oracle.features(b + 3.25, None, 0.05, 0.05)
    o1, _ = nn_oracle.forward(params, f1, bf16_mode=False)
    o2, _ = nn_oracle.forward(params, f2, bf16_mode=False)
    # invariant up to the float32 rounding of the shifted heights
    assert np.max(np.abs(o1 - o2)) < 1e-5


def _excess_brute(z, crit):
    """Edge-by-edge restatement with the relaxation's legacy edge table."""
    B, H, W = z.shape
    total = 0.0
    for n in range(B):
        for r in range(H):
            for c in range(W):
                edges = []
                if c + 1 < W:
                    edges.append((0, (r, c), (r, c + 1)))
                if r + 1 < H:
                    edges.append((1, (r, c), (r + 1, c)))
                if r + 1 < H and c + 1 < W:
                    edges.append((2, (r, c), (r + 1, c + 1)))
                    edges.append((3, (r + 1, c), (r, c + 1)))
                for d, a, b in edges:
                    ex = max(abs(z[n][a] - z[n][b]) - crit[d], 0.0)
                    total += ex * ex
    return total


def test_excess_loss_matches_edge_enumeration_and_finite_differences():
    rng = np.random.default_rng(5)
    pre = rng.uniform(0.0, 0.2, size=(2, 7, 9))
    pred = rng.uniform(-0.05, 0.05, size=(2, 7, 9))
    crit = np.array([0.02, 0.03, 0.04, 0.045])
    loss, grad = nn_oracle.excess_loss(pred, pre, crit, 3.0)
    n = pre.size
    assert abs(loss - 3.0 * _excess_brute(pre + pred, crit) / n) < 1e-14
    eps = 1e-7
    for idx in ((0, 0, 0), (1, 3, 4), (0, 6, 8), (1, 0, 8), (0, 5, 1)):
        p2 = pred.copy(); p2[idx] += eps
        p3 = pred.copy(); p3[idx] -= eps
        fd = (nn_oracle.excess_loss(p2, pre, crit, 3.0)[0] - nn_oracle.excess_loss(p3, pre, crit, 3.0)[0]) / (2 * eps)
        assert abs(fd - grad[idx]) <= 1e-6 * max(1.0, abs(fd)), (idx, fd, grad[idx])


def test_excess_loss_is_zero_on_subcritical_surfaces():
    dx = 0.05
    crit = np.tan(np.deg2rad(34.0)) * np.array([dx, dx, np.hypot(dx, dx), np.hypot(dx, dx)])
    x = np.arange(16)[None, None, :] * dx
    plane = np.tan(np.deg2rad(30.0)) * x + np.zeros((1, 12, 1))
    loss, grad = nn_oracle.excess_loss(np.zeros_like(plane), plane, crit, 1.0)
    assert loss == 0.0 and not grad.any()
    steep = np.tan(np.deg2rad(40.0)) * x + np.zeros((1, 12, 1))
    loss, grad = nn_oracle.excess_loss(np.zeros_like(steep), steep, crit, 1.0)
    assert loss > 0.0 and grad.any()


def _envelope_dijkstra(z, crit):
    """min over q of z(q) + shortest path length q -> p (float64, brute force)."""
    import heapq

    H, W = z.shape
    best = np.array(z, dtype=np.float64)
    heap = [(best[r, c], r, c) for r in range(H) for c in range(W)]
    heapq.heapify(heap)
    done = np.zeros((H, W), bool)
    while heap:
        v, r, c = heapq.heappop(heap)
        if done[r, c]:
            continue
        done[r, c] = True
        for dr, dc, d in nn_oracle.NEIGHBOURS:
            rr, cc = r + dr, c + dc
            if 0 <= rr < H and 0 <= cc < W and v + crit[d] < best[rr, cc]:
                best[rr, cc] = v + crit[d]
                heapq.heappush(heap, (best[rr, cc], rr, cc))
    return best


def test_envelope_is_the_maximal_lipschitz_minorant():
    rng = np.random.default_rng(9)
    z = rng.uniform(0.0, 0.4, size=(9, 11))
    crit = np.array([0.03, 0.035, 0.05, 0.045])
    env = nn_oracle.envelope(z, crit)
    np.testing.assert_allclose(env, _envelope_dijkstra(z, crit), rtol=0, atol=1e-15)
    assert np.all(env <= z)
    # every lowered cell is held down by a neighbour at exactly the limit
    for r in range(9):
        for c in range(11):
            if env[r, c] < z[r, c]:
                tight = [env[r + dr, c + dc] + crit[d] for dr, dc, d in nn_oracle.NEIGHBOURS
                         if 0 <= r + dr < 9 and 0 <= c + dc < 11]
                assert abs(min(tight) - env[r, c]) < 1e-15
    # the excess of the envelope is zero up to the rounding of z_b + c - z_b
    loss, _ = nn_oracle.excess_loss(np.zeros((1, 9, 11)), env[None], crit, 1.0)
    assert loss < 1e-30


def test_commit_surface_conserves_the_moved_mass():
    rng = np.random.default_rng(4)
    b = rng.uniform(0.5, 1.0, size=(20, 20)).astype(np.float32)
    z = b.copy()
    z[5:9, 5:9] -= 0.05
    z[12:15, 3:8] += 0.02
    out = nn_oracle.commit_surface(b, z, 1e-5)
    assert abs(float(out.astype(np.float64).sum() - b.astype(np.float64).sum())) < 1e-5
    assert np.array_equal(out[0:4, :], b[0:4, :])


def test_commit_surface_keeps_mass_when_the_floor_clamps():
    """Cells pushed below zero are held at zero and the shift makes up for it:
    the moved cells keep their total (to float rounding)."""
    rng = np.random.default_rng(5)
    b = rng.uniform(0.0, 0.3, size=(32, 32)).astype(np.float32)
    z = b - rng.uniform(0.0, 0.4, size=b.shape).astype(np.float32) * (rng.random(b.shape) < 0.5)
    z[:4] += 0.5                                  # some mass moves up elsewhere
    out = nn_oracle.commit_surface(b, z, 1e-5)
    assert out.min() >= 0.0 and (out == 0.0).any()
    moved = np.abs(z - b) > 1e-5
    total_in = float(b[moved].astype(np.float64).sum())
    total_out = float(out[moved].astype(np.float64).sum())
    assert abs(total_out - total_in) <= 1e-6 * total_in
    # the plain mean shift alone would have created mass
    plain = nn_oracle.commit_surface(b, z, 1e-5, corrections=0)
    assert float(plain[moved].astype(np.float64).sum()) - total_in > 1e-3


def _trench_case():
    from isaac_bulk_hip.relaxation import legacy_critical_differences

    crit = legacy_critical_differences(0.05, 0.05, 34.0).astype(np.float32)
    rng = np.random.default_rng(3)
    yy, xx = np.mgrid[0:64, 0:80]
    cone = (1.5 - 0.05 * np.hypot(yy - 32, xx - 40) * np.tan(np.deg2rad(30))).astype(np.float32)
    b = nn_oracle.envelope(cone, crit)
    b[20:40, 30:34] -= 0.3                                 # a dug trench with over-steep walls
    pred = (nn_oracle.envelope(b, crit) - b) * 0.9 + rng.normal(0, 0.003, b.shape).astype(np.float32)
    z = nn_oracle.envelope((b + pred).astype(np.float32), crit)
    return b, z, crit


def _max_excess(f, crit):
    f = f.astype(np.float64)
    worst = 0.0
    for dr, dc, d in nn_oracle.NEIGHBOURS:
        if dr < 0 or (dr == 0 and dc < 0):
            continue
        r1, c1 = f.shape[0] - dr, f.shape[1] - abs(dc)
        a = f[0:r1, max(0, -dc):max(0, -dc) + c1]
        o = f[dr:dr + r1, max(0, dc):max(0, dc) + c1]
        worst = max(worst, float((np.abs(a - o) - crit[d]).max()))
    return worst


def test_masked_commit_balances_the_whole_field():
    """Outside the mask the field takes z exactly; the shift over the mask pays
    for every change, so the total is kept to float32 rounding of the shift;
    with nothing moved outside the mask it reduces to commit_surface."""
    rng = np.random.default_rng(11)
    b = rng.uniform(0.5, 1.0, (40, 50)).astype(np.float32)
    z = (b - np.where(rng.random(b.shape) < 0.2, rng.uniform(0, 0.01, b.shape), 0.0)).astype(np.float32)
    mask = rng.random(b.shape) < 0.5
    out = nn_oracle.commit_masked(b, z, mask)
    np.testing.assert_array_equal(out[~mask], z[~mask])
    assert abs(out.astype(np.float64).sum() - b.astype(np.float64).sum()) <= 2e-7 * b.size
    moved = np.abs(z - b) > np.float32(1e-5)
    np.testing.assert_array_equal(nn_oracle.commit_masked(b, np.where(moved, z, b), moved),
                                  nn_oracle.commit_surface(b, np.where(moved, z, b), 1e-5))


def test_settle_rounds_remove_the_rim_step_and_keep_mass():
    """The uniform shift of the first commit leaves a step at the rim of the
    moved cells; each settle round cuts it by orders of magnitude (closed loop:
    the envelope takes the step, the masked projection returns the mass)."""
    b, z, crit = _trench_case()
    ex = [_max_excess(nn_oracle.commit_settled(b, z, crit, 1e-5, r), crit) for r in (1, 2, 3)]
    assert ex[0] > 1e-3 and ex[1] < 1e-2 * ex[0] and ex[2] < 1e-6
    np.testing.assert_array_equal(nn_oracle.commit_settled(b, z, crit, 1e-5, 1), nn_oracle.commit_surface(b, z, 1e-5))
    mass = b.astype(np.float64).sum()
    for r in (1, 2, 3):
        assert abs(nn_oracle.commit_settled(b, z, crit, 1e-5, r).astype(np.float64).sum() - mass) <= 1e-7 * mass


@pytest.mark.parametrize("dil", [1, 2, 4, 8])
def test_conv_matches_library_convolution(dil):
    """nn_oracle.conv with zero padding is the cross-correlation of a library
    convolution: CPU torch.nn.functional.conv2d (float64, dilation d,
    padding d) with the taps in row-major order of (dy, dx) in -1..1."""
    import torch

    rng = np.random.default_rng(dil)
    x = rng.standard_normal((2, 11, 13, 5))
    w = rng.standard_normal((9, 3, 5))
    bias = rng.standard_normal(3)
    got = nn_oracle.conv(x, w, bias, dil)
    wt = torch.from_numpy(w.reshape(3, 3, 3, 5).transpose(2, 3, 0, 1).copy())     # [Cout, Cin, ky, kx]
    ref = torch.nn.functional.conv2d(torch.from_numpy(x.transpose(0, 3, 1, 2).copy()), wt,
                                     torch.from_numpy(bias), padding=dil, dilation=dil)
    np.testing.assert_allclose(got, ref.numpy().transpose(0, 2, 3, 1), rtol=1e-12, atol=1e-12)


def test_conv_centre_fill_channel_by_brute_force():
    """Layer 1's b_eff channel takes the centre value for taps outside the
    terrain (a constant height offset then cancels through the zero-sum taps):
    checked against a direct loop on a small input."""
    rng = np.random.default_rng(7)
    B, H, W, dil = 1, 5, 6, 2
    x = rng.standard_normal((B, H, W, 3))
    w = rng.standard_normal((9, 4, 3))
    got = nn_oracle.conv(x, w, None, dil, replicate_channel0=True)
    ref = np.zeros((B, H, W, 4))
    for y in range(H):
        for xx in range(W):
            for t, (ty, tx) in enumerate(nn_oracle.TAPS):
                yy, xs = y + ty * dil, xx + tx * dil
                inside = 0 <= yy < H and 0 <= xs < W
                v = x[0, yy, xs] if inside else np.array([x[0, y, xx, 0], 0.0, 0.0])
                ref[0, y, xx] += w[t] @ v
    np.testing.assert_allclose(got, ref, rtol=1e-12, atol=1e-12)
