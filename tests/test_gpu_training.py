"""GPU training of the warm-start network: the Adam kernel against a NumPy
Adam step on the same gradients, one backward pass against the oracle's,
and a short seeded training run whose loss falls."""

from __future__ import annotations

import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_adam_kernel_matches_numpy_adam():
    import torch
    from isaac_bulk_hip.relaxation import _runtime

    rt = _runtime("cuda:0")
    rng = np.random.default_rng(0)
    n = 10007
    p = rng.standard_normal(n).astype(np.float32)
    m = np.zeros(n, np.float32)
    v = np.zeros(n, np.float32)
    dp, dm, dv = (torch.from_numpy(a.copy()).to("cuda:0") for a in (p, m, v))
    lr, b1, b2, eps = 2e-3, 0.9, 0.999, 1e-8
    for step in range(1, 6):
        g = rng.standard_normal(n).astype(np.float32) * np.float32(10.0 ** (-step))
        dg = torch.from_numpy(g).to("cuda:0")
        rt.bind_stream()
        rt.check(rt.lib.mv2_nn_adam(rt.ctx, dp.data_ptr(), dg.data_ptr(), dm.data_ptr(), dv.data_ptr(), n, lr,
                                    b1, b2, eps, step))
        # Adam (Kingma & Ba 2015, Algorithm 1) in float32
        m = np.float32(b1) * m + np.float32(1.0 - b1) * g
        v = np.float32(b2) * v + np.float32(1.0 - b2) * g * g
        mh = m / np.float32(1.0 - b1 ** step)
        vh = v / np.float32(1.0 - b2 ** step)
        p = p - np.float32(lr) * mh / (np.sqrt(vh) + np.float32(eps))
        torch.cuda.synchronize()
        np.testing.assert_allclose(dm.cpu().numpy(), m, rtol=1e-6, atol=1e-12)
        np.testing.assert_allclose(dv.cpu().numpy(), v, rtol=1e-5, atol=1e-20)
        np.testing.assert_allclose(dp.cpu().numpy(), p, rtol=1e-6, atol=1e-7)


def test_training_reduces_the_loss():
    """A fresh network (seed 0) trained for 300 Adam steps (batch 16) on 512
    seeded pre/post pairs of the plain relaxation: the mean loss of the last
    30 steps is well below that of the first 30, every loss is finite, and
    the projection keeps layer 1's b_eff taps at zero sum."""
    from isaac_bulk_hip.relaxation import _runtime
    from isaac_bulk_hip.surrogate import WarmStartNet
    from isaac_bulk_hip.warm_start_training import make_pairs, train

    rt = _runtime("cuda:0")
    pre, change, iters = make_pairs(rt, 512, size=64, seed=3)
    assert iters.max() > 10
    net = WarmStartNet(rt, seed=0)
    losses = train(net, pre, change, 0.05, steps=300, batch=16, lr=2e-3, seed=1, log=None)
    losses = np.asarray(losses)
    assert np.all(np.isfinite(losses))
    first, last = losses[:30].mean(), losses[-30:].mean()
    assert last < 0.5 * first, (first, last)
    w1 = net.params_numpy()["w1"]
    assert np.max(np.abs(w1[:, :, 0].sum(axis=0))) < 1e-5


@pytest.mark.parametrize("dil", [1, 2])
def test_mfma_weight_gradient_matches_float64_sum(dil):
    """The 32 -> 32 layer weight gradient on the matrix cores:
    dW[t][co][ci] = sum_p dZ[p][co] X[p + off(t)][ci] (zero padding),
    db[co] = sum_p dZ[p][co], against a float64 sum over the same BF16
    inputs on an uneven batch (FP32 accumulation: relative error ~1e-6)."""
    import torch
    from isaac_bulk_hip.relaxation import _runtime

    rt = _runtime("cuda:0")
    rt.bind_stream()
    B, H, W = 3, 37, 45
    rng = np.random.default_rng(20 + dil)
    dz = torch.from_numpy(rng.standard_normal((B, H, W, 32)).astype(np.float32)).to("cuda:0", torch.bfloat16)
    x = torch.from_numpy(rng.uniform(0.0, 1.0, (B, H, W, 32)).astype(np.float32)).to("cuda:0", torch.bfloat16)
    dw = torch.zeros(9 * 32 * 32, device="cuda:0")
    db = torch.zeros(32, device="cuda:0")
    rt.check(rt.lib.mv2_nn_weight_grad(rt.ctx, dz.data_ptr(), 0, x.data_ptr(), 0, B, H, W, dil, 32, 32,
                                       dw.data_ptr(), db.data_ptr()))
    torch.cuda.synchronize()
    z64 = dz.double().cpu().numpy()
    x64 = x.double().cpu().numpy()
    ref = np.zeros((9, 32, 32))
    for t in range(9):
        dy, dx = (t // 3 - 1) * dil, (t % 3 - 1) * dil
        xs = np.zeros_like(x64)
        ys0, ys1 = max(0, -dy), min(H, H - dy)
        xs0, xs1 = max(0, -dx), min(W, W - dx)
        xs[:, ys0:ys1, xs0:xs1] = x64[:, ys0 + dy:ys1 + dy, xs0 + dx:xs1 + dx]
        ref[t] = np.einsum("bhwo,bhwi->oi", z64, xs)
    got = dw.double().cpu().numpy().reshape(9, 32, 32)
    assert np.max(np.abs(got - ref)) <= 2e-5 * np.max(np.abs(ref))
    np.testing.assert_allclose(db.double().cpu().numpy(), z64.sum(axis=(0, 1, 2)), rtol=1e-5, atol=1e-4)
