"""GPU checks of the learned warm start: MFMA conv layout, forward/backward vs the
NumPy oracle, batched patch relaxation vs the relaxation oracle."""

from __future__ import annotations

import numpy as np
import pytest

import nn_oracle
import oracle

pytestmark = pytest.mark.gpu


def _t(a, dtype):
    import torch

    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda:0", dtype)


def test_mfma_conv_lane_layout_against_oracle():
    import torch
    from isaac_bulk_hip.relaxation import _runtime

    rt = _runtime("cuda:0")
    rng = np.random.default_rng(0)
    B, H, W = 2, 40, 70
    x = nn_oracle.bf16(rng.standard_normal((B, H, W, 32)).astype(np.float32))
    w = nn_oracle.bf16((rng.standard_normal((9, 32, 32)) * 0.1).astype(np.float32))
    bias = rng.standard_normal(32).astype(np.float32)
    wf = torch.zeros((9, 32, 32), dtype=torch.bfloat16, device="cuda:0")
    wt = _t(w, torch.float32)
    rt.bind_stream()
    rt.check(rt.lib.mv2_nn_pack_weights(rt.ctx, wt.data_ptr(), 32, 32, 32, 32, wf.data_ptr(), None))
    xt = _t(x, torch.bfloat16)
    bt = _t(bias, torch.float32)
    y = torch.zeros((B, H, W, 32), dtype=torch.bfloat16, device="cuda:0")
    rt.check(rt.lib.mv2_nn_conv_mfma(rt.ctx, xt.data_ptr(), 32, wf.data_ptr(), bt.data_ptr(),
                                     None, y.data_ptr(), B, H, W, 3, 1, 0))
    ref = np.maximum(nn_oracle.conv(x, w, bias, 3), 0.0)
    got = y.float().cpu().numpy()
    np.testing.assert_allclose(got, ref, rtol=1e-2, atol=2e-2)


def test_network_forward_backward_against_oracle():
    import torch
    from isaac_bulk_hip.relaxation import _runtime
    from isaac_bulk_hip.surrogate import WarmStartNet

    rt = _runtime("cuda:0")
    params = {k: v.astype(np.float64) for k, v in nn_oracle.init_params(5).items()}
    params["w5"] *= 10.0
    net = WarmStartNet(rt, params=params)
    rng = np.random.default_rng(1)
    B, H, W = 2, 64, 64
    b = rng.uniform(0.5, 1.5, size=(B, H, W)).astype(np.float32)
    h = np.where(rng.random((B, H, W)) < 0.2, rng.uniform(0, 0.1, (B, H, W)), 0.0).astype(np.float32)
    pred = net.forward(_t(b, torch.float32), _t(h, torch.float32), 0.05, 0.05).cpu().numpy()
    feat = nn_oracle.features(b, h, 0.05, 0.05)
    ref, acts = nn_oracle.forward(net.params_numpy(), feat, bf16_mode=True)
    scale = np.max(np.abs(ref))
    assert np.max(np.abs(pred - ref)) <= 3e-2 * scale
    go = rng.standard_normal((B, H, W)).astype(np.float32) / (B * H * W)
    grads = net.backward(_t(go, torch.float32), (B, H, W))
    rgrads = nn_oracle.backward(net.params_numpy(), acts, go, bf16_mode=True)
    for k in ("w5", "b5", "w4", "b4", "w3", "w2", "w1", "b1"):
        g = grads[k].cpu().numpy().astype(np.float64)
        if k == "w1":
            g = g[:, :, :3]
        r = rgrads[k].reshape(g.shape)
        err = np.linalg.norm(g - r) / max(np.linalg.norm(r), 1e-30)
        assert err <= 5e-2, (k, err)


def test_batched_patch_relaxation_bit_exact():
    import torch
    from isaac_bulk_hip.relaxation import _runtime
    from isaac_bulk_hip.surrogate import make_pre_patches, relax_patches

    rt = _runtime("cuda:0")
    pre = make_pre_patches(12, 64, seed=3)
    t = _t(pre, torch.float32)
    its = relax_patches(rt, t, 0.05, 34.0, 3000, 1e-6).cpu().numpy()
    got = t.cpu().numpy()
    crit = float(np.tan(np.deg2rad(34.0)))
    for i in range(12):
        o, it, conv, _ = oracle.relax_yx(pre[i], 0.05, 0.05, crit, 3000, 1e-6, dtype=np.float32)
        assert its[i] == it
        np.testing.assert_array_equal(got[i], o)


def test_excess_loss_kernel_against_oracle():
    import ctypes

    import torch
    from isaac_bulk_hip.relaxation import _runtime, legacy_critical_differences

    rt = _runtime("cuda:0")
    rng = np.random.default_rng(11)
    B, H, W = 3, 33, 47
    pre = rng.uniform(0.0, 0.3, size=(B, H, W)).astype(np.float32)
    pred = rng.uniform(-0.03, 0.03, size=(B, H, W)).astype(np.float32)
    g0 = rng.standard_normal((B, H, W)).astype(np.float32) * 1e-3
    crit = legacy_critical_differences(0.05, 0.05, 34.0)
    grad = _t(g0, torch.float32)
    pred_t, pre_t = _t(pred, torch.float32), _t(pre, torch.float32)   # held across the launch
    loss = ctypes.c_double(0.0)
    rt.bind_stream()
    rt.check(rt.lib.mv2_nn_excess_loss(rt.ctx, pred_t.data_ptr(), pre_t.data_ptr(), B, H, W, crit.ctypes.data,
                                       2.5, grad.data_ptr(), ctypes.byref(loss)))
    ref_loss, ref_grad = nn_oracle.excess_loss(pred.astype(np.float64), pre.astype(np.float64), crit, 2.5)
    assert ref_loss > 0.0
    # FP32 surface and edge differences on the device
    assert abs(loss.value - ref_loss) <= 1e-4 * ref_loss
    np.testing.assert_allclose(grad.cpu().numpy(), g0 + ref_grad, rtol=1e-3, atol=1e-9)


def test_envelope_kernel_bit_exact_against_oracle():
    import ctypes

    import torch
    from isaac_bulk_hip.relaxation import _runtime, legacy_critical_differences

    rt = _runtime("cuda:0")
    rng = np.random.default_rng(21)
    for (H, W) in ((37, 53), (130, 77)):
        b = rng.uniform(0.0, 0.5, size=(H, W)).astype(np.float32)
        delta = (rng.standard_normal((H, W)) * 0.05).astype(np.float32)
        crit = legacy_critical_differences(0.05, 0.05, 34.0)
        bt, dt = _t(b, torch.float32), _t(delta, torch.float32)
        z = torch.empty_like(bt)
        launches = (ctypes.c_int * 1)()
        rt.bind_stream()
        rt.check(rt.lib.mv2_envelope(rt.ctx, bt.data_ptr(), dt.data_ptr(), z.data_ptr(), H, W, crit.ctypes.data, 200,
                                     launches))
        ref = nn_oracle.envelope(b + delta, crit.astype(np.float32))
        np.testing.assert_array_equal(z.cpu().numpy(), ref)
        assert launches[0] >= 1


def test_commit_surface_against_oracle():
    import torch
    from isaac_bulk_hip.relaxation import _runtime

    rt = _runtime("cuda:0")
    rng = np.random.default_rng(22)
    b = rng.uniform(0.5, 1.0, size=(64, 96)).astype(np.float32)
    z = b + np.where(rng.random(b.shape) < 0.3, rng.standard_normal(b.shape) * 0.01, 0.0).astype(np.float32)
    bt, zt = _t(b, torch.float32), _t(z, torch.float32)
    rt.bind_stream()
    rt.check(rt.lib.mv2_nn_commit_surface(rt.ctx, bt.data_ptr(), zt.data_ptr(), b.size, 1e-5))
    np.testing.assert_allclose(bt.cpu().numpy(), nn_oracle.commit_surface(b, z, 1e-5), rtol=0, atol=2e-7)


def test_commit_surface_with_floor_clamp_against_oracle():
    """Cells pushed below zero: the GPU's shift corrections match the oracle's
    and the moved cells keep their total."""
    import torch
    from isaac_bulk_hip.relaxation import _runtime

    rt = _runtime("cuda:0")
    rng = np.random.default_rng(23)
    b = rng.uniform(0.0, 0.3, size=(64, 96)).astype(np.float32)
    z = b - (rng.uniform(0.0, 0.4, size=b.shape) * (rng.random(b.shape) < 0.5)).astype(np.float32)
    z[:8] += 0.5
    bt, zt = _t(b, torch.float32), _t(z, torch.float32)
    rt.bind_stream()
    rt.check(rt.lib.mv2_nn_commit_surface(rt.ctx, bt.data_ptr(), zt.data_ptr(), b.size, 1e-5))
    got = bt.cpu().numpy()
    np.testing.assert_allclose(got, nn_oracle.commit_surface(b, z, 1e-5), rtol=0, atol=2e-7)
    moved = np.abs(z - b) > 1e-5
    assert got.min() >= 0.0 and (got == 0.0).any()
    total = float(b[moved].astype(np.float64).sum())
    assert abs(float(got[moved].astype(np.float64).sum()) - total) <= 1e-6 * total


def test_commit_settled_against_oracle():
    """mv2_nn_commit_settled / mv2_batch_commit_settled against the oracle's
    settle rounds (float32 shifts from double sums in a different order: the
    same 2e-7 m bound as the plain commit), single field and stacked fields."""
    import ctypes

    import torch
    from isaac_bulk_hip.relaxation import _runtime, legacy_critical_differences

    rt = _runtime("cuda:0")
    crit = legacy_critical_differences(0.05, 0.05, 34.0)
    c32 = crit.astype(np.float32)
    rng = np.random.default_rng(23)
    fields = []
    for k in range(3):
        yy, xx = np.mgrid[0:48, 0:72]
        cone = (1.2 + 0.1 * k - 0.05 * np.hypot(yy - 24, xx - 30 - 4 * k) * np.tan(np.deg2rad(30)))
        b = nn_oracle.envelope(cone.astype(np.float32), c32)
        b[10 + k:30, 20:24 + k] -= 0.25
        pred = (nn_oracle.envelope(b, c32) - b) * 0.9 + rng.normal(0, 0.003, b.shape).astype(np.float32)
        z = nn_oracle.envelope((b + pred).astype(np.float32), c32)
        fields.append((b, z))
    launches = (ctypes.c_int * 1)()
    rt.bind_stream()
    for rounds in (1, 3):
        b, z = fields[0]
        bt, zt = _t(b, torch.float32), _t(z, torch.float32)
        rt.check(rt.lib.mv2_nn_commit_settled(rt.ctx, bt.data_ptr(), zt.data_ptr(), 48, 72, crit.ctypes.data, 1e-5,
                                              rounds, 256, launches))
        ref = nn_oracle.commit_settled(b, z, c32, 1e-5, rounds)
        np.testing.assert_allclose(bt.cpu().numpy(), ref, rtol=0, atol=2e-7)
        assert (launches[0] == 0) == (rounds == 1)
    B = np.concatenate([f[0] for f in fields])
    Z = np.concatenate([f[1] for f in fields])
    bt, zt = _t(B, torch.float32), _t(Z, torch.float32)
    rt.check(rt.lib.mv2_batch_commit_settled(rt.ctx, bt.data_ptr(), zt.data_ptr(), 3, 48, 72, crit.ctypes.data, 1e-5,
                                             3, 256, launches))
    ref = nn_oracle.commit_settled(B, Z, c32, 1e-5, 3, seam=48)
    np.testing.assert_allclose(bt.cpu().numpy(), ref, rtol=0, atol=2e-7)
