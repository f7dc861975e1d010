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
    y = torch.zeros((B, H, W, 32), dtype=torch.bfloat16, device="cuda:0")
    rt.check(rt.lib.mv2_nn_conv_mfma(rt.ctx, xt.data_ptr(), 32, wf.data_ptr(), _t(bias, torch.float32).data_ptr(),
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
