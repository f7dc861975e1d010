"""Pin the NumPy network oracle: finite differences of its float64 backward pass."""

from __future__ import annotations

import numpy as np

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
