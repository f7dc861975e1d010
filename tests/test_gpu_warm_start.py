"""The trained warm start (committed weights) against the plain relaxation it
accelerates: same tolerance, same mass, far fewer exact iterations."""

from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("seed,face_deg", [(7, 42.0), (8, 52.0)])
def test_learned_warm_start_beats_plain_relaxation(seed, face_deg):
    import torch
    from isaac_bulk_hip.relaxation import _runtime, relax_device
    from isaac_bulk_hip.surrogate import warm_start_relax
    from isaac_bulk_hip.warm_start_training import load_net
    from isaac_bulk_hip.workloads import cut_face_terrain

    rt = _runtime("cuda:0")
    net, meta = load_net(rt)
    b0, _, _ = cut_face_terrain(351, 0.05, seed=seed, face_deg=face_deg)
    base = torch.from_numpy(b0.astype(np.float32)).to("cuda:0")
    tol = 1e-6
    plain = base.clone()
    it_plain, conv_plain, _ = relax_device(plain, 0.05, 0.05, 34.0, 20000, tol, runtime=rt)
    assert conv_plain and it_plain > 300
    learned = base.clone()
    res = warm_start_relax(net, learned, None, 0.05, 34.0, tol)
    assert res.converged
    exact = res.iterations_exact + res.fallback_iterations
    # the trained network removes most of the plain relaxation's iterations
    assert exact * 3 <= it_plain, (exact, it_plain)
    # judge: the plain relaxation finds nothing above the tolerance to move
    it, conv, last = relax_device(learned.clone(), 0.05, 0.05, 34.0, 1, tol, runtime=rt)
    assert conv and last <= tol
    m0 = float(base.double().sum())
    assert abs(float(learned.double().sum()) - m0) <= 1e-6 * m0
    # same neighbourhood of rest states as the plain relaxation (0.5 deg margin)
    assert float((learned - plain).abs().max()) < 0.1
