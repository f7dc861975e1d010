"""The allocation-free ledger entry point (mv2_reservoirs_noalloc) against
mv2_reservoirs, through the C-ABI: bit-identical sums on two calls in a row,
the second on a smaller grid, for FP64 and FP32 storage."""

from __future__ import annotations

import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _state(rt, n, seed, precision):
    from isaac_bulk_hip import DeviceBulkState, TerrainGrid
    from isaac_bulk_hip.workloads import pile_with_patch

    b, h, _, _ = pile_with_patch(n, seed=seed)
    state = DeviceBulkState(TerrainGrid(n, n, 0.05, 0.05), b, precision=precision, runtime=rt)
    rng = np.random.default_rng(seed)
    state.set_fields(mobile=h, momentum_x=rng.normal(0.0, 0.01, (n, n)) * (h > 0),
                     momentum_y=rng.normal(0.0, 0.01, (n, n)) * (h > 0),
                     z_base=np.minimum(b, 0.1))
    return state


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_reservoirs_noalloc_matches_reservoirs_on_two_calls(precision):
    import torch
    from isaac_bulk_hip import HipRuntime

    rt = HipRuntime("cuda:0")
    out = torch.full((4,), float("nan"), dtype=torch.float64, device=rt.device)
    for n, seed in ((300, 1), (77, 2)):                       # second call: a smaller grid
        state = _state(rt, n, seed, precision)
        for weight_mode in (1, 0):
            ref = np.zeros(4)
            rt.bind_stream()
            rt.check(rt.lib.mv2_reservoirs(rt.ctx, weight_mode, 1370.0, ref.ctypes.data))
            out.fill_(float("nan"))
            rt.check(rt.lib.mv2_reservoirs_noalloc(rt.ctx, weight_mode, 1370.0, ctypes.c_void_p(out.data_ptr())))
            torch.cuda.synchronize()
            np.testing.assert_array_equal(out.cpu().numpy(), ref, err_msg=f"{precision} n={n} w={weight_mode}")
            assert ref[0] > 0.0 and ref[1] > 0.0
        # the Python surface over the same entry point
        state.reservoir_reduction_into(out, 1370.0)
        torch.cuda.synchronize()
        r = state.reservoir_reduction(1370.0)
        np.testing.assert_array_equal(out.cpu().numpy(), [r["resting_volume_m3"], r["mobile_volume_m3"],
                                                          r["mobile_momentum_kg_m_s"][0],
                                                          r["mobile_momentum_kg_m_s"][1]])
    rt.close()


def test_reservoirs_noalloc_rejects_bad_arguments():
    from isaac_bulk_hip import HipRuntime

    rt = HipRuntime("cuda:0")
    state = _state(rt, 32, 3, "fp64")
    with pytest.raises(ValueError, match="density"):
        rt.check(rt.lib.mv2_reservoirs_noalloc(rt.ctx, 1, 1370.0, None))
    import torch

    with pytest.raises(ValueError):
        state.reservoir_reduction_into(torch.zeros(4, dtype=torch.float32, device=rt.device), 1370.0)
    rt.close()
