"""Learned warm start for the angle-of-repose relaxation (MI355X).

A 5-layer dilated 3x3 conv-net (3 -> 32 -> 32 -> 32 -> 32 -> 1 channels,
dilations 1, 2, 4, 8, 16, receptive field 63 cells) reads (b_eff, h_mobile,
|grad H|) and predicts the change of b_eff that the plain relaxation
(slope_model.relax_critical_slope, slope_model.py:273-343) would make.
Layer 1 runs in FP32 (terrain heights need it); layers 2-5 are BF16 MFMA
implicit GEMMs with FP32 accumulation (csrc/mv2_surrogate.hip).  Training,
inference, loss, Adam and the training-pair generator (batched plain
relaxation of 64 x 64 patches) all run as HIP kernels; torch only holds the
buffers.

``warm_start_relax`` is the accelerated relaxation: inference, a mass
projection (the predicted change is shifted to zero sum), at most four exact
relaxation iterations, and -- if the residual excess still exceeds the
fallback threshold -- the plain relaxation continued to the same tolerance.
The plain relaxation is both the fallback and the error judge.
"""

from __future__ import annotations

import ctypes
import time
from dataclasses import dataclass

import numpy as np

from . import _lib
from .relaxation import legacy_critical_differences
from .runtime import HipRuntime

DILATIONS = (1, 2, 4, 8, 16)


def _torch():
    import torch

    return torch


class WarmStartNet:
    """Parameters, BF16 packed weights, Adam state and activation buffers on one GPU."""

    def __init__(self, runtime: HipRuntime, *, seed: int = 0, params: dict | None = None) -> None:
        torch = _torch()
        self.rt = runtime
        dev = runtime.device
        if params is None:
            import sys
            from pathlib import Path

            rng = np.random.default_rng(seed)
            params = {}
            w1 = rng.standard_normal((9, 32, 3)) * np.sqrt(2.0 / 27.0)
            w1[:, :, 0] -= w1[:, :, 0].mean(axis=0, keepdims=True)
            params["w1"] = w1
            params["b1"] = np.zeros(32)
            for k in (2, 3, 4):
                params[f"w{k}"] = rng.standard_normal((9, 32, 32)) * np.sqrt(2.0 / 288.0)
                params[f"b{k}"] = np.zeros(32)
            params["w5"] = rng.standard_normal((9, 1, 32)) * np.sqrt(1.0 / 288.0) * 0.1
            params["b5"] = np.zeros(1)
        self.p: dict[str, object] = {}
        for name, value in params.items():
            v = np.asarray(value, dtype=np.float32)
            if name == "w1":       # master layout [9][32][4] (channel 3 unused)
                v4 = np.zeros((9, 32, 4), np.float32)
                v4[:, :, :3] = v
                v = v4
            self.p[name] = torch.from_numpy(np.ascontiguousarray(v)).to(dev)
        self.m = {k: torch.zeros_like(v) for k, v in self.p.items()}
        self.v = {k: torch.zeros_like(v) for k, v in self.p.items()}
        self.step_count = 0
        bf = torch.bfloat16
        self.wf = {k: torch.zeros((9, 32, 32), dtype=bf, device=dev) for k in (2, 3, 4, 5)}
        self.wb = {k: torch.zeros((9, 32, 32), dtype=bf, device=dev) for k in (2, 3, 4)}
        self.wb[5] = torch.zeros((9, 32, 16), dtype=bf, device=dev)
        self._bufs: dict = {}
        self.pack()

    # ---- parameter handling -------------------------------------------------------
    def params_numpy(self) -> dict:
        out = {k: v.cpu().numpy().astype(np.float64) for k, v in self.p.items()}
        out["w1"] = out["w1"][:, :, :3]
        return out

    def pack(self) -> None:
        rt = self.rt
        rt.bind_stream()
        for k in (2, 3, 4, 5):
            cout = 1 if k == 5 else 32
            cxb = 16 if k == 5 else 32
            rt.check(rt.lib.mv2_nn_pack_weights(rt.ctx, self.p[f"w{k}"].data_ptr(), cout, 32, 32, cxb,
                                                self.wf[k].data_ptr(), self.wb[k].data_ptr()))

    # ---- buffers -------------------------------------------------------------------
    def _buffers(self, B: int, H: int, W: int) -> dict:
        key = (B, H, W)
        if key not in self._bufs:
            torch = _torch()
            dev = self.rt.device
            bf = torch.bfloat16
            n = B * H * W
            self._bufs = {key: {
                "feat": torch.zeros((n, 4), dtype=torch.float32, device=dev),
                "y": [None] + [torch.zeros((n, 32), dtype=bf, device=dev) for _ in range(4)],
                "out": torch.zeros(n, dtype=torch.float32, device=dev),
                "surface": torch.zeros(n, dtype=torch.float32, device=dev),
                "g5": torch.zeros(n, dtype=torch.float32, device=dev),
                "g5x": torch.zeros((n, 16), dtype=bf, device=dev),
                "dz": [torch.zeros((n, 32), dtype=bf, device=dev) for _ in range(2)],
            }}
        return self._bufs[key]

    # ---- forward / backward --------------------------------------------------------
    def forward(self, b, h, dx: float, dy: float, events: list | None = None):
        """b, h: device float32 [B, H, W] -> device float32 prediction [B, H, W].

        ``events``: optional list of 6 torch.cuda.Event recorded on the
        library's stream around the launches (features | layer 1 | layers
        2, 3, 4 | layer 5), for per-layer timing.
        """
        rt = self.rt
        rt.bind_stream()
        B, H, W = b.shape
        buf = self._buffers(B, H, W)
        lib, ctx = rt.lib, rt.ctx
        mark = (lambda i: events[i].record()) if events is not None else (lambda i: None)
        hp = h.data_ptr() if h is not None else None
        rt.check(lib.mv2_nn_features(ctx, b.data_ptr(), hp, B, H, W, dx, dy, buf["feat"].data_ptr()))
        mark(0)
        rt.check(lib.mv2_nn_conv1(ctx, buf["feat"].data_ptr(), self.p["w1"].data_ptr(),
                                  self.p["b1"].data_ptr(), B, H, W, DILATIONS[0], buf["y"][1].data_ptr()))
        mark(1)
        for k in (2, 3, 4):
            rt.check(lib.mv2_nn_conv_mfma(ctx, buf["y"][k - 1].data_ptr(), 32, self.wf[k].data_ptr(),
                                          self.p[f"b{k}"].data_ptr(), None, buf["y"][k].data_ptr(),
                                          B, H, W, DILATIONS[k - 1], 1, 0))
            mark(k)
        rt.check(lib.mv2_nn_conv_mfma(ctx, buf["y"][4].data_ptr(), 32, self.wf[5].data_ptr(),
                                      self.p["b5"].data_ptr(), None, buf["out"].data_ptr(),
                                      B, H, W, DILATIONS[4], 0, 1))
        mark(5)
        return buf["out"].view(B, H, W)

    def backward(self, grad_out, shape) -> dict:
        """Gradients of sum(grad_out * prediction) for the last forward call."""
        torch = _torch()
        rt = self.rt
        lib, ctx = rt.lib, rt.ctx
        B, H, W = shape
        n = B * H * W
        buf = self._buffers(B, H, W)
        grads = {k: torch.zeros_like(v) for k, v in self.p.items()}
        rt.check(lib.mv2_nn_weight_grad(ctx, grad_out.data_ptr(), 1, buf["y"][4].data_ptr(), 0, B, H, W,
                                        DILATIONS[4], 1, 32, grads["w5"].data_ptr(), grads["b5"].data_ptr()))
        rt.check(lib.mv2_nn_expand_grad(ctx, grad_out.data_ptr(), n, buf["g5x"].data_ptr()))
        dz = buf["dz"][0]
        rt.check(lib.mv2_nn_conv_mfma(ctx, buf["g5x"].data_ptr(), 16, self.wb[5].data_ptr(), None,
                                      buf["y"][4].data_ptr(), dz.data_ptr(), B, H, W, DILATIONS[4], 0, 0))
        cur = 0
        for k in (4, 3, 2):
            dzk = buf["dz"][cur]
            rt.check(lib.mv2_nn_weight_grad(ctx, dzk.data_ptr(), 0, buf["y"][k - 1].data_ptr(), 0, B, H, W,
                                            DILATIONS[k - 1], 32, 32, grads[f"w{k}"].data_ptr(),
                                            grads[f"b{k}"].data_ptr()))
            nxt = buf["dz"][1 - cur]
            rt.check(lib.mv2_nn_conv_mfma(ctx, dzk.data_ptr(), 32, self.wb[k].data_ptr(), None,
                                          buf["y"][k - 1].data_ptr(), nxt.data_ptr(), B, H, W,
                                          DILATIONS[k - 1], 0, 0))
            cur = 1 - cur
        rt.check(lib.mv2_nn_weight_grad(ctx, buf["dz"][cur].data_ptr(), 0, buf["feat"].data_ptr(), 1, B, H, W,
                                        DILATIONS[0], 32, 4, grads["w1"].data_ptr(), grads["b1"].data_ptr()))
        return grads

    def train_step(self, b, h, target, dx: float, dy: float, *, lr: float = 1e-3,
                   excess_weight: float = 0.0, excess_crit=None) -> float:
        """One Adam step on MSE(pred, target) [+ excess_weight * slope excess of b + pred].

        ``excess_crit`` is the legacy critical-difference vector of the angle the
        warm-started surface must not exceed.  Returns the total loss; the two
        terms are kept in ``last_losses``.
        """
        rt = self.rt
        pred = self.forward(b, h, dx, dy)
        B, H, W = b.shape
        buf = self._buffers(B, H, W)
        loss = ctypes.c_double(0.0)
        rt.check(rt.lib.mv2_nn_mse(rt.ctx, pred.data_ptr(), target.data_ptr(), B * H * W,
                                   buf["g5"].data_ptr(), ctypes.byref(loss)))
        ex = ctypes.c_double(0.0)
        if excess_weight > 0.0:
            crit = np.ascontiguousarray(excess_crit, dtype=np.float64)
            rt.check(rt.lib.mv2_nn_excess_loss(rt.ctx, pred.data_ptr(), b.data_ptr(), B, H, W, crit.ctypes.data,
                                               float(excess_weight), buf["g5"].data_ptr(), ctypes.byref(ex)))
        self.last_losses = (float(loss.value), float(ex.value))
        grads = self.backward(buf["g5"], (B, H, W))
        self.step_count += 1
        for k, p in self.p.items():
            rt.check(rt.lib.mv2_nn_adam(rt.ctx, p.data_ptr(), grads[k].data_ptr(), self.m[k].data_ptr(),
                                        self.v[k].data_ptr(), p.numel(), lr, 0.9, 0.999, 1e-8,
                                        self.step_count))
        rt.check(rt.lib.mv2_nn_project_w1(rt.ctx, self.p["w1"].data_ptr()))
        self.pack()
        return float(loss.value) + float(ex.value)


# ---------------------------------------------------------------------------
# training pairs: seeded piles/cuts with 30-60 degree faces, relaxed by the
# plain relaxation (batched, one workgroup per 64 x 64 patch)
# ---------------------------------------------------------------------------
def make_pre_patches(count: int, size: int = 64, dx: float = 0.05, seed: int = 0) -> np.ndarray:
    rng = np.random.default_rng(seed)
    y, x = np.mgrid[0:size, 0:size].astype(np.float64) * dx
    out = np.empty((count, size, size), np.float32)
    L = size * dx
    for i in range(count):
        kind = rng.integers(0, 3)
        ang = np.deg2rad(rng.uniform(30.0, 60.0))
        s = np.tan(ang)
        base = rng.uniform(0.3, 1.0)
        if kind == 0:      # cone pile
            cy, cx = rng.uniform(0.2 * L, 0.8 * L, 2)
            r = np.hypot(y - cy, x - cx)
            top = rng.uniform(0.3, 1.5)
            z = base + np.maximum(top - s * r, 0.0)
        elif kind == 1:    # planar cut face between a floor and a plateau
            phi = rng.uniform(0, 2 * np.pi)
            d = (x - L / 2) * np.cos(phi) + (y - L / 2) * np.sin(phi) + rng.uniform(-0.3, 0.3) * L
            z = base + np.clip(s * d, 0.0, rng.uniform(0.3, 1.2))
        else:              # trench cut into a plateau
            phi = rng.uniform(0, 2 * np.pi)
            d = np.abs((x - L / 2) * np.cos(phi) + (y - L / 2) * np.sin(phi) + rng.uniform(-0.2, 0.2) * L)
            depth = rng.uniform(0.2, 0.8)
            z = base + depth - np.clip(depth - s * (d - rng.uniform(0.05, 0.4)), 0.0, depth)
        z += 0.01 * rng.standard_normal(z.shape) * rng.uniform(0.0, 1.0)
        out[i] = np.maximum(z, 0.0)
    return out


def relax_patches(runtime: HipRuntime, patches, dx: float, critical_angle_deg: float,
                  max_iterations: int, tolerance: float):
    """Relax a device float32 [N, H, W] tensor of independent patches in place."""
    torch = _torch()
    rt = runtime
    rt.bind_stream()
    N, H, W = patches.shape
    crit = legacy_critical_differences(dx, dx, critical_angle_deg)
    its = torch.zeros(N, dtype=torch.int32, device=rt.device)
    rt.check(rt.lib.mv2_relax_patches(rt.ctx, patches.data_ptr(), N, H, W, crit.ctypes.data,
                                      int(max_iterations), float(tolerance), its.data_ptr()))
    return its


@dataclass
class WarmStartResult:
    iterations_exact: int
    fallback_iterations: int
    converged: bool
    used_fallback: bool
    residual_after_sweeps: float
    elapsed_ms: float
    envelope_launches: int = 0
    settle_launches: int = 0


def _relax_marked(rt: HipRuntime, b_field, dx: float, critical_angle_deg: float, max_iterations: int,
                  tolerance: float, mark: int) -> tuple[int, bool, float]:
    """Plain relaxation of a device float32 [H, W] field in place (relax_device
    semantics, closed boundary); returns (iterations, converged, max excess of
    iteration `mark`, or of the last one when the loop ended before it)."""
    H, W = b_field.shape
    crit = legacy_critical_differences(dx, dx, critical_angle_deg)
    res = _lib.RelaxResult()
    marked = ctypes.c_double(0.0)
    rt.bind_stream()
    rt.check(rt.lib.mv2_relax_marked(rt.ctx, _lib.MV2_FP32, b_field.data_ptr(), int(H), int(W), crit.ctypes.data,
                                     int(max_iterations), float(tolerance), int(mark), ctypes.byref(res),
                                     ctypes.byref(marked)))
    return int(res.iterations), bool(res.converged), float(marked.value)


def warm_start_relax(net: WarmStartNet, b_field, h_field, dx: float, critical_angle_deg: float,
                     tolerance: float, *, max_iterations: int = 20000, sweeps: int = 4,
                     fallback_threshold: float = 1e-6, eps: float = 1e-5,
                     envelope_angle_deg: float | None = None, settle_rounds: int = 1) -> WarmStartResult:
    """In-place accelerated relaxation of a device float32 [H, W] b_eff field.

    1. the network predicts the relaxed change of b_eff;
    2. the Lipschitz envelope (slopes <= envelope angle, default the critical
       angle) removes the local over-steep spots of the predicted surface;
    3. mass projection: the change is shifted to zero sum over the moved cells;
       then ``settle_rounds - 1`` rounds of the envelope of the committed field
       and a projection onto the same moved cells that balances the whole
       field's mass (each round shrinks the step the shift leaves at the rim of
       the moved cells; one field of 701^2: 0.52 ms to rest with no settle
       round, 11 exact iterations, against 0.61 ms with two rounds and 1
       iteration -- each round costs an envelope and a host check);
    4. at most ``sweeps`` exact relaxation iterations; if the maximum excess
       still exceeds the tolerance, the plain relaxation continues to the same
       tolerance (fallback; ``used_fallback`` when the residual after the sweeps
       was above ``fallback_threshold``).
    """
    torch = _torch()
    rt = net.rt
    torch.cuda.synchronize(rt.device)
    t0 = time.perf_counter()
    H, W = b_field.shape
    pred = net.forward(b_field.view(1, H, W), None if h_field is None else h_field.view(1, H, W), dx, dx)
    env_deg = critical_angle_deg if envelope_angle_deg is None else envelope_angle_deg
    crit = legacy_critical_differences(dx, dx, env_deg)
    launches = (ctypes.c_int * 1)()
    # the surface gets its own buffer: an envelope whose surface aliases the
    # prediction needs one more pass (b + delta in place before the flag test)
    zbuf = net._buffers(1, H, W)["surface"].view(H, W)
    rt.check(rt.lib.mv2_envelope(rt.ctx, b_field.data_ptr(), pred.data_ptr(), zbuf.data_ptr(), H, W,
                                 crit.ctypes.data, 256, launches))
    settle = (ctypes.c_int * 1)()
    rt.check(rt.lib.mv2_nn_commit_settled(rt.ctx, b_field.data_ptr(), zbuf.data_ptr(), H, W, crit.ctypes.data, eps,
                                          int(settle_rounds), 256, settle))
    # the sweeps and the fallback in one relaxation: the same iterations as a
    # call limited to `sweeps` followed by one from its result (each
    # iteration depends on the field alone), plus the residual after the sweeps
    total, conv, last = _relax_marked(rt, b_field, dx, critical_angle_deg, max_iterations + sweeps, tolerance,
                                      sweeps)
    it = min(total, sweeps)
    fb_it = total - it
    used = total > sweeps and last > fallback_threshold
    torch.cuda.synchronize(rt.device)
    return WarmStartResult(it, fb_it, conv, used, last, (time.perf_counter() - t0) * 1e3, int(launches[0]),
                           int(settle[0]))
