"""NumPy oracle of the learned warm-start network (forward + backward).

TEST INFRASTRUCTURE ONLY.  Restates the layer definitions of the HIP kernels
in csrc/mv2_surrogate.hip: features (b_eff, h_mobile, |grad H|), layer 1 in
FP32 with replicate padding on the b_eff channel, layers 2-5 as dilated 3x3
convolutions with zero padding, ReLU after layers 1-4, a single-channel
output.  ``bf16=True`` rounds activations and weights to BF16 at the points
where the device does; ``bf16=False`` is plain float64 and is what the
finite-difference test pins.
"""

from __future__ import annotations

import numpy as np

DILATIONS = (1, 2, 4, 8, 16)
TAPS = [(t // 3 - 1, t % 3 - 1) for t in range(9)]


def bf16(x: np.ndarray) -> np.ndarray:
    """Round float32 values to bfloat16 (round to nearest even), returned as float32."""
    a = np.ascontiguousarray(np.asarray(x, dtype=np.float32)).view(np.uint32).astype(np.uint64)
    r = ((a + 0x7FFF + ((a >> 16) & 1)) >> 16) << 16
    return (r & 0xFFFFFFFF).astype(np.uint32).view(np.float32)


def features(b: np.ndarray, h: np.ndarray, dx: float, dy: float) -> np.ndarray:
    """[B,H,W] -> [B,H,W,3] (b_eff, h_mobile, |grad(b+h)|), float32 arithmetic."""
    b = np.asarray(b, np.float32)
    h = np.zeros_like(b) if h is None else np.asarray(h, np.float32)
    Hf = b + h
    B, H, W = b.shape
    xs = np.arange(W)
    xl, xr = np.maximum(xs - 1, 0), np.minimum(xs + 1, W - 1)
    ys = np.arange(H)
    yd, yu = np.maximum(ys - 1, 0), np.minimum(ys + 1, H - 1)
    gx = (Hf[:, :, xr] - Hf[:, :, xl]) / ((xr - xl).astype(np.float32) * np.float32(dx))[None, None, :]
    gy = (Hf[:, yu, :] - Hf[:, yd, :]) / ((yu - yd).astype(np.float32) * np.float32(dy))[None, :, None]
    slope = np.sqrt(gx * gx + gy * gy).astype(np.float32)
    return np.stack((b, h, slope), axis=-1)


def _shift(x: np.ndarray, dy: int, dx: int, fill_center: bool = False) -> np.ndarray:
    """out[n, y, x] = x[n, y+dy, x+dx] (zero outside, or the centre value)."""
    B, H, W = x.shape[:3]
    out = np.array(x, copy=True) if fill_center else np.zeros_like(x)
    ys0, ys1 = max(0, -dy), min(H, H - dy)
    xs0, xs1 = max(0, -dx), min(W, W - dx)
    if ys1 > ys0 and xs1 > xs0:
        out[:, ys0:ys1, xs0:xs1] = x[:, ys0 + dy:ys1 + dy, xs0 + dx:xs1 + dx]
    return out


def conv(x: np.ndarray, w: np.ndarray, bias, dil: int, *, replicate_channel0: bool = False) -> np.ndarray:
    """x [B,H,W,Cin], w [9,Cout,Cin] -> [B,H,W,Cout] (float64 accumulation)."""
    x = np.asarray(x, np.float64)
    out = np.zeros(x.shape[:3] + (w.shape[1],))
    for t, (ty, tx) in enumerate(TAPS):
        xs = _shift(x, ty * dil, tx * dil)
        if replicate_channel0:
            xs[..., 0] = _shift(x[..., :1], ty * dil, tx * dil, fill_center=True)[..., 0]
        out += xs @ np.asarray(w[t], np.float64).T
    if bias is not None:
        out += np.asarray(bias, np.float64)
    return out


def init_params(seed: int = 0) -> dict:
    rng = np.random.default_rng(seed)
    p = {}
    p["w1"] = (rng.standard_normal((9, 32, 3)) * np.sqrt(2.0 / 27.0)).astype(np.float32)
    p["w1"][:, :, 0] -= p["w1"][:, :, 0].mean(axis=0, keepdims=True)
    p["b1"] = np.zeros(32, np.float32)
    for k in (2, 3, 4):
        p[f"w{k}"] = (rng.standard_normal((9, 32, 32)) * np.sqrt(2.0 / 288.0)).astype(np.float32)
        p[f"b{k}"] = np.zeros(32, np.float32)
    p["w5"] = (rng.standard_normal((9, 1, 32)) * np.sqrt(1.0 / 288.0) * 0.1).astype(np.float32)
    p["b5"] = np.zeros(1, np.float32)
    return p


def forward(params: dict, feat: np.ndarray, *, bf16_mode: bool = True):
    """Return (output [B,H,W], activations list) for features [B,H,W,3]."""
    rnd = bf16 if bf16_mode else (lambda a: np.asarray(a, np.float64))
    acts = [np.asarray(feat, np.float64)]
    z = conv(feat, params["w1"], params["b1"], DILATIONS[0], replicate_channel0=True)
    y = rnd(np.maximum(z, 0.0).astype(np.float32) if bf16_mode else np.maximum(z, 0.0))
    acts.append(np.asarray(y, np.float64))
    for k in (2, 3, 4):
        w = rnd(params[f"w{k}"]) if bf16_mode else params[f"w{k}"]
        z = conv(acts[-1], w, params[f"b{k}"], DILATIONS[k - 1])
        y = rnd(np.maximum(z, 0.0).astype(np.float32) if bf16_mode else np.maximum(z, 0.0))
        acts.append(np.asarray(y, np.float64))
    w5 = rnd(params["w5"]) if bf16_mode else params["w5"]
    out = conv(acts[-1], w5, params["b5"], DILATIONS[4])[..., 0]
    return out, acts


def _conv_backward_input(gz: np.ndarray, w: np.ndarray, dil: int) -> np.ndarray:
    """dX[p] = sum_t W[t]^T gz[p - off(t)]."""
    out = np.zeros(gz.shape[:3] + (w.shape[2],))
    for t, (ty, tx) in enumerate(TAPS):
        out += _shift(gz, -ty * dil, -tx * dil) @ np.asarray(w[t], np.float64)
    return out


def _conv_backward_weight(gz: np.ndarray, x: np.ndarray, dil: int, *, replicate_channel0=False):
    cout, cin = gz.shape[-1], x.shape[-1]
    dw = np.zeros((9, cout, cin))
    for t, (ty, tx) in enumerate(TAPS):
        xs = _shift(x, ty * dil, tx * dil)
        if replicate_channel0:
            xs[..., 0] = _shift(x[..., :1], ty * dil, tx * dil, fill_center=True)[..., 0]
        dw[t] = np.einsum("bhwo,bhwi->oi", gz, xs)
    return dw, gz.sum(axis=(0, 1, 2))


def backward(params: dict, acts: list, grad_out: np.ndarray, *, bf16_mode: bool = True) -> dict:
    """Gradients of sum(grad_out * output) w.r.t. every parameter."""
    rnd = bf16 if bf16_mode else (lambda a: np.asarray(a, np.float64))
    grads = {}
    g = np.asarray(grad_out, np.float64)[..., None]
    w5 = rnd(params["w5"]) if bf16_mode else params["w5"]
    grads["w5"], grads["b5"] = _conv_backward_weight(g, acts[4], DILATIONS[4])
    gx = _conv_backward_input(rnd(g.astype(np.float32)) if bf16_mode else g, w5, DILATIONS[4])
    for k in (4, 3, 2):
        gz = gx * (acts[k] > 0)
        if bf16_mode:
            gz = np.asarray(bf16(gz.astype(np.float32)), np.float64)
        grads[f"w{k}"], grads[f"b{k}"] = _conv_backward_weight(gz, acts[k - 1], DILATIONS[k - 1])
        wk = rnd(params[f"w{k}"]) if bf16_mode else params[f"w{k}"]
        gx = _conv_backward_input(gz, wk, DILATIONS[k - 1])
    gz = gx * (acts[1] > 0)
    if bf16_mode:
        gz = np.asarray(bf16(gz.astype(np.float32)), np.float64)
    grads["w1"], grads["b1"] = _conv_backward_weight(gz, acts[0], DILATIONS[0], replicate_channel0=True)
    return grads


EDGE_OFFSETS = ((0, 1), (1, 0), (1, 1), (-1, 1))   # legacy directions d0..d3 as (row, col) a -> b


def excess_loss(pred: np.ndarray, pre: np.ndarray, crit, weight: float):
    """Slope-excess penalty of z = pre + pred over [B, H, W] and its gradient.

    loss = weight / N * sum over edges (a, b = a + off_d) of max(|z_a - z_b| - crit[d], 0)^2.
    """
    z = np.asarray(pre, np.float64) + np.asarray(pred, np.float64)
    B, H, W = z.shape
    n = B * H * W
    grad = np.zeros_like(z)
    total = 0.0
    for d, (dr, dc) in enumerate(EDGE_OFFSETS):
        r0, r1 = max(0, -dr), min(H, H - dr)
        c0, c1 = max(0, -dc), min(W, W - dc)
        za = z[:, r0:r1, c0:c1]
        zb = z[:, r0 + dr:r1 + dr, c0 + dc:c1 + dc]
        diff = za - zb
        ex = np.maximum(np.abs(diff) - crit[d], 0.0)
        total += float((ex * ex).sum())
        g = 2.0 * ex * np.sign(diff)
        grad[:, r0:r1, c0:c1] += g
        grad[:, r0 + dr:r1 + dr, c0 + dc:c1 + dc] -= g
    return weight * total / n, grad * (weight / n)


NEIGHBOURS = ((0, 1, 0), (0, -1, 0), (1, 0, 1), (-1, 0, 1), (1, 1, 2), (-1, -1, 2), (-1, 1, 3), (1, -1, 3))


def envelope(z: np.ndarray, crit, *, max_sweeps: int = 100000) -> np.ndarray:
    """Largest surface below z with |z_a - z_b| <= crit[d] across every legacy edge.

    Jacobi iteration z <- min(z, z_nbr + crit) over the 8 neighbours (edge
    direction d of each neighbour as in NEIGHBOURS) until nothing changes, in the
    dtype of z (float32 in, float32 arithmetic).
    """
    z = np.array(z, copy=True)
    H, W = z.shape
    c = np.asarray(crit, dtype=z.dtype)
    big = np.array(np.inf, dtype=z.dtype)
    for _ in range(max_sweeps):
        pad = np.full((H + 2, W + 2), big, dtype=z.dtype)
        pad[1:-1, 1:-1] = z
        m = z.copy()
        for dr, dc, d in NEIGHBOURS:
            cand = pad[1 + dr:1 + dr + H, 1 + dc:1 + dc + W] + c[d]
            np.minimum(m, cand, out=m)
        if np.array_equal(m, z):
            return z
        z = m
    raise RuntimeError("envelope did not converge")


def commit_surface(b: np.ndarray, z: np.ndarray, eps: float, corrections: int = 4,
                   return_mask: bool = False):
    """b + (z - b) over the moved cells |z - b| > eps, minus a shift s, heights
    clamped at 0 (float32).  s starts as the mean change over the moved cells;
    each correction adds (mass the clamp creates) / (moved cells above zero),
    so the moved cells keep their total when some are held at zero."""
    b = np.asarray(b, np.float32)
    d = np.asarray(z, np.float32) - b
    m = np.abs(d) > np.float32(eps)
    shift = np.float32(d[m].astype(np.float64).sum() / m.sum()) if m.any() else np.float32(0.0)
    for _ in range(corrections):
        v = np.maximum(b[m] + (d[m] - shift), np.float32(0.0))
        gain = float((v.astype(np.float64) - b[m].astype(np.float64)).sum())   # mass gained at this shift
        above = int((v > 0).sum())
        if above > 0:
            shift = np.float32(np.float64(shift) + gain / above)
    out = b.copy()
    out[m] = np.maximum(b[m] + (d[m] - shift), np.float32(0.0))
    return (out, m) if return_mask else out


def commit_masked(b: np.ndarray, z: np.ndarray, mask: np.ndarray, corrections: int = 4) -> np.ndarray:
    """The masked projection (float32): cells outside ``mask`` take z, the
    cells of ``mask`` take b + (z - b) - s, clamped at 0, with s balancing the
    mass of the whole field: s starts as (sum of z - b over all cells) / |mask|
    and each correction adds (mass the field gains at s) / (mask cells above
    zero), as in commit_surface."""
    b = np.asarray(b, np.float32)
    z = np.asarray(z, np.float32)
    m = np.asarray(mask, bool)
    d = z - b
    outside = float((z[~m].astype(np.float64) - b[~m].astype(np.float64)).sum())
    total = float((z.astype(np.float64) - b.astype(np.float64)).sum())
    shift = np.float32(total / m.sum()) if m.any() else np.float32(0.0)
    for _ in range(corrections):
        v = np.maximum(b[m] + (d[m] - shift), np.float32(0.0))
        gain = float((v.astype(np.float64) - b[m].astype(np.float64)).sum()) + outside
        above = int((v > 0).sum())
        if above > 0:
            shift = np.float32(np.float64(shift) + gain / above)
    out = z.copy()
    out[m] = np.maximum(b[m] + (d[m] - shift), np.float32(0.0))
    return out


def commit_settled(b: np.ndarray, z: np.ndarray, crit, eps: float, rounds: int, corrections: int = 4,
                   seam: int = 0) -> np.ndarray:
    """commit_surface, then (rounds - 1) settle rounds: the envelope of the
    committed field (fields of ``seam`` rows stacked: each on its own) and the
    masked projection onto the first commit's moved cells; a round whose
    envelope changes nothing ends them.  With seam, every field keeps its own
    shifts."""
    b = np.asarray(b, np.float32)
    z = np.asarray(z, np.float32)
    rows = seam if seam else b.shape[0]
    out = np.empty_like(b)
    for r0 in range(0, b.shape[0], rows):
        cur, mask = commit_surface(b[r0:r0 + rows], z[r0:r0 + rows], eps, corrections, return_mask=True)
        for _ in range(rounds - 1):
            env = envelope(cur, np.asarray(crit, np.float32))
            if np.array_equal(env, cur):
                break
            cur = commit_masked(cur, env, mask, corrections)
        out[r0:r0 + rows] = cur
    return out
