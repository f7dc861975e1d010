"""Training and evaluation of the learned relaxation warm start (MI355X).

Training pairs: seeded 64 x 64 terrain patches (planar cut faces between a
noisy floor and plateau, cone piles and trenches, faces of 30-60 degrees)
relaxed by the plain relaxation -- the batched HIP patch relaxation, one
workgroup per patch -- at the *target angle* ``critical - margin``.  The
network learns that relaxed change of b_eff.  At run time its prediction is
mass-projected and then the exact relaxation at the critical angle takes
over: a surface predicted slightly below the critical slope leaves only the
network's own local errors above it, and those relax in a few sweeps, where
the plain relaxation from the over-steep terrain needs ~1000 iterations for
the slowly decaying long-wavelength excess.  The plain relaxation remains
the error judge: a warm-started field counts as at rest only when its maximum
excess is at or below the same tolerance.

Everything on the GPU runs through libmv2 (features, convolutions, loss,
Adam, patch relaxation); numpy synthesises the terrain patches on the host.

Run ``python -m isaac_bulk_hip.warm_start_training --help`` for the CLI.
"""

from __future__ import annotations

import argparse
import json
import time
from pathlib import Path

import numpy as np

from .relaxation import legacy_critical_differences, relax_device
from .runtime import HipRuntime
from .surrogate import WarmStartNet, relax_patches, warm_start_relax

DATA_DIR = Path(__file__).resolve().parent / "data"
DEFAULT_WEIGHTS = DATA_DIR / "warm_start_34deg.npz"


def _torch():
    import torch

    return torch


# ---------------------------------------------------------------------------
# terrain patches (host, vectorised over the batch)
# ---------------------------------------------------------------------------
def _interp_matrix(size: int, cells: int) -> np.ndarray:
    """[size, cells + 2] smoothstep interpolation weights of a 1-D lattice."""
    t = np.linspace(0.0, cells, size)
    i0 = np.floor(t).astype(int)
    f = t - i0
    f = f * f * (3 - 2 * f)
    m = np.zeros((size, cells + 2), np.float32)
    m[np.arange(size), i0] = 1.0 - f
    m[np.arange(size), i0 + 1] += f
    return m


def _smooth_noise(rng: np.random.Generator, count: int, size: int, cells: int) -> np.ndarray:
    """[count, size, size] value noise in [0, 1): a (cells+2)^2 lattice, smoothstep-interpolated."""
    lat = rng.uniform(0.0, 1.0, size=(count, cells + 2, cells + 2)).astype(np.float32)
    m = _interp_matrix(size, cells)
    return m @ lat @ m.T


def synth_patches(count: int, size: int = 64, dx: float = 0.05, seed: int = 0,
                  angle_range=(30.0, 60.0)) -> np.ndarray:
    """Seeded over-steep terrain patches [count, size, size] (float32, heights >= 0)."""
    rng = np.random.default_rng(seed)
    L = size * dx
    yy, xx = (np.mgrid[0:size, 0:size] * dx).astype(np.float32)
    f32 = np.float32

    def u(lo, hi):
        return rng.uniform(lo, hi, size=(count, 1, 1)).astype(f32)

    kind = rng.integers(0, 4, size=count)[:, None, None]   # 0, 1 planar faces; 2 cone; 3 trench
    s = np.tan(np.deg2rad(u(*angle_range)))
    phi = u(0.0, 2 * np.pi)
    floor = u(0.1, 0.5) + f32(0.3) * u(0.0, 1.0) * _smooth_noise(rng, count, size, 3)
    plateau = floor + u(0.2, 1.6) * (f32(0.6) + f32(0.4) * _smooth_noise(rng, count, size, 4))
    # signed distance across a face through a random offset from the centre
    off = u(-0.3, 0.3) * f32(L)
    d = (xx[None] - f32(L / 2)) * np.cos(phi) + (yy[None] - f32(L / 2)) * np.sin(phi) + off
    rise = plateau - floor
    face = floor + np.clip(s * d, 0.0, rise)
    r = np.hypot(yy[None] - u(0.25 * L, 0.75 * L), xx[None] - u(0.25 * L, 0.75 * L))
    cone = floor + np.clip(rise - s * r, 0.0, None)
    trench = plateau - np.clip(rise - s * (np.abs(d) - u(0.05, 0.4)), 0.0, rise)
    z = np.where(kind <= 1, face, np.where(kind == 2, cone, trench))
    z += f32(0.005) * u(0.0, 1.0) * rng.standard_normal(z.shape, dtype=f32)
    return np.maximum(z, 0.0).astype(np.float32)


def make_pairs(rt: HipRuntime, count: int, *, size: int = 64, dx: float = 0.05, seed: int = 0,
               target_angle_deg: float = 33.5, max_iterations: int = 6000, tolerance: float = 1.0e-5,
               chunk: int = 4096, workers: int = 8):
    """Device float32 (pre, change) tensors [count, size, size]: change = relaxed - pre.

    Host chunks are synthesised by a thread pool (numpy releases the GIL) while
    the GPU relaxes the previous chunk; chunk k is seeded from (seed, k) alone.
    """
    from concurrent.futures import ThreadPoolExecutor

    torch = _torch()
    pre = torch.empty((count, size, size), dtype=torch.float32, device=rt.device)
    change = torch.empty_like(pre)
    iters = []
    starts = list(range(0, count, chunk))
    with ThreadPoolExecutor(max_workers=workers) as pool:
        futures = [pool.submit(synth_patches, min(count, k0 + chunk) - k0, size, dx,
                               seed * 1000003 + k0) for k0 in starts]
        for k0, fut in zip(starts, futures):
            k1 = min(count, k0 + chunk)
            p = torch.from_numpy(fut.result()).to(rt.device)
            pre[k0:k1] = p
            post = p.clone()
            its = relax_patches(rt, post, dx, target_angle_deg, max_iterations, tolerance)
            change[k0:k1] = post - p
            iters.append(its.cpu().numpy())
    return pre, change, np.concatenate(iters)


# ---------------------------------------------------------------------------
# training
# ---------------------------------------------------------------------------
def train(net: WarmStartNet, pre, change, dx: float, *, steps: int, batch: int = 64, lr: float = 2e-3,
          seed: int = 0, log_every: int = 200, log=print, excess_weight: float = 0.0,
          critical_deg: float = 34.0) -> list:
    """Adam on the MSE of the predicted change (+ the slope-excess penalty of the
    warm-started surface at the critical angle); cosine learning-rate decay."""
    torch = _torch()
    crit = legacy_critical_differences(dx, dx, critical_deg)
    gen = torch.Generator(device="cpu").manual_seed(seed)
    n = pre.shape[0]
    losses = []
    t0 = time.perf_counter()
    for step in range(steps):
        idx = torch.randint(0, n, (batch,), generator=gen).to(pre.device)
        b = pre.index_select(0, idx).contiguous()
        target = change.index_select(0, idx).contiguous()
        rate = lr * 0.5 * (1.0 + np.cos(np.pi * step / max(steps, 1)))
        loss = net.train_step(b, None, target, dx, dx, lr=float(max(rate, lr * 0.02)),
                              excess_weight=excess_weight, excess_crit=crit)
        losses.append(loss)
        if log and (step % log_every == 0 or step == steps - 1):
            mse, ex = net.last_losses
            log(f"step {step:6d}  mse {mse:.3e} (rms {np.sqrt(mse):.3e} m)  excess {ex:.3e}  "
                f"{(time.perf_counter() - t0):.1f} s")
    return losses


def save_weights(net: WarmStartNet, path: Path, meta: dict) -> None:
    path.parent.mkdir(parents=True, exist_ok=True)
    params = {k: v.astype(np.float32) for k, v in net.params_numpy().items()}
    np.savez(path, meta=np.frombuffer(json.dumps(meta).encode(), dtype=np.uint8), **params)


def load_weights(path: Path = DEFAULT_WEIGHTS):
    """(params dict, meta dict) from a weights file written by ``save_weights``."""
    with np.load(path, allow_pickle=False) as f:
        meta = json.loads(bytes(f["meta"]).decode())
        params = {k: np.asarray(f[k], dtype=np.float64) for k in f.files if k != "meta"}
    return params, meta


def load_net(rt: HipRuntime, path: Path = DEFAULT_WEIGHTS):
    params, meta = load_weights(path)
    return WarmStartNet(rt, params=params), meta


# ---------------------------------------------------------------------------
# evaluation: plain vs learned relax-to-rest on a full terrain
# ---------------------------------------------------------------------------
def relax_to_rest_pair(rt: HipRuntime, net: WarmStartNet, b0: np.ndarray, dx: float, *,
                       critical_deg: float = 34.0, tolerance: float = 1.0e-6, sweeps: int = 4,
                       max_iterations: int = 20000) -> dict:
    """Plain and warm-started relaxation of the same field to the same tolerance."""
    torch = _torch()
    base = torch.from_numpy(np.ascontiguousarray(b0, dtype=np.float32)).to(rt.device)
    plain = base.clone()
    torch.cuda.synchronize(rt.device)
    t0 = time.perf_counter()
    it_p, conv_p, last_p = relax_device(plain, dx, dx, critical_deg, max_iterations, tolerance, runtime=rt)
    torch.cuda.synchronize(rt.device)
    plain_ms = (time.perf_counter() - t0) * 1e3
    learned = base.clone()
    res = warm_start_relax(net, learned, None, dx, critical_deg, tolerance, max_iterations=max_iterations,
                           sweeps=sweeps)
    diff = (learned - plain).abs().max().item()
    mass0 = float(base.double().sum().item())
    return {
        "plain_ms": plain_ms, "plain_iterations": it_p, "plain_converged": bool(conv_p),
        "plain_last_excess": last_p,
        "learned_ms": res.elapsed_ms,
        "learned_iterations": res.iterations_exact + res.fallback_iterations,
        "learned_converged": bool(res.converged), "learned_used_fallback": bool(res.used_fallback),
        "learned_residual_after_sweeps": res.residual_after_sweeps,
        "envelope_launches": res.envelope_launches,
        "speedup": plain_ms / res.elapsed_ms if res.elapsed_ms > 0 else None,
        "max_abs_learned_minus_plain_m": diff,
        "mass_rel_change_learned": abs(float(learned.double().sum().item()) - mass0) / mass0,
        "mass_rel_change_plain": abs(float(plain.double().sum().item()) - mass0) / mass0,
    }


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--pairs", type=int, default=200000)
    ap.add_argument("--steps", type=int, default=6000)
    ap.add_argument("--batch", type=int, default=64)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--critical-deg", type=float, default=34.0)
    ap.add_argument("--margin-deg", type=float, default=1.0)
    ap.add_argument("--excess-weight", type=float, default=0.0)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--out", type=Path, default=DEFAULT_WEIGHTS)
    ap.add_argument("--eval-seeds", type=int, default=3)
    args = ap.parse_args(argv)
    torch = _torch()
    rt = HipRuntime("cuda:0")
    dx = 0.05
    target = args.critical_deg - args.margin_deg
    t0 = time.perf_counter()
    pre, change, its = make_pairs(rt, args.pairs, dx=dx, seed=args.seed, target_angle_deg=target)
    torch.cuda.synchronize()
    print(f"pairs {args.pairs}: {time.perf_counter() - t0:.1f} s, relaxation iterations "
          f"median {int(np.median(its))} max {int(its.max())}", flush=True)
    net = WarmStartNet(rt, seed=args.seed)
    losses = train(net, pre, change, dx, steps=args.steps, batch=args.batch, lr=args.lr, seed=args.seed,
                   log=lambda s: print(s, flush=True), excess_weight=args.excess_weight,
                   critical_deg=args.critical_deg)
    meta = {"critical_deg": args.critical_deg, "margin_deg": args.margin_deg, "target_deg": target,
            "pairs": args.pairs, "steps": args.steps, "excess_weight": args.excess_weight, "batch": args.batch, "lr": args.lr,
            "seed": args.seed, "final_loss": float(np.mean(losses[-50:])), "dx": dx, "patch": 64}
    save_weights(net, args.out, meta)
    print(f"saved {args.out}", flush=True)
    from .workloads import cut_face_terrain

    for s in range(args.eval_seeds):
        b0, _, _ = cut_face_terrain(701, dx, seed=100 + s, face_deg=40.0 + 5.0 * s)
        r = relax_to_rest_pair(rt, net, b0, dx, critical_deg=args.critical_deg)
        print(json.dumps({"seed": 100 + s, **r}), flush=True)
    rt.close()
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
