"""Benchmark: Mobile V2 cell-substeps/s on the 701x701 full-size frame workload.

Workload (one "step" = one 1/60 s frame of the hot path, all inputs resident
in HBM before timing starts):
  terrain  : 701 x 701 cells at dx = 0.05 m, seeded multi-octave value noise
             (amplitude 1.5 m) with a 40 degree cut face; h_mobile
             uniform(0.05, 0.3) m on the face
  per frame: Mobile V2 step_resident (CFL substeps), angle-of-repose relaxation
             of b_eff at 34 degrees, failure mask + union-find labelling
  default  : 300 frames = 5 s simulated, FP32 fields (int32 mobile quanta)
value = 701^2 * (substeps over the timed frames) / (wall time of the V2 step calls of
those frames; each call returns after a device sync).  The frame rate including
relaxation and labelling is reported beside it (frame_cell_substeps_per_s).

Also printed: relax-to-rest ms of the initial terrain, the V2 kernel's
roofline (28 B per cell-substep vs 8 TB/s) and the CPU oracle timed on a
bounded sample of the same workload.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "iron-soil-physical-simulation-algorithom-for-isaac-sim_amd"))

N = 701
DX = 0.05
FRAME_DT = 1.0 / 60.0
HBM_PEAK_GBS = 8000.0
# HBM bytes per timed substep launch of the FP32 kernel on this workload,
# from rocprofv3 FETCH_SIZE / WRITE_SIZE scaled by a 64 MiB copy
# (profiles/v2_substep_701x701_cutface_traffic.md): the mean of the checking
# launch (15.46 MB) and the closing substep (15.63 MB), the launches a
# 3-substep step times; counters cannot be read from inside the timed run.
MEASURED_TRAFFIC_BYTES = {"fp32": 15.5e6}
# the same for a full launch of the batched substep (256 terrains, 18 274 304
# cells; profiles/batch_substep_256_terrains.md)
MEASURED_TRAFFIC_BATCHED_FULL = 519.5e6
BYTES_PER_CELL_SUBSTEP = {"fp32": 28, "fp64": 56}


def parse() -> argparse.Namespace:
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=300)
    p.add_argument("--warmup", type=int, default=3)
    p.add_argument("--precision", choices=("fp32", "fp64"), default="fp32")
    p.add_argument("--seed", type=int, default=0)
    p.add_argument("--relax-max-iterations", type=int, default=1000)
    p.add_argument("--cpu-sample-frames", type=int, default=100000,
                   help="CPU baseline: frames to run at most (each mode stops after ~10 s)")
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--parity-frames", type=int, default=30,
                   help="frames of the FP32-vs-FP64-oracle parity check (cpu_baseline leg)")
    p.add_argument("--skip-configs", action="store_true",
                   help="omit the batched (256 terrains, 2 s) and digging (64 x 256^2, 3 s) legs of the default line")
    p.add_argument("--skip-relax-to-rest", action="store_true",
                   help="omit the relax-to-rest measurement (short profiling runs)")
    p.add_argument("--workload", choices=("frame", "batched", "digging", "deposition", "frontier"),
                   default="frame",
                   help="frame: the 701x701 frame (headline); batched: 256 uneven terrains; "
                        "digging: 64 terrains of 256^2 with a bucket sweep; deposition: the "
                        "Mobile -> Resting deposition step on the 701x701 frame state; frontier: "
                        "residual projection rounds after landings on the 701x701 terrain")
    p.add_argument("--dig-count", type=int, default=64)
    p.add_argument("--dig-relax", choices=("learned", "plain"), default="learned",
                   help="digging: learned warm start + exact relaxation, or the plain relaxation")
    p.add_argument("--batch-count", type=int, default=256)
    p.add_argument("--calibrate-copy", action="store_true",
                   help="only copy a 64 MiB float32 tensor 10 times (scale of the HBM byte counters)")
    return p.parse_args()


def build_state(precision: str, seed: int):
    from isaac_bulk_hip import (DeviceBulkState, DeviceLargeAvalancheBridge,
                                LargeAvalancheTransitionConfig, MaterialScenario, MobileV2Config,
                                TerrainGrid, TerrainVolumeIntegrator, WarpProductionMobileV2Solver)
    from isaac_bulk_hip.workloads import cut_face_terrain

    b, h, _ = cut_face_terrain(N, DX, seed=seed)
    grid = TerrainGrid(N, N, DX, DX)
    state = DeviceBulkState(grid, b, precision=precision)
    state.set_fields(mobile=h)
    material = MaterialScenario("iron_ore_fines_uncalibrated", 1370.0, 29.8, 0.0, 0.5, 34.0, 30.0, 0.55)
    config = MobileV2Config()
    integrator = TerrainVolumeIntegrator.from_grid(grid)
    solver = WarpProductionMobileV2Solver(runtime=state.runtime, config=config)
    solver.bind_device_state(state, material, grid, integrator)
    bridge = DeviceLargeAvalancheBridge(state, material, grid, integrator,
                                        LargeAvalancheTransitionConfig())
    return state, solver, bridge, b, h


def relax_tol(precision: str) -> float:
    # 1e-8 m is the reference tolerance (FP64); float32 heights of O(1 m) resolve ~1e-7 m
    return 1.0e-8 if precision == "fp64" else 1.0e-6


def run_frame(state, solver, bridge, precision, relax_iters):
    from isaac_bulk_hip import relax_device

    rec = solver.step_resident(FRAME_DT)
    b = state.runtime.arrays["b_eff"].view(N, N)
    it, conv, _ = relax_device(b, DX, DX, 34.0, relax_iters, relax_tol(precision), runtime=state.runtime)
    lab = bridge.label()
    state.advance_time(FRAME_DT)
    return rec.substeps, it, int(lab.diag_i[3])


def learned_relax_to_rest(state, plain_rest) -> dict:
    """Learned warm start (committed weights) + exact sweeps to the plain tolerance."""
    import torch
    from isaac_bulk_hip import relax_device
    from isaac_bulk_hip.surrogate import warm_start_relax
    from isaac_bulk_hip.warm_start_training import load_net

    net, meta = load_net(state.runtime)
    b0 = state.runtime.arrays["b_eff"].view(N, N)
    runs = []
    for _ in range(3):                     # the first call also sizes the network buffers
        b = b0.clone()
        res = warm_start_relax(net, b, None, DX, 34.0, relax_tol("fp32"))
        runs.append(res)
    res = runs[-1]
    # the plain relaxation as the judge: one more exact iteration finds nothing to move
    check_it, check_conv, check_last = relax_device(b.clone(), DX, DX, 34.0, 1, relax_tol("fp32"),
                                                    runtime=state.runtime)
    mass0 = float(b0.double().sum().item())
    # the same inputs as the warm start above (b_eff only)
    roof = surrogate_roofline(net, b0.view(1, N, N).contiguous(), None)
    return {"ms": res.elapsed_ms, "ms_runs": [r.elapsed_ms for r in runs],
            "surrogate_roofline": roof,
            "exact_iterations": res.iterations_exact + res.fallback_iterations,
            "exact_sweeps": res.iterations_exact, "fallback_iterations": res.fallback_iterations,
            "used_fallback": bool(res.used_fallback),
            "envelope_launches": res.envelope_launches, "converged": bool(res.converged),
            "judge_max_excess": check_last, "judge_converged": bool(check_conv),
            "mass_rel_change": abs(float(b.double().sum().item()) - mass0) / mass0,
            "max_abs_vs_plain_m": float((b - plain_rest).abs().max().item()),
            "weights": {"margin_deg": meta.get("margin_deg"), "pairs": meta.get("pairs"),
                        "steps": meta.get("steps")}}


# learned warm start: algorithmic work per pixel of one forward pass
NN_FLOP_LAYER1 = 9 * 3 * 32 * 2            # FP32 MFMA (v_mfma_f32_32x32x2_f32)
NN_FLOP_MID = 9 * 32 * 32 * 2              # each of layers 2-4, BF16 MFMA
NN_FLOP_LAST = 9 * 32 * 1 * 2              # layer 5 (one output channel)
NN_BYTES_MID = 32 * 2 * 2                  # BF16 NHWC32 activation read + written
MFMA_PEAK_BF16_TFLOPS = 2500.0
MFMA_PEAK_FP32_TFLOPS = 157.3


def surrogate_roofline(net, b, h, reps: int = 20) -> dict:
    """Per-layer device time of the warm-start network's forward pass (HIP
    events on the library's stream) and the matrix-core fraction of the 32 ->
    32 BF16 layers (conv_mfma_kernel), the network's dominant kernel."""
    import torch

    B, H, W = b.shape
    px = B * H * W
    net.forward(b, h, DX, DX)                                  # sizes the buffers
    torch.cuda.synchronize()
    tot = np.zeros(5)
    for _ in range(reps):
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(6)]
        net.forward(b, h, DX, DX, events=ev)
        torch.cuda.synchronize()
        tot += [ev[i].elapsed_time(ev[i + 1]) for i in range(5)]
    layer_us = tot / reps * 1e3                               # layer 1, 2, 3, 4, 5
    mid_us = float(np.mean(layer_us[1:4]))
    achieved = NN_FLOP_MID * px / (mid_us * 1e-6) / 1e12
    fwd_us = float(layer_us.sum())
    return {"kernel": "conv_mfma_kernel<2> (layers 2-4, 32 -> 32 channels, 3x3 dilated, BF16 MFMA)",
            "bound": "mfma", "achieved": achieved, "peak": MFMA_PEAK_BF16_TFLOPS, "unit": "TFLOP/s",
            "frac": achieved / MFMA_PEAK_BF16_TFLOPS,
            "hbm_GBps_algorithmic": NN_BYTES_MID * px / (mid_us * 1e-6) / 1e9,
            "layer1_fp32_mfma_frac": NN_FLOP_LAYER1 * px / (layer_us[0] * 1e-6) / 1e12 / MFMA_PEAK_FP32_TFLOPS,
            "per_layer_us": [round(float(v), 2) for v in layer_us],
            "forward_us": fwd_us,
            "forward_tflops": (NN_FLOP_LAYER1 + 3 * NN_FLOP_MID + NN_FLOP_LAST) * px / (fwd_us * 1e-6) / 1e12,
            "pixels": px, "shape": [B, H, W],
            "flop_per_pixel": {"layer1": NN_FLOP_LAYER1, "layers2_4_each": NN_FLOP_MID, "layer5": NN_FLOP_LAST}}


def cpu_baseline(b, h, frames: int) -> dict:
    """FP64 CPU oracle (the C restatement) on the first frames of the same
    workload, on every host core the process may use (OpenMP build; fields
    identical to the one-thread build) and on one core."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import oracle

    def run(threads: bool, budget_s: float):
        qx = np.zeros_like(h)
        qy = np.zeros_like(h)
        cur_h = h.copy()
        substeps, done = 0, 0
        t0 = time.perf_counter()
        for _ in range(frames):
            cur_h, qx, qy, rec = oracle.v2_step(b, cur_h, qx, qy, FRAME_DT, limit_mode=1, threads=threads)
            substeps += rec["substeps"]
            done += 1
            if time.perf_counter() - t0 > budget_s:
                break
        elapsed = time.perf_counter() - t0
        return N * N * substeps / elapsed, substeps, done, elapsed

    cores = int(os.environ.get("OMP_NUM_THREADS", "0")) or len(os.sched_getaffinity(0))
    os.environ["OMP_NUM_THREADS"] = str(cores)
    v_all, s_all, f_all, e_all = run(True, 10.0)
    v_one, s_one, f_one, e_one = run(False, 10.0)
    return {"value": v_all, "unit": "cell-substeps/s", "cores": cores, "kind": "port",
            "sample": f"oracle FP64 Mobile V2 step (C restatement, OpenMP over cells), first {f_all} frames "
                      f"(1/60 s each) of the 701x701 cut-face terrain: {s_all} substeps in {e_all:.2f} s",
            "one_core": {"value": v_one, "cores": 1,
                         "sample": f"same, one thread: {f_one} frames, {s_one} substeps in {e_one:.2f} s"}}


PARITY_TOL = {"linf_m": 1.0e-3, "rmse_m": 1.0e-5, "v2_mass_rel": 1.0e-10}


def frame_parity(seed: int, frames: int) -> dict:
    """FP32 device frames against the FP64 oracle's frames from the same
    (FP32-rounded, quantised) initial fields: per-frame L-inf / RMSE of h,
    qx, qy, b_eff and H_free, the Mobile V2 mass error |M_t - M_0| / M_0 of
    both, the failure mask against the oracle on the device's own fields
    (must be bit-exact) and against the FP64 sequence, and the substep counts.
    Tolerances: L-inf <= 1e-3 m, RMSE <= 1e-5 m (the reference's terrain
    acceptance, tools/run_terrain_hierarchy_acceptance.py:42-43), mass <= 1e-10."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import oracle
    from frame_sequence import OracleFrameSequence, field_errors
    from isaac_bulk_hip import relax_device

    state, solver, bridge, b, h = build_state("fp32", seed)
    q = 2.0 ** -24
    b0 = b.astype(np.float32).astype(np.float64)
    h0 = np.rint(h / q) * q
    aval = {"tan_start": float(np.tan(np.deg2rad(34.0))), "tan_stop": float(np.tan(np.deg2rad(30.0))),
            "density": 1370.0, "gravity": 9.81, "cohesion": 0.0, "mobilization_depth": 0.08,
            "dry": 1.0e-8, "activity_speed": 0.05, "release": True}
    seq = OracleFrameSequence(b0, h0, dx=DX, dt=FRAME_DT, v2={"limit_mode": 1}, relax_angle_deg=34.0,
                              relax_max_iterations=1000, relax_tolerance=1.0e-6, avalanche=aval)
    quanta0 = int(np.sum(np.rint(h0 / q).astype(np.int64)))
    worst = {}
    mass_dev = mass_orc = 0.0
    mask_fp64 = 0
    mask_own_exact = labels_exact = True
    sub_diff = 0
    for _ in range(frames):
        rec = solver.step_resident(FRAME_DT)
        relax_device(state.runtime.arrays["b_eff"].view(N, N), DX, DX, 34.0, 1000, 1.0e-6, runtime=state.runtime)
        bridge.label()
        ref = seq.step()
        view = state.explicit_host_view(source="acceptance")
        dev = {"h": view.H_mobile_m, "qx": view.mobile_momentum_m2_s[..., 0],
               "qy": view.mobile_momentum_m2_s[..., 1], "b_eff": view.b_eff_m, "H_free": view.H_free_m}
        err = field_errors(dev, {"h": seq.h, "qx": seq.qx, "qy": seq.qy, "b_eff": seq.b, "H_free": seq.b + seq.h})
        for k, v in err.items():
            w = worst.setdefault(k, {"linf": 0.0, "rmse": 0.0})
            w["linf"], w["rmse"] = max(w["linf"], v["linf"]), max(w["rmse"], v["rmse"])
        unstable = state.runtime.arrays["avalanche_unstable"].cpu().numpy().reshape(N, N)
        labels = state.runtime.arrays["avalanche_parent"].cpu().numpy().reshape(N, N)
        own = seq.mask(dev["b_eff"], dev["h"], dev["qx"], dev["qy"])
        mask_own_exact &= bool(np.array_equal(unstable, own["unstable"]))
        labels_exact &= bool(np.array_equal(labels, oracle.label(unstable, 8)))
        mask_fp64 = max(mask_fp64, int(np.count_nonzero(unstable != ref["unstable"])))
        quanta = int(np.sum(np.rint(dev["h"] / q).astype(np.int64)))
        mass_dev = max(mass_dev, abs(quanta - quanta0) / quanta0)
        mass_orc = max(mass_orc, seq.mass_error())
        sub_diff = max(sub_diff, abs(rec.substeps - ref["substeps"]))
    passed = (all(v["linf"] <= PARITY_TOL["linf_m"] and v["rmse"] <= PARITY_TOL["rmse_m"] for v in worst.values())
              and mass_dev <= PARITY_TOL["v2_mass_rel"] and mask_own_exact and labels_exact and sub_diff <= 1)
    state.runtime.close()
    return {"frames": frames, "reference": "FP64 oracle frames (C restatement: V2 step, relaxation, mask, labels)",
            "tolerance": PARITY_TOL, "worst_per_field": worst,
            "v2_mass_error_rel": {"device_fp32": mass_dev, "oracle_fp64": mass_orc},
            "total_volume_error_rel_oracle_fp64": seq.total_error(),
            "mask_bit_exact_on_device_fields": mask_own_exact, "labels_bit_exact": labels_exact,
            "mask_cells_differing_from_fp64_sequence_max": mask_fp64, "substep_count_diff_max": sub_diff,
            "passed": bool(passed)}


def batched_result(count: int, steps: int, warmup: int, seed: int) -> dict:
    """Uneven batch (sizes uniform over {64..512}^2, own pile and mobile patch each),
    1/60 s frames, FP32; value = sum of cells x substeps over the wall time."""
    import torch
    from isaac_bulk_hip import MobileV2Config
    from isaac_bulk_hip.batched import TILE_SHAPES, BatchedMobileV2
    from isaac_bulk_hip.workloads import batched_piles

    terrains = [(b.astype(np.float32), h) for b, h in batched_piles(count, DX, seed=seed)]
    batch = BatchedMobileV2(terrains, DX, config=MobileV2Config(cfl=0.18))
    tune = batch.autotune(FRAME_DT, frames=2)
    for _ in range(warmup):
        batch.step(FRAME_DT)
    cells = (batch.ny.astype(np.int64) * batch.nx.astype(np.int64))
    cell_sub = 0
    sub_ms = 0.0
    launches = 0
    exact_mass = True
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        recs = batch.step(FRAME_DT)
        cell_sub += int(np.dot(cells, recs.substeps.astype(np.int64)))
        sub_ms += batch.last_substep_ms
        launches += batch.last_launches
        exact_mass &= bool(np.array_equal(recs.mass_quanta_before, recs.mass_quanta_after))
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    bpc = BYTES_PER_CELL_SUBSTEP["fp32"]
    achieved = bpc * cell_sub / (sub_ms * 1e-3) / 1e9 if sub_ms > 0 else 0.0
    # a full launch (every tile of every terrain working): a 1 us step takes
    # exactly one substep everywhere; the second such step enqueues that
    # launch plus one that finds every loop ended, both inside the timing
    full = []
    for _ in range(4):
        recs = batch.step(1.0e-6)
        if np.all(recs.substeps == 1):
            full.append((batch.last_substep_ms, batch.last_launches))
    full_ms = min(ms for ms, _ in full[1:]) if len(full) > 1 else None
    full_bytes = bpc * int(cells.sum())
    out = {
        "value": cell_sub / wall, "unit": "cell-substeps/s", "steps": steps, "warmup": warmup,
        "ms_per_step": wall * 1e3 / max(steps, 1), "simulated_s": steps * FRAME_DT,
        "config": {"workload": f"{count} terrains, sizes uniform over {{64..512}}^2, 1/60 s frames",
                   "cells": int(cells.sum()), "frame_dt_s": FRAME_DT, "tile": TILE_SHAPES[tune["best"]]},
        "autotune_ms": tune["ms"], "mass_quanta_exact_every_instance_and_frame": exact_mass,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": MEASURED_TRAFFIC_BATCHED_FULL,
                     "traffic_note": "HBM bytes of one full launch (every instance active)",
                     "kernel": "batch_substep_kernel",
                     "bytes_per_cell_substep": bpc, "substep_launches": launches,
                     "substep_launch_ms_total": sub_ms,
                     "note": "achieved = algorithmic bytes of the instances' substeps / all substep launch time "
                             "(late launches carry few active terrains)"},
        "full_launch_roofline": None if full_ms is None else {
            "bound": "hbm", "achieved": full_bytes / (full_ms * 1e-3) / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": full_bytes / (full_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, "traffic": MEASURED_TRAFFIC_BATCHED_FULL,
            "kernel": "batch_substep_kernel", "algorithmic_bytes": full_bytes, "ms": full_ms,
            "note": "one substep of every terrain; the timing also holds the following launch that finds every "
                    "loop ended (a few microseconds), so the fraction is a lower bound"},
    }
    batch.rt.close()
    return out


def run_batched(args) -> int:
    res = batched_result(args.batch_count, args.steps, args.warmup, args.seed)
    out = {"metric": "batched Mobile V2 cell-substeps/s (uneven terrains)", "value": res.pop("value"),
           "unit": res.pop("unit"), "n_gpus": 1, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "f32", "data": "synthetic (seeded Gaussian piles and mobile patches)", **res}
    print(json.dumps(out), flush=True)
    return 0


def digging_result(count: int, steps: int, warmup: int, seed: int, relax_mode: str,
                   relax_max_iterations: int) -> dict:
    """Digging workload: 64 terrains of 256^2 (Gaussian piles on a 0.2 m floor,
    relaxed to rest at 34 degrees before timing), each with a 1.5 x 0.8 m
    bucket swept 0.02 m per frame along a seeded line;
    per 1/60 s frame and for every terrain at once: entrainment of 0.01 m under
    the bucket (entrain contract), the Mobile V2 step, relaxation of b_eff at
    34 degrees to 1e-6 m (learned warm start + exact sweeps, or plain), failure
    labelling.  Run it with --steps 180 for the 3 s sweep."""
    import torch
    from isaac_bulk_hip import MobileV2Config
    from isaac_bulk_hip.batched import BatchedMobileV2
    from isaac_bulk_hip.workloads import bucket_cells, digging_sites

    n = 256
    terrains, lines = digging_sites(count, n, DX, seed=seed)
    batch = BatchedMobileV2([(b.astype(np.float32), h) for b, h in terrains], DX, config=MobileV2Config(cfl=0.18))
    frames = warmup + steps
    # bucket index lists of every frame, uploaded before timing
    plan = []
    for f in range(frames):
        cells = [bucket_cells(line, f * 0.02, n, DX) for line in lines]
        inst = np.concatenate([np.full(len(c), i, np.int32) for i, c in enumerate(cells)])
        plan.append((torch.from_numpy(inst).to("cuda:0"), torch.from_numpy(np.concatenate(cells)).to("cuda:0"),
                     torch.full((len(inst),), 0.01, dtype=torch.float64, device="cuda:0")))
    # the piles start at rest (relaxed to 34 degrees), then the bucket digs
    rest = batch.relax(34.0, 1e-6, max_iterations=20000)
    net = None
    if relax_mode == "learned":
        from isaac_bulk_hip.warm_start_training import load_net

        net, _ = load_net(batch.rt)
    env_launches = 0
    batch.entrain(*plan[0], check=True)            # validates the index layout once (and digs frame 0)

    def total_volume():
        return float(batch.b.double().sum().item() + batch.h.double().sum().item() * 2.0 ** -24) * DX * DX

    cells_total = int(batch.cells)
    relax_it, relax_max, cell_sub, dug = 0, 0, 0, 0.0
    frame_max = []                     # per frame: the most iterations any terrain needed
    v2_s = 0.0
    t_parts = {"entrain": 0.0, "v2": 0.0, "relax": 0.0, "label": 0.0}
    for f in range(frames):
        timed = f >= warmup
        if f == warmup:
            torch.cuda.synchronize()
            volume0 = total_volume()
            t0 = time.perf_counter()
        ta = time.perf_counter()
        if f > 0:
            vol = batch.entrain(*plan[f], check=False)
            if timed:
                dug += float(vol.sum())
        tb = time.perf_counter()
        recs = batch.step(FRAME_DT)
        tc = time.perf_counter()
        if net is not None:
            ws = batch.warm_start_relax(net, 34.0, 1e-6, max_iterations=relax_max_iterations)
            rr = ws["relax"]
            if timed:
                env_launches += ws["envelope_launches"]
        else:
            rr = batch.relax(34.0, 1e-6, max_iterations=relax_max_iterations)
        td = time.perf_counter()
        _, stats = batch.label(start_angle_deg=38.0, stop_angle_deg=30.0, density=1370.0, cohesion=0.0,
                               mobilization_depth=0.05, dry_tolerance=1e-8, connectivity=8)
        te = time.perf_counter()
        if timed:
            cell_sub += int(np.dot(batch.ny.astype(np.int64) * batch.nx.astype(np.int64),
                                   recs.substeps.astype(np.int64)))
            relax_it += int(rr["iterations"].sum())
            relax_max = max(relax_max, int(rr["iterations"].max()))
            frame_max.append(int(rr["iterations"].max()))
            for k, v in zip(t_parts, (tb - ta, tc - tb, td - tc, te - td)):
                t_parts[k] += v
            v2_s += tc - tb
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    volume1 = total_volume()
    out = {
        "metric": "digging frames/s (entrain + Mobile V2 + relaxation + labelling, 64 terrains)",
        "value": steps / wall, "unit": "frames/s", "n_gpus": 1, "steps": steps, "warmup": warmup,
        "ms_per_step": wall * 1e3 / max(steps, 1), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic (seeded Gaussian piles, seeded bucket lines)",
        "config": {"workload": f"{count} terrains of 256^2, bucket 1.5 x 0.8 m at 0.02 m/frame, "
                               "0.01 m entrained per frame, relaxation 34 deg to 1e-6 m, 1/60 s frames",
                   "cells": cells_total, "frame_dt_s": FRAME_DT},
        "v2_cell_substeps_per_s": cell_sub / v2_s if v2_s > 0 else None,
        "ms_per_frame_parts": {k: v * 1e3 / max(steps, 1) for k, v in t_parts.items()},
        "relax_iterations_per_terrain_frame": relax_it / max(steps * count, 1),
        "relax_iterations_max": relax_max,
        "relax_iterations_frame_max": {"mean": float(np.mean(frame_max)) if frame_max else None,
                                       "p50": float(np.percentile(frame_max, 50)) if frame_max else None,
                                       "p90": float(np.percentile(frame_max, 90)) if frame_max else None},
        "relaxation": relax_mode,
        "envelope_launches_per_frame": env_launches / max(steps, 1),
        "setup_relax_to_rest_iterations_max": int(rest["iterations"].max()),
        "entrained_m3": dug,
        "volume_rel_change": (volume1 - volume0) / volume0,
        "unstable_cells_last_frame": int(stats[:, 3].sum()),
    }
    if net is not None:
        # the network on the final fields of all terrains, after the timed frames
        b3 = batch.b.view(count, n, n)
        h3 = (batch.h.to(torch.float32) * 2.0 ** -24).view(count, n, n)
        out["surrogate_roofline"] = surrogate_roofline(net, b3, h3)
    batch.rt.close()
    return out


def run_digging(args) -> int:
    out = digging_result(args.dig_count, args.steps, args.warmup, args.seed, args.dig_relax,
                         args.relax_max_iterations)
    print(json.dumps(out), flush=True)
    return 0


def run_deposition(args) -> int:
    """Mobile -> Resting deposition step (WarpDepositionOperator.step_resident)
    on the 701x701 cut-face state after 30 Mobile V2 frames (developed depth
    and momentum).  Every call starts from that snapshot, restored outside
    the timed region; value = cells / device time of one call (HIP events on
    the library's stream around the two kernels, the result copy and sync)."""
    import torch
    from isaac_bulk_hip import MaterialScenario, WarpDepositionOperator

    state, solver, bridge, b0, h0 = build_state(args.precision, args.seed)
    for _ in range(30):
        solver.step_resident(FRAME_DT)
    material = MaterialScenario("iron_ore_fines_uncalibrated", 1370.0, 29.8, 0.0, 0.5, 34.0, 30.0, 0.55)
    rt = state.runtime
    op = WarpDepositionOperator(runtime=rt)
    op.bind_device_state(state)
    names = ("b_eff", "mobile", "momentum_x", "momentum_y", "avalanche_m2r_cumulative")
    snap = {k: rt.arrays[k].clone() for k in names}

    def restore():
        for k in names:
            rt.arrays[k].copy_(snap[k])

    for _ in range(args.warmup):
        restore()
        op.step_resident(material, FRAME_DT)
    # cells that the call touches (constant: every call starts from the snapshot)
    restore()
    op.step_resident(material, FRAME_DT)
    dep = rt.arrays["deposition_work"]
    n_dep = int((dep > 0).sum().item())
    m_snap, f = snap["mobile"], rt.arrays["material_mask"]
    n_wet = int(((m_snap > 0) & (f == 0)).sum().item())
    stream = torch.cuda.current_stream()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
          for _ in range(args.steps)]
    deposited = 0.0
    for k in range(args.steps):
        restore()
        ev[k][0].record(stream)
        step = op.step_resident(material, FRAME_DT)
        ev[k][1].record(stream)
        deposited = step.deposited_volume_m3
    torch.cuda.synchronize()
    ms = [a.elapsed_time(b) for a, b in ev]
    avg_ms = float(np.mean(ms))
    cells = N * N
    # algorithmic bytes per call: pass 1 reads b, mobile, forcing and (wet,
    # unforced cells) momentum, writes deposited and removed momentum; pass 2
    # reads deposited and, where material settles, reads and writes b, mobile,
    # momentum and the cumulative field
    f_b = 4 if args.precision == "fp32" else 8
    per_cell = 2 * f_b + 4 + 3 * f_b + f_b
    per_wet = 2 * f_b
    per_dep = 2 * (4 * f_b + 8)
    algo = cells * per_cell + n_wet * per_wet + n_dep * per_dep
    achieved = algo / (avg_ms * 1e-3) / 1e9
    out = {
        "metric": "Mobile -> Resting deposition cells/s",
        "value": cells / (avg_ms * 1e-3), "unit": "cells/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": avg_ms, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32" if args.precision == "fp32" else "f64",
        "data": "synthetic (701x701 cut-face terrain after 30 Mobile V2 frames)",
        "config": {"workload": "701x701 deposition step, 1/60 s, cohesion 0, stop 30 deg",
                   "cells": cells, "wet_unforced_cells": n_wet, "settling_cells": n_dep,
                   "deposited_volume_m3": deposited, "precision": args.precision},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "deposition_decide_kernel + deposition_apply_kernel (one call)",
                     "algorithmic_bytes_per_call": algo, "avg_call_ms": avg_ms,
                     "min_call_ms": float(np.min(ms))},
    }
    if not args.no_cpu_baseline and args.precision == "fp32":
        sys.path.insert(0, str(ROOT / "oracle"))
        import oracle
        from isaac_bulk_hip.config import triangle_vertex_weights

        host = {k: v.cpu().numpy().reshape(N, N) for k, v in snap.items()}
        forcing = f.cpu().numpy().reshape(N, N)
        w = triangle_vertex_weights(N, N, DX, DX)
        calls, t0 = 0, time.perf_counter()
        while time.perf_counter() - t0 < 10.0:
            oracle.deposition(host["b_eff"], host["mobile"], host["momentum_x"], host["momentum_y"],
                              host["avalanche_m2r_cumulative"], w, dx=DX, dy=DX, dt=FRAME_DT,
                              density=1370.0, cohesion=0.0, tan_stop=float(np.tan(np.deg2rad(30.0))),
                              forcing=forcing, fp32=True)
            calls += 1
        elapsed = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": cells * calls / elapsed, "unit": "cells/s", "cores": 1,
                               "kind": "port",
                               "sample": f"oracle FP32-storage deposition step (C restatement, one "
                                         f"thread) on the same snapshot: {calls} calls in {elapsed:.2f} s"}
    print(json.dumps(out), flush=True)
    torch.cuda.synchronize()
    rt.close()
    return 0


def run_frontier(args) -> int:
    """Residual MiniSlope projection (the frame scheduler's compact-frontier
    rounds, v2_physics_core.py:1172-1254) on the 701x701 cut-face terrain
    relaxed to rest at 30 degrees, after 24 seeded landings of 0.3 m; tile 64,
    FP64, vertex weights.  Timed: the device round loop (one call), and the
    host-driven loop of the reference (per colour batch one relax call and
    one scan call, each with a host read of the tile flags) on the same
    input; both end on identical fields."""
    import torch
    from isaac_bulk_hip import (CompactTileFrontier, HipCompactActiveEdgeOperator, neighbor_pairs,
                                relax_device)
    from isaac_bulk_hip.config import triangle_vertex_weights
    from isaac_bulk_hip.workloads import cut_face_terrain

    ts, stop_deg = 64, 30.0
    b, _, _ = cut_face_terrain(N, DX, seed=args.seed)
    rest = torch.from_numpy(np.ascontiguousarray(b)).to("cuda:0")
    relax_device(rest, DX, DX, stop_deg, 20000, 1e-8)
    base = rest.cpu().numpy()
    rng = np.random.default_rng(args.seed + 1)
    cells = rng.choice(N * N, size=24, replace=False)
    disturbed = base.copy().ravel()
    disturbed[cells] += 0.3
    disturbed = disturbed.reshape(N, N)
    reached = np.zeros(N * N, dtype=np.int32)
    reached[cells] = 1
    reached = reached.reshape(N, N)
    w = triangle_vertex_weights(N, N, DX, DX)
    frontier = CompactTileFrontier((N, N), ts)
    seeds = frontier.tiles_around_cells(*np.divmod(cells, N))
    nb = neighbor_pairs(DX, DX)
    tan = float(np.tan(np.deg2rad(stop_deg)))
    op = HipCompactActiveEdgeOperator((N, N), tile_size=ts)

    def device_run():
        op.initialize(disturbed, reached, w)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        res = op.advance_rounds(seeds, neighbors=nb, critical_tangent=tan, tolerance_m=1e-8, max_rounds=4096)
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) * 1e3, res

    def host_run():
        op.initialize(disturbed, reached, w)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        active, rounds = seeds.astype(np.int32), 0
        while active.size:
            rounds += 1
            phase = active
            for d, (di, dj, dist) in enumerate(nb):
                for p in (0, 1):
                    r = op.relax_phase(frontier.edge_batch(phase, direction_index=d, di=di, dj=dj, phase=p),
                                       critical_difference_m=tan * dist)
                    if r.triggered_tile_ids.size:
                        phase = np.union1d(phase, r.triggered_tile_ids).astype(np.int32)
            owners = []
            for d, (di, dj, dist) in enumerate(nb):
                for p in (0, 1):
                    r = op.scan_phase(frontier.edge_batch(phase, direction_index=d, di=di, dj=dj, phase=p),
                                      critical_difference_m=tan * dist, tolerance_m=1e-8)
                    if r.unstable_owner_tile_ids.size:
                        owners.append(r.unstable_owner_tile_ids)
            active = np.unique(np.concatenate(owners)).astype(np.int32) if owners else np.empty(0, np.int32)
        op.runtime.arrays["reached"].zero_()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) * 1e3, rounds

    for _ in range(max(args.warmup, 1)):
        device_run()
        host_run()
    dev_ms, host_ms = [], []
    for _ in range(args.steps):
        ms, res = device_run()
        dev_ms.append(ms)
    h_dev = op.download_state()[0]
    for _ in range(max(1, min(args.steps, 5))):
        ms, host_rounds = host_run()
        host_ms.append(ms)
    h_host = op.download_state()[0]
    out = {
        "metric": "residual projection rounds to rest (device loop vs host-driven loop)",
        "value": float(np.median(dev_ms)), "unit": "ms", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": float(np.median(dev_ms)), "higher_is_better": False,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (701x701 cut-face terrain at rest at 30 deg, 24 landings of 0.3 m)",
        "config": {"workload": "701x701 residual projection, tile 64, 4 directions, tol 1e-8",
                   "rounds": res.rounds, "moved_edges": res.moved_edge_count,
                   "seed_tiles": int(seeds.size)},
        "host_driven_loop_ms": float(np.median(host_ms)), "host_driven_rounds": host_rounds,
        "fields_identical": bool(np.array_equal(h_dev, h_host)),
        "speedup_vs_host_loop": float(np.median(host_ms) / np.median(dev_ms)),
    }
    print(json.dumps(out), flush=True)
    torch.cuda.synchronize()
    op.runtime.close()
    return 0


def main() -> int:
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU sharding is a design "
                                   "note (row strips / per-terrain batches), not a launch mode",
                          "n_gpus": args.gpus}))
        return 2
    import torch

    torch.cuda.set_device(0)
    if args.calibrate_copy:
        src = torch.rand(16 << 20, device="cuda:0")
        dst = torch.empty_like(src)
        for _ in range(10):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"calibrate_copy_bytes_read": src.numel() * 4, "bytes_written": dst.numel() * 4,
                          "launches": 10}))
        return 0
    if args.workload == "batched":
        return run_batched(args)
    if args.workload == "digging":
        return run_digging(args)
    if args.workload == "deposition":
        return run_deposition(args)
    if args.workload == "frontier":
        return run_frontier(args)
    state, solver, bridge, b0, h0 = build_state(args.precision, args.seed)

    # relax-to-rest of the initial terrain (plain relaxation, same tolerance as the frames)
    from isaac_bulk_hip import relax_device

    relax_to_rest_ms, rest_it, rest_conv = None, None, None
    learned = None
    if not args.skip_relax_to_rest:
        for _ in range(2):                 # the second run is reported (the first sizes scratch)
            rest = state.runtime.arrays["b_eff"].view(N, N).clone()
            torch.cuda.synchronize()
            t_rest = time.perf_counter()
            rest_it, rest_conv, _ = relax_device(rest, DX, DX, 34.0, 20000, relax_tol(args.precision),
                                                 runtime=state.runtime)
            torch.cuda.synchronize()
            relax_to_rest_ms = (time.perf_counter() - t_rest) * 1e3
        if args.precision == "fp32":
            learned = learned_relax_to_rest(state, rest)

    for _ in range(args.warmup):
        run_frame(state, solver, bridge, args.precision, args.relax_max_iterations)
    torch.cuda.synchronize()

    total_sub = 0
    total_relax_it = 0
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    loop_s = 0.0         # wall time of the V2 step calls (each returns after a device sync)
    for _ in range(args.steps):
        t_step = time.perf_counter()
        rec = solver.step_resident(FRAME_DT)
        loop_s += time.perf_counter() - t_step
        b = state.runtime.arrays["b_eff"].view(N, N)
        it, _, _ = relax_device(b, DX, DX, 34.0, args.relax_max_iterations, relax_tol(args.precision),
                                runtime=state.runtime)
        bridge.label()
        state.advance_time(FRAME_DT)
        total_sub += rec.substeps
        total_relax_it += it
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0

    # Roofline of the substep kernel: a second run of frames in timing mode,
    # where each step brackets its working substep launches with one HIP event
    # pair on the library's stream (frames whose substep count differed from
    # the bracket report 0 and are left out; timing stays out of the run that
    # gives `value`).
    sub_ms = 0.0
    timed_sub = 0
    timed_frames = 0
    v2_ms = 0.0          # device time of the whole V2 step calls (HIP events inside mv2_step)
    v2_sub = 0
    v2_frames = min(args.steps, 60)
    solver.set_kernel_timing(True)
    for _ in range(v2_frames):
        subs, _, _ = run_frame(state, solver, bridge, args.precision, args.relax_max_iterations)
        v2_ms += solver.last_record.kernel_ms
        v2_sub += subs
        if solver.last_record.substep_kernel_ms > 0.0:
            sub_ms += solver.last_record.substep_kernel_ms
            timed_sub += solver.last_record.timed_launches
            timed_frames += 1
    solver.set_kernel_timing(False)
    torch.cuda.synchronize()

    cells = N * N
    value = cells * total_sub / loop_s
    bpc = BYTES_PER_CELL_SUBSTEP[args.precision]
    avg_substep_ms = sub_ms / max(timed_sub, 1)
    achieved = bpc * cells / (avg_substep_ms * 1e-3) / 1e9 if avg_substep_ms > 0 else 0.0
    out = {
        "metric": "Mobile V2 cell-substeps/s + relax-to-rest ms, learned vs plain @ same error",
        "value": value,
        "unit": "cell-substeps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / max(args.steps, 1),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if args.precision == "fp32" else "f64",
        "data": "synthetic (seeded value-noise terrain with a 40 degree cut face)",
        "config": {"workload": "701x701 cut-face frame: Mobile V2 + relaxation 34 deg + labelling",
                   "grid": [N, N], "dx_m": DX, "frame_dt_s": FRAME_DT,
                   "simulated_s": args.steps * FRAME_DT, "precision": args.precision},
        "substeps_total": total_sub,
        "frame_cell_substeps_per_s": cells * total_sub / wall,
        "v2_step_device_cell_substeps_per_s": cells * v2_sub / (v2_ms * 1e-3) if v2_ms > 0 else None,
        "v2_step_device_ms_per_frame": v2_ms / max(v2_frames, 1),
        "relax_to_rest_ms": relax_to_rest_ms,
        "relax_to_rest_iterations": rest_it,
        "relax_to_rest_converged": rest_conv,
        "relax_iterations_timed": total_relax_it,
        "learned_relax_to_rest_ms": learned["ms"] if learned else None,
        "learned_relax_to_rest": learned,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": MEASURED_TRAFFIC_BYTES.get(args.precision),
                     "kernel": "v2_substep_kernel", "bytes_per_cell_substep": bpc,
                     "cells_per_launch": cells, "avg_launch_ms": avg_substep_ms,
                     "launches_timed": timed_sub, "frames_timed": timed_frames,
                     "note": "one HIP event pair per step around its substep launches after launch 0 "
                             "(which also runs the begin pass): the launch that checks the CFL "
                             "prediction, plain working launches and the substep that ends the loop "
                             "and stores the closing ledger, gaps between dependent launches included"},
    }
    if args.precision == "fp32" and not args.skip_configs:
        out["c3_batched"] = batched_result(256, 120, 3, args.seed)          # 2 s simulated
        dig = {mode: digging_result(64, 180, 3, args.seed, mode, args.relax_max_iterations)
               for mode in ("learned", "plain")}                            # 3 s simulated each
        out["c5_digging"] = {
            "workload": dig["learned"]["config"]["workload"],
            "frames": 180, "unit": "frames/s",
            "learned_frames_per_s": dig["learned"]["value"], "plain_frames_per_s": dig["plain"]["value"],
            "learned_over_plain": dig["learned"]["value"] / dig["plain"]["value"],
            "volume_rel_change": {m: dig[m]["volume_rel_change"] for m in dig},
            "relax_iterations_per_terrain_frame": {m: dig[m]["relax_iterations_per_terrain_frame"] for m in dig},
            "v2_cell_substeps_per_s": dig["learned"]["v2_cell_substeps_per_s"],
            "ms_per_frame_parts": {m: dig[m]["ms_per_frame_parts"] for m in dig},
            "surrogate_roofline": dig["learned"].get("surrogate_roofline"),
        }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(b0, h0, args.cpu_sample_frames)
        if args.precision == "fp32" and args.parity_frames > 0:
            out["parity"] = frame_parity(args.seed, args.parity_frames)
    print(json.dumps(out), flush=True)
    # release device resources before interpreter teardown (profilers hook exit)
    from isaac_bulk_hip.runtime import close_all_runtimes

    close_all_runtimes()
    return 0


if __name__ == "__main__":
    sys.exit(main())
