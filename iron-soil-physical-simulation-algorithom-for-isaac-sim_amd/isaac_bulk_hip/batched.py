"""Batched Mobile V2: many independent terrains of uneven size on one GPU.

Each instance keeps the semantics of ``WarpMobileV2ReferenceSolver.step`` /
``step_resident`` (its own CFL substep loop, static early exit and limits,
pkg/experimental/mobile_v2_reference.py:226-307) but all instances advance in
the same launches: the fields of every instance are packed back to back in one
set of device arrays (FP32 b_eff and momentum, int32 mobile quanta of 2^-24 m),
one launch per substep covers the tiles of the whole batch, and eight substep
launches are replayed from a captured HIP graph.  ``autotune`` picks the tile
shape of the substep kernel by timing the candidates on a copy of the batch.

Records per instance: substeps, CFL, exact mobile quanta before/after,
momentum and kinetic energy before/after (uniform cell weights, per unit
density).
"""

from __future__ import annotations

import ctypes
import json
from dataclasses import dataclass
from pathlib import Path

import numpy as np

from . import _lib
from .bulk_state import MOBILE_QUANTUM_M
from .config import MobileV2Config
from .mobile_v2 import _step_config
from .runtime import HipRuntime

TILE_SHAPES = {0: (8, 62), 1: (16, 62), 2: (32, 62), 3: (64, 62)}    # rows x columns


def _torch():
    import torch

    return torch


@dataclass
class BatchStepRecord:
    substeps: int
    numeric_error: int
    static_exit: bool
    maximum_cfl: float
    remaining: float
    mass_quanta_before: int
    mass_quanta_after: int
    momentum_before: tuple
    momentum_after: tuple
    kinetic_before: float
    kinetic_after: float


# numpy view of mv2_batch_record (include/mv2.h)
_RECORD_DTYPE = np.dtype({
    "names": ["substeps", "numeric_error", "static_exit", "complete", "maximum_cfl", "remaining",
              "mass_quanta_before", "mass_quanta_after", "momentum_before", "momentum_after",
              "kinetic_before", "kinetic_after"],
    "formats": ["<i4", "<i4", "<i4", "<i4", "<f8", "<f8", "<f8", "<f8", ("<f8", 2), ("<f8", 2), "<f8", "<f8"],
    "offsets": [getattr(_lib.BatchRecord, f).offset for f, _ in _lib.BatchRecord._fields_],
    "itemsize": ctypes.sizeof(_lib.BatchRecord),
})


class BatchRecords:
    """The records of one batched step: a sequence of BatchStepRecord, built on
    access, over a numpy array whose columns (``substeps``, ``mass_quanta_before``
    ...) are available as attributes without building the records."""

    def __init__(self, table: np.ndarray) -> None:
        self.table = table

    def __len__(self) -> int:
        return int(self.table.shape[0])

    def __getitem__(self, i: int) -> BatchStepRecord:
        r = self.table[i]
        return BatchStepRecord(int(r["substeps"]), int(r["numeric_error"]), bool(r["static_exit"]),
                               float(r["maximum_cfl"]), float(r["remaining"]), int(r["mass_quanta_before"]),
                               int(r["mass_quanta_after"]), tuple(float(v) for v in r["momentum_before"]),
                               tuple(float(v) for v in r["momentum_after"]), float(r["kinetic_before"]),
                               float(r["kinetic_after"]))

    def __iter__(self):
        return (self[i] for i in range(len(self)))

    def __getattr__(self, name: str) -> np.ndarray:
        if name in _RECORD_DTYPE.names:
            return self.table[name]
        raise AttributeError(name)


def to_quanta(h: np.ndarray) -> np.ndarray:
    q = np.rint(np.asarray(h, dtype=np.float64) * (1.0 / MOBILE_QUANTUM_M))
    if np.any(q < 0) or np.any(q > np.iinfo(np.int32).max):
        raise ValueError("[BatchedMobileV2] mobile thickness outside 0 .. 128 m")
    return q.astype(np.int32)


class BatchedMobileV2:
    """Independent terrains advanced together; ``terrains`` is a list of (b_eff, h_mobile)."""

    backend_name = "GPU_HIP_MOBILE_V2_BATCHED"

    def __init__(self, terrains, dx: float, dy: float | None = None, *, runtime: HipRuntime | None = None,
                 config: MobileV2Config | None = None, use_graph: bool = False, tile_shape: int = 0) -> None:
        torch = _torch()
        self.rt = runtime if runtime is not None else HipRuntime("cuda:0")
        self.dx = float(dx)
        self.dy = float(dx if dy is None else dy)
        self.config = config or MobileV2Config(dx_m=self.dx, dy_m=self.dy)
        self.use_graph = bool(use_graph)
        shapes = [np.asarray(b).shape for b, _ in terrains]
        self.ny = np.asarray([s[0] for s in shapes], np.int32)
        self.nx = np.asarray([s[1] for s in shapes], np.int32)
        sizes = self.ny.astype(np.int64) * self.nx.astype(np.int64)
        self.offset = np.concatenate([[0], np.cumsum(sizes)[:-1]]).astype(np.int64)
        self.cells = int(sizes.sum())
        b = np.concatenate([np.asarray(t[0], np.float32).ravel() for t in terrains])
        h = np.concatenate([to_quanta(t[1]).ravel() for t in terrains])
        dev = self.rt.device
        self.b = torch.from_numpy(b).to(dev)
        self.h = torch.from_numpy(h).to(dev)
        self.qx = torch.zeros(self.cells, dtype=torch.float32, device=dev)
        self.qy = torch.zeros(self.cells, dtype=torch.float32, device=dev)
        self.tile_shape = int(tile_shape)
        self._token = object()            # marks the runtime's context as bound to this batch
        self._bind(self.b, self.h, self.qx, self.qy)
        self.last_kernel_ms = 0.0

    @property
    def count(self) -> int:
        return int(self.ny.size)

    def _bind(self, b, h, qx, qy) -> None:
        rt = self.rt
        rt.bind_stream()
        rt.check(rt.lib.mv2_batch_bind(rt.ctx, self.count, self.ny.ctypes.data, self.nx.ctypes.data,
                                       self.offset.ctypes.data, self.dx, self.dy, b.data_ptr(), h.data_ptr(),
                                       qx.data_ptr(), qy.data_ptr()))
        rt.check(rt.lib.mv2_batch_set_tile_shape(rt.ctx, self.tile_shape))
        rt.batch_owner = self._token

    def _claim(self) -> None:
        """A runtime's context holds one bound batch: rebind this one when
        another batch on the same runtime was bound since (the batches' loop
        state restarts; fields live in each batch's own tensors)."""
        if getattr(self.rt, "batch_owner", None) is not self._token:
            self._bind(self.b, self.h, self.qx, self.qy)

    def set_tile_shape(self, shape: int) -> None:
        if shape not in TILE_SHAPES:
            raise ValueError(f"[BatchedMobileV2] unknown tile shape {shape}")
        self._claim()
        self.tile_shape = int(shape)
        self.rt.check(self.rt.lib.mv2_batch_set_tile_shape(self.rt.ctx, self.tile_shape))

    def step(self, dt: float, *, limit_mode: int = 0) -> BatchRecords:
        """Advance every instance by dt; one record per instance."""
        self._claim()
        return self._step(dt, limit_mode)

    def _step(self, dt: float, limit_mode: int = 0) -> BatchRecords:
        rt = self.rt
        rt.bind_stream()
        cfg = _step_config(self.config, dt, density=1.0, limit_mode=limit_mode, weight_mode=0)
        table = np.zeros(self.count, _RECORD_DTYPE)
        ms = (ctypes.c_double * 3)()
        rc = rt.lib.mv2_batch_step(rt.ctx, ctypes.byref(cfg), int(self.use_graph), table.ctypes.data, ms)
        self.last_kernel_ms, self.last_substep_ms, self.last_launches = ms[0], ms[1], int(ms[2])
        if rc == _lib.MV2_ENUMERIC:
            raise RuntimeError(_lib.last_error(rt.ctx))
        rt.check(rc)
        return BatchRecords(table)

    # ---- relaxation, exchange and labelling of every instance at once -------
    def relax(self, critical_angle_deg: float, tolerance: float, *, max_iterations: int = 20000,
              field=None) -> np.ndarray:
        """Angle-of-repose relaxation of every instance (of b_eff, or of `field`
        laid out like it), each to its own convergence as if relaxed alone
        (relax_critical_slope, slope_model.py:273-343).  Returns a structured
        array per instance: iterations, converged, max_excess_last."""
        from .relaxation import legacy_critical_differences

        self._claim()
        rt = self.rt
        rt.bind_stream()
        target = self.b if field is None else field
        if target.dtype != _torch().float32 or target.numel() != self.cells or not target.is_contiguous():
            raise ValueError("[BatchedMiniSlope] field must be a contiguous float32 array of the batch's cells")
        crit = (ctypes.c_double * 4)(*legacy_critical_differences(self.dx, self.dy, critical_angle_deg))
        out = (_lib.RelaxResult * self.count)()
        rt.check(rt.lib.mv2_batch_relax(rt.ctx, target.data_ptr(), crit, int(max_iterations), float(tolerance),
                                        out))
        res = np.zeros(self.count, [("iterations", "<i4"), ("converged", "?"), ("max_excess_last", "<f8")])
        for i, r in enumerate(out):
            res[i] = (r.iterations, bool(r.converged), r.max_excess_last)
        return res

    def warm_start_relax(self, net, critical_angle_deg: float, tolerance: float, *, max_iterations: int = 20000,
                         eps: float = 1e-5, envelope_angle_deg: float | None = None, settle_rounds: int = 3) -> dict:
        """Learned warm start of the relaxation of b_eff for a batch of
        equal-size terrains: the network on all of them at once, the Lipschitz
        envelope of every predicted surface, the mass projection per terrain
        with its settle rounds (as ``surrogate.warm_start_relax``), then the
        exact relaxation of every terrain to `tolerance` (each to its own
        convergence, as ``relax``).  Batched, the relaxation runs until its
        slowest terrain is at rest, so the settle rounds pay off: digging 64 x
        256^2, 22 exact iterations per terrain and frame and 245.6 frames/s
        without them, 1.0 and 420.3 frames/s with two."""
        from .relaxation import legacy_critical_differences

        torch = _torch()
        if len(set(zip(self.ny.tolist(), self.nx.tolist()))) != 1:
            raise ValueError("[BatchedWarmStart] the terrains of the batch must have equal sizes")
        self._claim()
        rt = self.rt
        rt.bind_stream()
        B, H, W = self.count, int(self.ny[0]), int(self.nx[0])
        b3 = self.b.view(B, H, W)
        h3 = (self.h.to(torch.float32) * MOBILE_QUANTUM_M).view(B, H, W)
        pred = net.forward(b3, h3, self.dx, self.dy)
        crit = legacy_critical_differences(self.dx, self.dy,
                                           critical_angle_deg if envelope_angle_deg is None else envelope_angle_deg)
        launches = (ctypes.c_int * 1)()
        rt.check(rt.lib.mv2_batch_envelope(rt.ctx, self.b.data_ptr(), pred.data_ptr(), pred.data_ptr(), B, H, W,
                                           crit.ctypes.data, 256, launches))
        settle = (ctypes.c_int * 1)()
        rt.check(rt.lib.mv2_batch_commit_settled(rt.ctx, self.b.data_ptr(), pred.data_ptr(), B, H, W,
                                                 crit.ctypes.data, eps, int(settle_rounds), 256, settle))
        res = self.relax(critical_angle_deg, tolerance, max_iterations=max_iterations)
        return {"relax": res, "envelope_launches": int(launches[0]), "settle_launches": int(settle[0])}

    def entrain(self, instance, cell, depth_m, *, z_base=None, check: bool = True) -> np.ndarray:
        """The entrain contract (bulk_state_authority.py:104-127) at (instance,
        instance-local flat cell) pairs; returns the volume moved per instance
        (m^3, whole quanta, uniform cell area).  (instance, cell) pairs must be
        unique.  Inputs may be host arrays or device tensors; `check=False`
        skips the range and uniqueness checks (which read the indices back)
        for inputs already validated."""
        torch = _torch()
        self._claim()
        rt = self.rt
        rt.bind_stream()
        dev = rt.device

        def dev_tensor(x, dtype):
            if isinstance(x, torch.Tensor):
                return x.to(device=dev, dtype=dtype).contiguous()
            return torch.as_tensor(np.asarray(x), dtype=dtype, device=dev)

        inst = dev_tensor(instance, torch.int32)
        cells = dev_tensor(cell, torch.int32)
        depth = dev_tensor(depth_m, torch.float64)
        n = int(inst.numel())
        if cells.numel() != n or depth.numel() != n:
            raise ValueError("[BatchedDeviceBulkState] instance, cell and depth must have the same length")
        if check and n and (int(inst.min()) < 0 or int(inst.max()) >= self.count):
            raise ValueError("[BatchedDeviceBulkState] instance index out of range")
        if check and n:
            sizes = torch.as_tensor((self.ny.astype(np.int64) * self.nx.astype(np.int64)), device=dev)
            if bool((cells < 0).any()) or bool((cells.to(torch.int64) >= sizes[inst.to(torch.int64)]).any()):
                raise ValueError("[BatchedDeviceBulkState] cell index out of range")
            # the kernel updates each (instance, cell) with a plain read-modify-write
            key = inst.to(torch.int64) * int(sizes.max().item()) + cells.to(torch.int64)
            if int(torch.unique(key).numel()) != n:
                raise ValueError("[BatchedDeviceBulkState] duplicate (instance, cell) entrainment pairs")
        zb = 0 if z_base is None else z_base.data_ptr()
        vol = np.zeros(self.count, np.float64)
        rt.check(rt.lib.mv2_batch_entrain(rt.ctx, zb, inst.data_ptr(), cells.data_ptr(), depth.data_ptr(), n,
                                          vol.ctypes.data))
        return vol

    def label(self, *, start_angle_deg: float, stop_angle_deg: float, density: float, cohesion: float,
              mobilization_depth: float, dry_tolerance: float = 1e-8, connectivity: int = 8,
              gravity: float = 9.81, z_base=None, labels=None):
        """Failure mask (gpu_large_avalanche.py:59-162) and connected components
        (:165-252) of every instance.  Returns (labels, stats): labels is the
        packed int32 array (instance-local root index, -1 outside the mask),
        stats an int32 [count, 4] array of region count, largest component,
        winner root (-1 if none) and unstable cells per instance."""
        torch = _torch()
        self._claim()
        rt = self.rt
        rt.bind_stream()
        if labels is None:
            labels = torch.empty(self.cells, dtype=torch.int32, device=rt.device)
        start, stop = float(np.deg2rad(start_angle_deg)), float(np.deg2rad(stop_angle_deg))
        cfg = _lib.AvalancheConfig(tan_start=float(np.tan(start)), tan_stop=float(np.tan(stop)), start_rad=start,
                                   stop_rad=stop, density=float(density), gravity=float(gravity),
                                   cohesion=float(cohesion), mobilization_depth=float(mobilization_depth),
                                   dry_tolerance=float(dry_tolerance), activity_speed=0.0,
                                   connectivity=int(connectivity), release_settled_latches=0, weight_mode=0,
                                   reserved=0)
        stats = np.zeros((self.count, 4), np.int32)
        zb = 0 if z_base is None else z_base.data_ptr()
        rt.check(rt.lib.mv2_batch_label(rt.ctx, zb, ctypes.byref(cfg), labels.data_ptr(), stats.ctypes.data))
        return labels, stats

    def fields(self, i: int):
        """(b_eff, h_mobile [m], qx, qy) of instance i as numpy arrays."""
        o, n = int(self.offset[i]), int(self.ny[i]) * int(self.nx[i])
        shape = (int(self.ny[i]), int(self.nx[i]))
        b = self.b[o:o + n].cpu().numpy().reshape(shape)
        h = self.h[o:o + n].cpu().numpy().reshape(shape).astype(np.float64) * MOBILE_QUANTUM_M
        return b, h, self.qx[o:o + n].cpu().numpy().reshape(shape), self.qy[o:o + n].cpu().numpy().reshape(shape)

    def autotune(self, dt: float, *, shapes=tuple(TILE_SHAPES), frames: int = 2, cache: Path | None = None) -> dict:
        """Time each tile shape on a copy of the batch; keep the fastest (and cache it)."""
        torch = _torch()
        key = f"{self.count}:{int(self.ny.sum())}:{int(self.cells)}:{dt:.6g}"
        if cache is not None and Path(cache).exists():
            known = json.loads(Path(cache).read_text())
            if key in known:
                self.set_tile_shape(int(known[key]["best"]))
                return known[key]
        times = {}
        b, h, qx, qy = self.b.clone(), self.h.clone(), self.qx.clone(), self.qy.clone()
        try:
            for s in shapes:
                self.tile_shape = int(s)
                self._bind(b, h, qx, qy)          # the copies keep their addresses across frames
                for _ in range(frames + 1):        # the first frame also captures the graph
                    b.copy_(self.b); h.copy_(self.h); qx.copy_(self.qx); qy.copy_(self.qy)
                    torch.cuda.synchronize()
                    self._step(dt)
                    times.setdefault(s, []).append(self.last_kernel_ms * 1e-3)
        finally:
            best = min(times, key=lambda s: min(times[s][1:] or times[s]))
            self.tile_shape = int(best)
            self._bind(self.b, self.h, self.qx, self.qy)
        result = {"best": int(best), "tile": TILE_SHAPES[int(best)],
                  "ms": {int(s): 1e3 * min(v[1:] or v) for s, v in times.items()}}
        if cache is not None:
            known = json.loads(Path(cache).read_text()) if Path(cache).exists() else {}
            known[key] = result
            Path(cache).write_text(json.dumps(known, indent=1))
        return result
