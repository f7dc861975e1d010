"""Device runtime: named torch tensors + one libmv2 context per runtime.

Replaces WarpRuntime (pkg/performance/warp_backend.py:41-190 in the reference):
the runtime owns every named device allocation (``arrays``), counts host/device
transfers, and hands raw pointers plus the current HIP stream to the C-ABI.
"""

from __future__ import annotations

import atexit
import ctypes
import sys
import weakref
from dataclasses import dataclass

import numpy as np

from . import _lib
from ._lib import HipBackendUnavailable


@dataclass
class TransferTelemetry:
    """Host/device traffic counters (WarpTransferTelemetry equivalent)."""

    h2d_bytes: int = 0
    d2h_bytes: int = 0
    h2d_transfer_count: int = 0
    d2h_transfer_count: int = 0
    synchronization_count: int = 0
    kernel_launch_count: int = 0

    def record_h2d(self, value) -> None:
        self.h2d_bytes += int(np.asarray(value).nbytes)
        self.h2d_transfer_count += 1

    def record_d2h(self, value) -> None:
        self.d2h_bytes += int(np.asarray(value).nbytes)
        self.d2h_transfer_count += 1

    def record_d2h_bytes(self, nbytes: int) -> None:
        self.d2h_bytes += int(nbytes)
        self.d2h_transfer_count += 1

    def record_call(self, launches: int = 1, syncs: int = 1) -> None:
        self.kernel_launch_count += launches
        self.synchronization_count += syncs

    def reset(self) -> None:
        self.h2d_bytes = self.d2h_bytes = 0
        self.h2d_transfer_count = self.d2h_transfer_count = 0
        self.synchronization_count = self.kernel_launch_count = 0


def _torch():
    import torch  # local import: the CPU test suite never needs a device

    return torch


def _raw_stream(device_index: int) -> int:
    """Handle of torch's current HIP stream on a device (the raw query avoids
    building a Stream object on every call)."""
    torch = _torch()
    raw = getattr(torch._C, "_cuda_getCurrentRawStream", None)
    if raw is not None:
        return int(raw(device_index))
    return int(torch.cuda.current_stream(device_index).cuda_stream)


# Live runtimes in creation order.  At interpreter exit they are closed by an
# atexit hook registered when the first runtime is created -- after torch was
# imported, so it runs before torch's own exit handlers and while the HIP
# runtime and the borrowed torch streams still exist.  Closing from __del__
# during interpreter finalisation instead (module globals torn down in
# arbitrary order, after torch/HIP teardown) synchronised a stream handle that
# no longer existed and freed device memory after the runtime's own teardown;
# that is what crashed profiled runs in exit/__cxa_finalize.
_LIVE: list = []
_EXIT_HOOK = False


def close_all_runtimes() -> None:
    """Close every live runtime, newest first, after draining the device."""
    live = [r() for r in reversed(_LIVE)]
    _LIVE.clear()
    live = [r for r in live if r is not None and getattr(r, "ctx", None)]
    if not live:
        return
    try:
        _torch().cuda.synchronize()
    except Exception:
        pass
    for rt in live:
        try:
            rt.close()
        except Exception:
            pass


def _register(rt) -> None:
    global _EXIT_HOOK
    if not _EXIT_HOOK:
        atexit.register(close_all_runtimes)
        _EXIT_HOOK = True
    _LIVE.append(weakref.ref(rt))


class HipRuntime:
    """Named device arrays on one MI355X plus the libmv2 context that computes on them."""

    backend_name = "GPU_HIP_MI355X"

    def __init__(self, device: str | int = "cuda:0") -> None:
        torch = _torch()
        if not torch.cuda.is_available():
            raise HipBackendUnavailable("HIP backend unavailable: torch sees no ROCm GPU")
        self.device = torch.device(device if isinstance(device, str) else f"cuda:{device}")
        index = self.device.index if self.device.index is not None else torch.cuda.current_device()
        self.lib = _lib.load()
        handle = ctypes.c_void_p()
        rc = self.lib.mv2_create(int(index), ctypes.byref(handle))
        if rc != _lib.MV2_OK:
            raise HipBackendUnavailable(_lib.last_error(None))
        self.ctx = handle.value
        self._device_index = int(index)
        self._bound_stream = None
        self.arrays: dict[str, object] = {}
        self.telemetry = TransferTelemetry()
        _register(self)

    # -- stream / sync ---------------------------------------------------------
    def bind_stream(self) -> None:
        """Hand torch's current stream on this device to the library (only
        when it changed since the last call)."""
        stream = _raw_stream(self._device_index)
        if stream != self._bound_stream:
            self.lib.mv2_set_stream(self.ctx, ctypes.c_void_p(stream))
            self._bound_stream = stream

    def synchronize(self) -> None:
        _torch().cuda.synchronize(self.device)
        self.telemetry.synchronization_count += 1

    def check(self, rc: int, *, value_error=ValueError, runtime_error=RuntimeError) -> None:
        _lib.raise_for(rc, self.ctx, value_error=value_error, runtime_error=runtime_error)

    # -- named arrays ----------------------------------------------------------
    def upload(self, name: str, values, dtype) -> object:
        torch = _torch()
        host = np.ascontiguousarray(values)
        tensor = torch.from_numpy(host).to(device=self.device, dtype=dtype).contiguous()
        self.arrays[name] = tensor
        self.telemetry.record_h2d(host)
        return tensor

    def zeros(self, name: str, size: int, dtype) -> object:
        tensor = _torch().zeros(int(size), dtype=dtype, device=self.device)
        self.arrays[name] = tensor
        return tensor

    def empty(self, name: str, size: int, dtype) -> object:
        tensor = _torch().empty(int(size), dtype=dtype, device=self.device)
        self.arrays[name] = tensor
        return tensor

    def download(self, name: str) -> np.ndarray:
        host = self.arrays[name].detach().cpu().numpy().copy()
        self.telemetry.record_d2h(host)
        return host

    def release(self, name: str) -> None:
        self.arrays.pop(name, None)

    def ptr(self, name_or_tensor) -> int:
        t = self.arrays[name_or_tensor] if isinstance(name_or_tensor, str) else name_or_tensor
        return int(t.data_ptr())

    def close(self) -> None:
        """Release the library context: the stream handle is dropped first (the
        library then drains the whole device), then its scratch is freed."""
        if getattr(self, "ctx", None):
            ctx, self.ctx = self.ctx, None
            self._bound_stream = None
            self.lib.mv2_set_stream(ctx, None)
            self.lib.mv2_destroy(ctx)

    def __del__(self) -> None:  # pragma: no cover - interpreter shutdown order
        # Once the interpreter is finalising, HIP and torch may already be torn
        # down: make no device call (the driver reclaims the memory at exit).
        if sys.is_finalizing():
            self.ctx = None
            return
        try:
            self.close()
        except Exception:
            pass

    def diagnostics(self, backend_name: str | None = None) -> dict[str, object]:
        return {
            "backend_name": backend_name or self.backend_name,
            "device": str(self.device),
            "resident_array_names": sorted(self.arrays),
            "h2d_bytes": self.telemetry.h2d_bytes,
            "d2h_bytes": self.telemetry.d2h_bytes,
        }
