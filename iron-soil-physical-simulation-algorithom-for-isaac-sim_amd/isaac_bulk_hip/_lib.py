"""ctypes binding of libmv2.so (declarations in include/mv2.h).

The HIP library is the only compute path: if it is missing or cannot find a
GPU, the call raises -- there is no CPU fallback.
"""

from __future__ import annotations

import ctypes
from pathlib import Path

_HERE = Path(__file__).resolve().parent
LIB_PATH = _HERE / "libmv2.so"

_D = ctypes.c_double
_I = ctypes.c_int
_I32 = ctypes.c_int32
_P = ctypes.c_void_p

MV2_OK, MV2_EINVAL, MV2_ENUMERIC, MV2_EDEVICE, MV2_ESTATE = 0, 1, 2, 3, 4
MV2_FP64, MV2_FP32 = 0, 1


class HipBackendUnavailable(RuntimeError):
    """The HIP library or a GPU is unavailable (WarpBackendUnavailable equivalent)."""


class StepConfig(ctypes.Structure):
    _fields_ = [
        ("dt", _D), ("gravity", _D), ("earth_pressure", _D), ("cfl", _D),
        ("dry_tolerance", _D), ("basal_friction", _D), ("tan_start", _D),
        ("external_ax", _D), ("external_ay", _D), ("density", _D),
        ("maximum_substeps", _I), ("limit_mode", _I), ("weight_mode", _I), ("static_exit", _I),
        ("diagnostics", _I),
    ]


class StepRecord(ctypes.Structure):
    _fields_ = [
        ("status", _I), ("substeps", _I), ("cfl_limited", _I), ("static_exit", _I),
        ("numeric_error", _I), ("complete", _I),
        ("maximum_cfl", _D), ("remaining", _D),
        ("summary_before", _D * 8), ("summary_after", _D * 8), ("diag", _D * 10),
        ("mass_quanta_before", _D), ("mass_quanta_after", _D), ("kernel_ms", _D),
        ("substep_kernel_ms", _D), ("launches", _I), ("timed_launches", _I),
    ]


class BatchRecord(ctypes.Structure):
    _fields_ = [
        ("substeps", _I), ("numeric_error", _I), ("static_exit", _I), ("complete", _I),
        ("maximum_cfl", _D), ("remaining", _D), ("mass_quanta_before", _D), ("mass_quanta_after", _D),
        ("momentum_before", _D * 2), ("momentum_after", _D * 2), ("kinetic_before", _D), ("kinetic_after", _D),
    ]


class RelaxResult(ctypes.Structure):
    _fields_ = [("iterations", _I), ("converged", _I), ("max_excess_last", _D), ("complete", _I),
                ("launches", _I)]


class AvalancheConfig(ctypes.Structure):
    _fields_ = [
        ("tan_start", _D), ("tan_stop", _D), ("start_rad", _D), ("stop_rad", _D),
        ("density", _D), ("gravity", _D), ("cohesion", _D), ("mobilization_depth", _D),
        ("dry_tolerance", _D), ("activity_speed", _D),
        ("connectivity", _I), ("release_settled_latches", _I), ("weight_mode", _I), ("reserved", _I),
    ]


class AvalancheFields(ctypes.Structure):
    _fields_ = [(name, _P) for name in (
        "unstable", "labels", "latch", "previous", "activation_count", "counts",
        "owned_surface", "owned_export_baseline", "export_cumulative", "slope", "grad_x", "grad_y",
        "severity", "first_activation_time", "last_activation_time", "r2m_cumulative")]


class LabelResult(ctypes.Structure):
    _fields_ = [("diag_i", _I32 * 8), ("diag_f", _D * 8), ("union_passes", _I), ("reserved", _I)]


class DepositionConfig(ctypes.Structure):
    _fields_ = [
        ("dt", _D), ("settling_rate", _D), ("speed_threshold", _D), ("density", _D),
        ("gravity", _D), ("cohesion", _D), ("tan_stop", _D),
        ("settle_subcell_tail", _I), ("weight_mode", _I),
    ]


class FrontierRoundConfig(ctypes.Structure):
    _fields_ = [
        ("rows", _I), ("cols", _I), ("tile_size", _I), ("tiles_x", _I), ("tile_count", _I),
        ("n_directions", _I), ("di", _I * 4), ("dj", _I * 4), ("critical_difference", _D * 4),
        ("tolerance", _D), ("max_rounds", _I), ("reserved", _I),
    ]


class FrontierRoundResult(ctypes.Structure):
    _fields_ = [("rounds", _I), ("converged", _I), ("active_tiles", _I), ("cleared_reached", _I),
                ("moved_edges", ctypes.c_longlong)]


class ActivationResult(ctypes.Structure):
    _fields_ = [("activated_volume", _D), ("forcing_count", _I), ("reserved", _I)]


class BatchDesc(ctypes.Structure):
    _fields_ = [
        ("ny", _I), ("nx", _I), ("offset", ctypes.c_int64),
    ]


# exported symbol -> (restype, argtypes); the CPU test suite checks that every
# symbol declared in include/mv2.h is exported
SIGNATURES: dict[str, tuple[object, list[object]]] = {
    "mv2_abi_version": (_I, []),
    "mv2_mobile_quantum": (_D, []),
    "mv2_create": (_I, [_I, ctypes.POINTER(_P)]),
    "mv2_destroy": (None, [_P]),
    "mv2_last_error": (ctypes.c_char_p, [_P]),
    "mv2_set_stream": (_I, [_P, _P]),
    "mv2_synchronize": (_I, [_P]),
    "mv2_bind": (_I, [_P, _I, _I, _I, _D, _D, _P, _P, _P, _P, _P]),
    "mv2_step": (_I, [_P, ctypes.POINTER(StepConfig), ctypes.POINTER(StepRecord)]),
    "mv2_summary": (_I, [_P, ctypes.POINTER(StepConfig), _P]),
    "mv2_static_diagnostic": (_I, [_P, _P]),
    "mv2_set_timing": (_I, [_P, _I]),
    "mv2_set_graphs": (_I, [_P, _I]),
    "mv2_set_speculation": (_I, [_P, _I]),
    "mv2_speculation_stats": (_I, [_P, _P]),
    "mv2_graph_replays": (ctypes.c_longlong, [_P]),
    "mv2_set_external": (_I, [_P, _P, _P, _D]),
    "mv2_batch_bind": (_I, [_P, _I, _P, _P, _P, _D, _D, _P, _P, _P, _P]),
    "mv2_batch_set_tile_shape": (_I, [_P, _I]),
    "mv2_batch_tile_shape": (_I, [_I, _P, _P]),
    "mv2_batch_step": (_I, [_P, ctypes.POINTER(StepConfig), _I, _P, _P]),
    "mv2_batch_relax": (_I, [_P, _P, ctypes.POINTER(_D), _I, _D, _P]),
    "mv2_batch_entrain": (_I, [_P, _P, _P, _P, _P, _I, _P]),
    "mv2_batch_label": (_I, [_P, _P, ctypes.POINTER(AvalancheConfig), _P, _P]),
    "mv2_entrain": (_I, [_P, _P, _P, _I, _I, ctypes.POINTER(_D)]),
    "mv2_deposit": (_I, [_P, _P, _P, _I, _I, _D, _P]),
    "mv2_reservoirs": (_I, [_P, _I, _D, _P]),
    "mv2_reservoirs_noalloc": (_I, [_P, _I, _D, _P]),
    "mv2_patch_scatter": (_I, [_P, _I, _I, _I, _I, _I, _P]),
    "mv2_index_scatter_int": (_I, [_P, _P, _P, _P, _I32, _I]),
    "mv2_index_add": (_I, [_P, _I, _P, _P, _I]),
    "mv2_region_gather": (_I, [_P, _I, _I, _I, _I, _P]),
    "mv2_failure_activate": (_I, [_P, _I, _I, _I, _I, _P, _P, _P, _P, ctypes.POINTER(ActivationResult)]),
    "mv2_surface_capture": (_I, [_P, _P]),
    "mv2_surface_dirty_tiles": (_I, [_P, _P, _D, _I, _I, _I, _P, _P, ctypes.POINTER(_I)]),
    "mv2_surface_sample": (_I, [_P, _P, _I, _P]),
    "mv2_surface_gather_tiles": (_I, [_P, _P, _I, _I, _I, _I, _P]),
    "mv2_deposition_step": (_I, [_P, ctypes.POINTER(DepositionConfig), _P, _P, _P, _P, _P, _P]),
    "mv2_relax": (_I, [_P, _I, _P, _I, _I, _P, _I, _D, ctypes.POINTER(RelaxResult)]),
    "mv2_relax_marked": (_I, [_P, _I, _P, _I, _I, _P, _I, _D, _I, ctypes.POINTER(RelaxResult), _P]),
    "mv2_frontier_relax": (_I, [_P, _P, _P, _P, _P, _P, _P, _P, _I, _I, _I, _I, _I, _I, _D, _P,
                                ctypes.POINTER(_I32)]),
    "mv2_frontier_scan": (_I, [_P, _P, _P, _P, _P, _P, _P, _P, _I, _I, _I, _D, _D, _P]),
    "mv2_frontier_rounds": (_I, [_P, ctypes.POINTER(FrontierRoundConfig), _P, _P, _P, _P,
                                 ctypes.POINTER(FrontierRoundResult)]),
    "mv2_avalanche_label": (_I, [_P, ctypes.POINTER(AvalancheConfig), ctypes.POINTER(AvalancheFields),
                                 ctypes.POINTER(LabelResult)]),
    "mv2_avalanche_previous": (_I, [_P, ctypes.POINTER(AvalancheFields), _I32, _I, _I, _P, _P, _I, _I,
                                    _I]),
    "mv2_avalanche_persistence": (_I, [_P, ctypes.POINTER(AvalancheConfig), ctypes.POINTER(AvalancheFields),
                                       _P, _P, _I, _I, _I, _P, _P, _P, ctypes.POINTER(_I32)]),
    "mv2_avalanche_mobilize": (_I, [_P, ctypes.POINTER(AvalancheConfig),
                                    ctypes.POINTER(AvalancheFields), _I32, _D, _P]),
    "mv2_label_mask": (_I, [_P, _P, _P, _I, _I, _I]),
    "mv2_nn_features": (_I, [_P, _P, _P, _I, _I, _I, _D, _D, _P]),
    "mv2_nn_conv1": (_I, [_P, _P, _P, _P, _I, _I, _I, _I, _P]),
    "mv2_nn_conv_mfma": (_I, [_P, _P, _I, _P, _P, _P, _P, _I, _I, _I, _I, _I, _I]),
    "mv2_nn_pack_weights": (_I, [_P, _P, _I, _I, _I, _I, _P, _P]),
    "mv2_nn_expand_grad": (_I, [_P, _P, ctypes.c_int64, _P]),
    "mv2_nn_weight_grad": (_I, [_P, _P, _I, _P, _I, _I, _I, _I, _I, _I, _I, _P, _P]),
    "mv2_nn_mse": (_I, [_P, _P, _P, ctypes.c_int64, _P, ctypes.POINTER(_D)]),
    "mv2_nn_adam": (_I, [_P, _P, _P, _P, _P, _I, _D, _D, _D, _D, _I]),
    "mv2_nn_project_w1": (_I, [_P, _P]),
    "mv2_nn_apply_delta": (_I, [_P, _P, _P, ctypes.c_int64, _D]),
    "mv2_nn_commit_surface": (_I, [_P, _P, _P, ctypes.c_int64, _D]),
    "mv2_envelope": (_I, [_P, _P, _P, _P, _I, _I, _P, _I, _P]),
    "mv2_batch_envelope": (_I, [_P, _P, _P, _P, _I, _I, _I, _P, _I, _P]),
    "mv2_nn_commit_settled": (_I, [_P, _P, _P, _I, _I, _P, _D, _I, _I, _P]),
    "mv2_batch_commit_settled": (_I, [_P, _P, _P, _I, _I, _I, _P, _D, _I, _I, _P]),
    "mv2_batch_commit_surface": (_I, [_P, _P, _P, _I, ctypes.c_int64, _D]),
    "mv2_nn_excess_loss": (_I, [_P, _P, _P, _I, _I, _I, _P, _D, _P, ctypes.POINTER(_D)]),
    "mv2_relax_patches": (_I, [_P, _P, _I, _I, _I, _P, _I, _D, _P]),
}

_lib: ctypes.CDLL | None = None


def load() -> ctypes.CDLL:
    """Load libmv2.so and declare its signatures (raises if it is not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise HipBackendUnavailable(
            f"libmv2.so is not built at {LIB_PATH}; run __graft_entry__.build()")
    lib = ctypes.CDLL(str(LIB_PATH))
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name, None)
        if fn is None:
            continue
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def exported_symbols() -> list[str]:
    lib = load()
    return [name for name in SIGNATURES if hasattr(lib, name)]


def last_error(ctx: int | None) -> str:
    msg = load().mv2_last_error(ctx)
    return msg.decode() if msg else ""


def raise_for(code: int, ctx: int | None, *, value_error=ValueError, runtime_error=RuntimeError) -> None:
    """Map an mv2_status to the reference's exception types."""
    if code == MV2_OK:
        return
    msg = last_error(ctx)
    if code == MV2_EINVAL:
        raise value_error(msg)
    if code == MV2_ENUMERIC:
        raise runtime_error(msg)
    if code == MV2_ESTATE:
        raise RuntimeError(msg)
    raise HipBackendUnavailable(msg)
