"""MI355X-native Mobile V2 granular-terrain hot path (HIP kernels behind a C-ABI).

Python surface mirrors the reference plugin interface (src/isaac_bulk_pipeline):
DeviceBulkState, WarpProductionMobileV2Solver, DeviceLargeAvalancheBridge,
WarpCompactActiveEdgeOperator, relax_critical_slope -- backed by libmv2.so.
"""

from ._lib import HipBackendUnavailable
from .bulk_state import (
    MOBILE_QUANTUM_M,
    BulkStateAuthority,
    BulkStateAuthorityError,
    DeviceBulkState,
    DeviceMaterialLedger,
    DeviceMaterialLedgerSnapshot,
    DeviceTransferSnapshot,
    HostBulkStatePatch,
    HostBulkStateView,
)
from .chain import GpuBulkOperatorChain, GpuBulkOperatorStep
from .config import (
    LargeAvalancheTransitionConfig,
    MaterialScenario,
    MobileV2Config,
    SolverConfig,
    TerrainGrid,
    TerrainVolumeIntegrator,
)
from .checkpoint import (load_frozen_boundary, restore_device_checkpoint, sha256_array,
                         write_device_checkpoint)
from .deposition import WarpDepositionOperator, WarpDepositionStep
from .failure_zone import DeviceFailureZoneBridge, DeviceFailureZoneResult, FailureGeometry
from .large_avalanche import DeviceLargeAvalancheBridge, DeviceLargeAvalancheResult, label_mask
from .mobile_v2 import (
    HipMobileV2ReferenceSolver,
    HipProductionMobileV2Solver,
    MobileV2State,
    MobileV2Step,
    WarpMobileStep,
    WarpMobileV2ReferenceSolver,
    WarpProductionMobileV2Solver,
)
from .relaxation import (
    CompactActiveEdgeBatch,
    CompactTileFrontier,
    FrontierRoundsResult,
    GpuResidualFrontier,
    HipCompactActiveEdgeOperator,
    HipMinimumSlopeAdapter,
    RelaxationStats,
    WarpCompactActiveEdgeOperator,
    legacy_critical_differences,
    neighbor_pairs,
    relax_critical_slope,
    relax_device,
)
from .runtime import HipRuntime

__all__ = [name for name in dir() if not name.startswith("_")]
