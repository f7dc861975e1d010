/*
 * mv2.h -- C-ABI of the MI355X Mobile V2 granular-terrain library (libmv2.so).
 *
 * Plain pointers and sizes only.  Field pointers are device pointers on the
 * context's GPU; records are host structs filled after one stream sync.
 * Every entry point returns an mv2_status; mv2_last_error() gives the message.
 *
 * Each entry point names the reference interface it replaces (paths relative
 * to the reference repository; pkg/ = src/isaac_bulk_pipeline/).
 *
 * Field layout: row-major H[y, x], ny rows by nx columns, one value per cell.
 *   MV2_FP64: z_base, b_eff, mobile, momentum_x, momentum_y are double.
 *   MV2_FP32: z_base, b_eff, momentum_x, momentum_y are float; mobile is int32
 *             in quanta of mv2_mobile_quantum() metres (exact mass transfer).
 */
#ifndef MV2_H
#define MV2_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MV2_ABI_VERSION 1

typedef enum mv2_status {
    MV2_OK = 0,
    MV2_EINVAL = 1,    /* ValueError in the reference: shape/config/binding */
    MV2_ENUMERIC = 2,  /* RuntimeError: positivity, maximum_substeps, non-convergence */
    MV2_EDEVICE = 3,   /* HIP runtime failure */
    MV2_ESTATE = 4     /* call order (e.g. step before bind) */
} mv2_status;

typedef enum mv2_precision { MV2_FP64 = 0, MV2_FP32 = 1 } mv2_precision;

typedef struct mv2_context mv2_context;

/* ---- lifecycle: replaces WarpRuntime (pkg/performance/warp_backend.py:131-190) ---- */
int mv2_abi_version(void);
double mv2_mobile_quantum(void);
int mv2_create(int device, mv2_context **out);
void mv2_destroy(mv2_context *ctx);
/* last error of ctx (or of the calling thread when ctx is NULL) */
const char *mv2_last_error(const mv2_context *ctx);
int mv2_set_stream(mv2_context *ctx, void *hip_stream);
int mv2_synchronize(mv2_context *ctx);

/* ---- state binding: WarpProductionMobileV2Solver.bind_device_state
 *      (pkg/bulk_interaction/warp_mobile_v2.py:103-115) over the DeviceBulkState
 *      named arrays (pkg/runtime/bulk_state_authority.py:543-557). Scratch is
 *      allocated here once; no per-substep allocation. ---- */
int mv2_bind(mv2_context *ctx, int precision, int ny, int nx, double dx, double dy,
             void *z_base, void *b_eff, void *mobile, void *momentum_x, void *momentum_y);

/* ---- Mobile V2 step ---- */
typedef struct mv2_step_config {
    double dt;                        /* frame dt (s) */
    double gravity;                   /* g */
    double earth_pressure;            /* K */
    double cfl;
    double dry_tolerance;
    double basal_friction;            /* mu */
    double tan_start;                 /* tan(start angle) for the static early exit */
    double external_ax, external_ay;  /* uniform external acceleration (see mv2_set_external) */
    double density;                   /* ledger density (1.0 = per-density) */
    int maximum_substeps;
    int limit_mode;                   /* 0: mobile_v2_reference.py:237, 1: warp_mobile_v2.py:215 */
    int weight_mode;                  /* 0: uniform dx*dy, 1: Triangle-A-C vertex weights */
    int static_exit;                  /* apply the global static early exit */
    int diagnostics;                  /* per-substep sums kept: 0 all ten (mobile_v2_reference.py ledgers),
                                         1 the five of the production record, 0 1 4 5 7
                                         (warp_mobile_v2.py:178-241); the others read 0 */
} mv2_step_config;

/* summary[]: 0 mass(m^3), 1 px, 2 py, 3 max speed, 4 kinetic, 5 gravitational,
 *            6 internal (each x density), 7 min h */
typedef struct mv2_step_record {
    int status;                       /* mv2_status of the step */
    int substeps;
    int cfl_limited;
    int static_exit;
    int numeric_error;                /* 1 positivity, 2 maximum_substeps, 3 overflow */
    int complete;                     /* 1: the substep loop ended (always 1 on return) */
    double maximum_cfl;
    double remaining;
    double summary_before[8];
    double summary_after[8];
    /* per-substep sums accumulated over the step:
     * 0,1 sum(q_new - q_old) before dry zeroing (mobile_v2_warp.py:198-199)
     * 2,3 external impulse  4,5 friction impulse  6 external work (x dt)
     * 7 twice the friction dissipation  8,9 sum(q_new_zeroed - q_old) (mobile_v2_reference.py:258-272) */
    double diag[10];
    double mass_quanta_before;        /* FP32: exact sum of mobile quanta (as double) */
    double mass_quanta_after;
    double kernel_ms;                 /* device time of the whole call (HIP events) */
    double substep_kernel_ms;         /* sum over the substep launches (timing mode only, else 0) */
    int launches;                     /* substep launches enqueued (substeps + ended-loop launches) */
    int timed_launches;               /* substep launches inside substep_kernel_ms (timing mode) */
} mv2_step_record;

/* One step_resident call (warp_mobile_v2.py:148-245) or standalone step
 * (pkg/experimental/mobile_v2_warp.py:251-344): static check, device-side CFL
 * loop with no host round trip per substep, ledgers before/after. */
int mv2_step(mv2_context *ctx, const mv2_step_config *cfg, mv2_step_record *out);

/* Per-cell external acceleration of the following mv2_step calls: ax / ay are
 * DEVICE arrays of ny*nx values in the state's precision (double for
 * MV2_FP64, float for MV2_FP32), used instead of cfg->external_ax/ay; NULL
 * (both) returns to the uniform values.  max_abs = max |external| over the
 * cells (the static early exit needs it <= 1e-14).  Replaces the
 * external_acceleration_m_s2 array of MobileV2ReferenceSolver.step
 * (pkg/experimental/mobile_v2_reference.py:191-210, applied :276-304).  The
 * arrays are borrowed: keep them alive while set. */
int mv2_set_external(mv2_context *ctx, const void *ax, const void *ay, double max_abs);

/* Timing mode (no counterpart in the reference; a measurement aid): when
 * enabled, mv2_step records one HIP event before its first substep launch and
 * one after the last launch expected to work (the previous call's substep
 * count); substep_kernel_ms is the time between them when the step took
 * exactly that many substeps, else 0 (events between launches would add
 * their own latency to each). */
int mv2_set_timing(mv2_context *ctx, int enable);

/* Launch graphs (no counterpart in the reference; the reference's step is a
 * sequence of Warp launches): when enabled, the first round of an mv2_step
 * call -- begin pass, predicted substeps, end / fold -- is replayed from a
 * cached HIP graph once the same launch arguments (state pointers,
 * configuration, predicted count) have been seen twice, instead of being
 * launched one kernel at a time.  Results are identical either way.  Off by
 * default: on ROCm 7.2 a replay of the 701 x 701 sequence took 73 us of
 * device time per call against 66 us for plain launches.
 * mv2_graph_replays returns how many calls were served by a replay. */
int mv2_set_graphs(mv2_context *ctx, int enable);
long long mv2_graph_replays(const mv2_context *ctx);

/* Speculative FP32 steps (no counterpart in the reference; its step begins
 * with a host-synchronised CFL reduction): the end pass of a step also
 * records the CFL maximum of its result, and the next step's first launch
 * runs the begin pass and substep 0 together with that predicted maximum;
 * the second launch keeps the substep only if the true maximum of the input
 * equals the prediction bit for bit (and the input is not static), else it
 * restarts from the untouched input.  Results are identical either way.  A
 * step speculates when the previous one left a prediction and no library
 * call has written mobile thickness or momentum since; on by default.
 * mv2_set_speculation(ctx, 0) restores a separate begin launch;
 * mv2_speculation_stats gives the steps that speculated (out[0]) and those
 * whose prediction did not hold (out[1]). */
int mv2_set_speculation(mv2_context *ctx, int enable);
int mv2_speculation_stats(const mv2_context *ctx, long long out[2]);

/* ledger reduction only (warp_mobile_v2.py:28-54 summarize) */
int mv2_summary(mv2_context *ctx, const mv2_step_config *cfg, double out[8]);

/* static diagnostic (warp_mobile_v2.py:56-85): out[0] max |grad H|, out[1] max |q| */
int mv2_static_diagnostic(mv2_context *ctx, double out[2]);

/* ---- entrainment / deposition contracts
 *      DeviceBulkState.entrain_host_indices / deposit_host_indices
 *      (pkg/runtime/bulk_state_authority.py:104-150, :920-1002).
 *      idx and depth are DEVICE arrays of length n (unique indices). ---- */
int mv2_entrain(mv2_context *ctx, const int32_t *idx, const double *depth, int n,
                int weight_mode, double *volume_out);
int mv2_deposit(mv2_context *ctx, const int32_t *idx, const double *depth, int n,
                int weight_mode, double density, double out[3]);
/* DeviceBulkState.reservoir_reduction (bulk_state_authority.py:60-75, :1004-1034):
 * out = resting volume, mobile volume, density*px, density*py */
int mv2_reservoirs(mv2_context *ctx, int weight_mode, double density, double out[4]);
/* Allocation-free variant of mv2_reservoirs.  No counterpart in the reference:
 * it is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * kernels and sums as mv2_reservoirs, bit for bit; out_device is a DEVICE
 * double[4] (same order) written asynchronously on the context's stream (no
 * sync, no host copy).  Its workgroup partials live in a file-scope device
 * array of 256 x 4 doubles (8 KiB per GPU, held until the process ends) that
 * every grid size fits; the library allocates nothing in this call.  Calls
 * that overlap on one GPU share that array and race. */
int mv2_reservoirs_noalloc(mv2_context *ctx, int weight_mode, double density, double *out_device);

/* ---- surface publication: DeviceBulkState.capture_surface_for_dirty_tracking
 *      / collect_surface_dirty_tiles (pkg/runtime/bulk_state_authority.py:733-781,
 *      kernels :156-189), sample_surface_bilinear (:1138-1181, kernel :196-228)
 *      and download_dirty_tiles (:1280-1328).  Surface = b_eff + mobile. ---- */
/* surface_before (device, ny*nx double) = the current free surface */
int mv2_surface_capture(mv2_context *ctx, double *surface_before);
/* tiles whose surface moved by more than tolerance since the capture:
 * tile_flags (device, tile_count) rewritten; tile_ids_out (HOST, tile_count
 * entries) receives the ascending ids, count_out their number */
int mv2_surface_dirty_tiles(mv2_context *ctx, const double *surface_before, double tolerance,
                            int tile_size, int tiles_x, int tile_count, int32_t *tile_flags,
                            int32_t *tile_ids_out, int *count_out);
/* bilinear free-surface samples at n fractional (row, col) pairs (device
 * double[n][2], clamped to the grid) into out (device double[n]) */
int mv2_surface_sample(mv2_context *ctx, const double *rows_cols, int n, double *out);
/* free surface of n_tiles tiles (device ids) with their one-cell overlap:
 * tile j fills out[j*(tile_size+1)^2 ...] row-major over
 * (min(r0+tile_size+1, ny) - r0) x (min(c0+tile_size+1, nx) - c0) cells */
int mv2_surface_gather_tiles(mv2_context *ctx, const int32_t *tile_ids, int n_tiles, int tile_size,
                             int tiles_x, int tile_count, double *out);

/* ---- compact state bridges: DeviceBulkState.apply_host_patch / apply_host_indices /
 *      add_host_indices / download_patch / download_region
 *      (pkg/runtime/bulk_state_authority.py:804-918, :1036-1137).
 *      field: 0 b_eff, 1 mobile, 2 momentum_x, 3 momentum_y; values in metres
 *      (FP32 mobile stored as rint(value / quantum)).  All arrays are DEVICE
 *      arrays; the calls are asynchronous on the context's stream. ---- */
/* values: row-major (row1-row0) x (col1-col0) doubles for rows [row0,row1) x cols [col0,col1) */
int mv2_patch_scatter(mv2_context *ctx, int field, int row0, int row1, int col0, int col1,
                      const double *values);
/* target[idx[k]] = values ? values[k] : scalar, for an int32 field of ny*nx cells
 * (material_mask, active_mask, deposition_mask, frontier_reached); indices
 * outside the grid are skipped */
int mv2_index_scatter_int(mv2_context *ctx, int32_t *target, const int32_t *idx, const int32_t *values,
                          int32_t scalar, int n);
/* field[idx[k]] += values[k] (atomic; indices may repeat) */
int mv2_index_add(mv2_context *ctx, int field, const int32_t *idx, const double *values, int n);
/* out (device double[5][h*w]): b_eff, z_base, mobile (m), momentum_x, momentum_y of the region */
int mv2_region_gather(mv2_context *ctx, int row0, int row1, int col0, int col1, double *out);

/* ---- FailureZone activation: DeviceFailureZoneBridge.execute after the CAD
 *      geometry (pkg/runtime/gpu_failure_bridge.py:114-136).  activated and
 *      affected are DEVICE arrays over the patch rows [row0,row1) x cols
 *      [col0,col1), row-major.  Entrains min(max(activated, 0), b_eff - z_base)
 *      per cell (Triangle-A-C ledger), then sets material_mask = 1 on every
 *      cell with mobile_before + activated > 0 that is, or is a 4-neighbour
 *      inside the patch of, an affected cell with mobile_before + activated > 0;
 *      forcing_out (device, capacity = patch cells) receives those flat indices
 *      in row-major order.  One stream sync. ---- */
typedef struct mv2_activation_result {
    double activated_volume;          /* m^3 moved resting -> mobile */
    int forcing_count;                /* entries written to forcing_out */
    int reserved;
} mv2_activation_result;
int mv2_failure_activate(mv2_context *ctx, int row0, int row1, int col0, int col1, const double *activated,
                         const int32_t *affected, int32_t *material_mask, int32_t *forcing_out,
                         mv2_activation_result *out);

/* ---- Mobile -> Resting deposition: WarpDepositionOperator.step_resident
 *      (pkg/bulk_interaction/warp_deposition.py:180-235) with its kernels
 *      deposit (:23-115) and weighted_sum (:117-129).  A cell with mobile
 *      material, forcing[i] == 0 (material_mask; forcing may be NULL), speed
 *      at or below speed_threshold and drive <= cohesion + stop resistance on
 *      the free surface b_eff + mobile moves min(h, settling_rate dt) into
 *      b_eff (all of h when settle_subcell_tail != 0) and the same fraction of
 *      its momentum out.  Every cell is judged on the pre-call fields.
 *      Per-cell outputs (DEVICE arrays of ny*nx): deposited = deposition_work
 *      (FP64 double m, FP32 int32 quanta), removed_x/y =
 *      deposition_removed_momentum_x/y (momentum precision);
 *      m2r_cumulative (double, avalanche_m2r_cumulative) += deposited * w.
 *      out = sum(deposited w), density sum(removed_x w), density sum(removed_y w).
 *      FP32 storage deposits whole quanta, min(h, floor(rate dt / quantum)). ---- */
typedef struct mv2_deposition_config {
    double dt;                        /* dt_s (> 0) */
    double settling_rate;             /* settling_rate_m_s (> 0) */
    double speed_threshold;           /* speed_threshold_m_s (>= 0) */
    double density;                   /* assumed_bulk_density_kg_m3 */
    double gravity;                   /* 9.81 in the reference */
    double cohesion;                  /* cohesion_proxy_pa */
    double tan_stop;                  /* tan(stop_angle_deg) */
    int settle_subcell_tail;
    int weight_mode;                  /* 0: uniform dx*dy, 1: Triangle-A-C vertex weights */
} mv2_deposition_config;

int mv2_deposition_step(mv2_context *ctx, const mv2_deposition_config *cfg, const int32_t *forcing,
                        void *deposited, void *removed_x, void *removed_y, double *m2r_cumulative,
                        double out[3]);

/* ---- batched Mobile V2 (no single counterpart in the reference: the
 *      reference steps one terrain per solver, warp_mobile_v2.py:148-245;
 *      this advances many independent terrains, each with its own CFL loop,
 *      in the same launches).  FP32 fields packed back to back: instance i
 *      occupies [offset[i], offset[i] + ny[i]*nx[i]) of each device array;
 *      h_quanta in quanta of mv2_mobile_quantum() metres. ---- */
typedef struct mv2_batch_record {
    int substeps;
    int numeric_error;                /* 1 positivity, 2 maximum_substeps, 3 int32 range */
    int static_exit;
    int complete;
    double maximum_cfl;
    double remaining;
    double mass_quanta_before;        /* exact sums of the mobile quanta */
    double mass_quanta_after;
    double momentum_before[2];        /* sum of q * dx * dy (per unit density) */
    double momentum_after[2];
    double kinetic_before;            /* sum of 0.5 h |v|^2 dx dy (per unit density) */
    double kinetic_after;
} mv2_batch_record;
int mv2_batch_bind(mv2_context *ctx, int count, const int32_t *ny, const int32_t *nx,
                   const int64_t *offset, double dx, double dy, float *b, int32_t *h_quanta,
                   float *qx, float *qy);
/* tile shape of the batched kernels (rows x columns): 0 8x62, 1 16x62, 2 32x62,
 * 3 64x62 (four wave bands of 2, 4, 8, 16 rows) */
int mv2_batch_set_tile_shape(mv2_context *ctx, int shape);
int mv2_batch_tile_shape(int shape, int *th, int *tw);
/* one frame of every instance (cfg->dt each); use_graph replays a HIP graph of
 * eight substep launches; out has one record per instance; timing_ms (NULL or
 * 3 doubles): device ms of the call, device ms of the substep launches, and
 * the number of substep launches */
int mv2_batch_step(mv2_context *ctx, const mv2_step_config *cfg, int use_graph,
                   mv2_batch_record *out, double *timing_ms);

/* ---- plain angle-of-repose relaxation: slope_model.relax_critical_slope
 *      (slope_model.py:273-343) behind MinimumSlopeAdapter._legacy_call
 *      (pkg/solvers/minimum_slope_adapter.py:153-182, closed boundary),
 *      on a canonical H[y, x] field of the given precision (double or float). */
typedef struct mv2_relax_result {
    int iterations;
    int converged;
    double max_excess_last;
    int complete;                     /* 1: the iteration loop ended (always 1 on return) */
    int launches;                     /* iteration launches enqueued */
} mv2_relax_result;
/* crit_diff[4] = critical slope * distance for the legacy directions
 * (1,0) dx, (0,1) dy, (1,1) and (1,-1) hypot(dx, dy), as the reference forms them. */
int mv2_relax(mv2_context *ctx, int precision, void *height, int ny, int nx,
              const double crit_diff[4], int max_iterations, double tolerance,
              mv2_relax_result *out);

/* mv2_relax that also reports the maximum excess of iteration mark_iteration
 * (>= 1), or of the last iteration when the loop ended before it.  It serves
 * the warm start's "at most N exact sweeps, then fall back to the plain
 * relaxation if the residual exceeds the threshold" (the learned avalanche
 * warm start has no counterpart in the reference; its fallback is
 * relax_critical_slope, slope_model.py:273-343): one call with mark_iteration
 * = N gives the field, iteration count and convergence of the sweeps followed
 * by the fallback, and the residual after the sweeps, without a second call. */
int mv2_relax_marked(mv2_context *ctx, int precision, void *height, int ny, int nx,
                     const double crit_diff[4], int max_iterations, double tolerance, int mark_iteration,
                     mv2_relax_result *out, double *max_excess_at_mark);

/* the plain relaxation of mv2_relax (FP32) applied to every instance of the
 * bound batch at once, each instance with its own iteration count and
 * convergence test as if relaxed alone: `field` has the batch's packed layout
 * (the bound b_eff, or any array laid out like it); out has one result per
 * instance.  Replaces a loop of relax_critical_slope calls
 * (slope_model.py:273-343), one per terrain. */
int mv2_batch_relax(mv2_context *ctx, float *field, const double crit_diff[4], int max_iterations,
                    double tolerance, mv2_relax_result *out);

/* ---- compact colour-batch sweep: WarpCompactActiveEdgeOperator.relax_phase /
 *      scan_phase (pkg/solvers/warp_compact_frontier.py:57-130, :296-392).
 *      All arrays are device arrays; flags are zeroed by the call. ---- */
int mv2_frontier_relax(mv2_context *ctx, double *height, const double *weights,
                       int32_t *reached, const int32_t *a_i, const int32_t *a_j,
                       const int32_t *b_i, const int32_t *b_j, int n_edges,
                       int rows, int cols, int tile_size, int tiles_x, int tile_count,
                       double critical_difference, int32_t *moved_tile_flags,
                       int32_t *moved_count_out);
int mv2_frontier_scan(mv2_context *ctx, const double *height, const int32_t *reached,
                      const int32_t *tile_ids, const int32_t *a_i, const int32_t *a_j,
                      const int32_t *b_i, const int32_t *b_j, int n_edges, int cols,
                      int tile_count, double critical_difference, double tolerance,
                      int32_t *unstable_tile_flags);

/* ---- residual projection rounds on the device: the loop of
 *      V2PhysicsCore._advance_gpu_residual_frontier
 *      (pkg/runtime/v2_physics_core.py:1172-1254) -- per round the relax
 *      phases of every direction and parity over the active tiles (tiles a
 *      transfer touches join for the later phases), then the scan of the
 *      round's tiles; its unstable owner tiles are the next round's active
 *      set.  Ends when the set is empty (reached is then zeroed, :1246-1249)
 *      or after max_rounds rounds.  The edge batches are those of
 *      CompactTileFrontier.edge_batch (pkg/solvers/sparse_tile_slope.py:116-230).
 *      active_tile_flags (device, tile_count int32): in, the active tiles
 *      (nonzero); out, the tiles still active.  One host sync per chunk of
 *      rounds instead of two per colour batch. ---- */
typedef struct mv2_frontier_round_config {
    int rows, cols, tile_size, tiles_x, tile_count;
    int n_directions;                 /* 1..4, e.g. slope_model._neighbor_pairs */
    int di[4], dj[4];                 /* neighbour offsets (each in -1..1) */
    double critical_difference[4];    /* critical tangent x neighbour distance */
    double tolerance;                 /* scan tolerance (m) */
    int max_rounds;                   /* residual_projection_max_rounds (>= 0) */
    int reserved;
} mv2_frontier_round_config;

typedef struct mv2_frontier_round_result {
    int rounds;                       /* rounds run */
    int converged;                    /* 1: the active set is empty */
    int active_tiles;                 /* tiles still active */
    int cleared_reached;              /* 1: a round ended empty and reached was zeroed */
    long long moved_edges;            /* transfers over all rounds */
} mv2_frontier_round_result;

int mv2_frontier_rounds(mv2_context *ctx, const mv2_frontier_round_config *cfg, double *height,
                        const double *weights, int32_t *reached, int32_t *active_tile_flags,
                        mv2_frontier_round_result *out);

/* ---- LargeAvalanche failure mask + connected components
 *      DeviceLargeAvalancheBridge.observe_and_maybe_mobilize
 *      (pkg/runtime/gpu_large_avalanche.py:541-767) device part. ---- */
typedef struct mv2_avalanche_config {
    double tan_start, tan_stop;
    double start_rad, stop_rad;
    double density, gravity, cohesion;
    double mobilization_depth, dry_tolerance, activity_speed;
    int connectivity;                 /* 4 or 8 */
    int release_settled_latches;
    int weight_mode;
    int reserved;
} mv2_avalanche_config;

/* per-cell persistent/diagnostic arrays (double unless noted), all length ny*nx */
typedef struct mv2_avalanche_fields {
    int32_t *unstable;                /* int32 */
    int32_t *labels;                  /* int32: canonical root (min index) or -1 */
    int32_t *latch;                   /* int32 */
    int32_t *previous;                /* int32 */
    int32_t *activation_count;        /* int32 */
    int32_t *counts;                  /* int32 scratch: cells per root */
    double *owned_surface;
    double *owned_export_baseline;
    double *export_cumulative;
    double *slope;
    double *grad_x;
    double *grad_y;
    double *severity;
    double *first_activation_time;
    double *last_activation_time;
    double *r2m_cumulative;
} mv2_avalanche_fields;

/* diag_i: 0 regions, 1 largest count, 2 winner root, 3 unstable cells, 4 overlap
 * diag_f: 0 sum excess_start, 1 max excess_start, 2 sum excess_stop, 3 max excess_stop,
 *         4 area, 5 mobilizable volume, 6 root excess sum */
typedef struct mv2_label_result {
    int32_t diag_i[8];
    double diag_f[8];
    int union_passes;
    int reserved;
} mv2_label_result;

int mv2_avalanche_label(mv2_context *ctx, const mv2_avalanche_config *cfg,
                        const mv2_avalanche_fields *f, mv2_label_result *out);
/* batched counterparts for the bound batch (FP32 fields, int32 quanta), every
 * instance as if handled alone:
 * mv2_batch_entrain: the entrain contract (bulk_state_authority.py:104-127) at
 *   (instance, instance-local cell) pairs; z_base NULL reads as zero; indices
 *   unique per call; volume_out[count] = m^3 moved per instance (whole quanta,
 *   uniform cell area).
 * mv2_batch_label: failure mask (gpu_large_avalanche.py:59-162 start
 *   criterion) and connected components (:165-252) per instance; labels is
 *   the packed layout (instance-local root index or -1); stats_out[4*count] =
 *   regions, largest, winner (-1 if none), unstable cells per instance. */
int mv2_batch_entrain(mv2_context *ctx, const float *z_base, const int32_t *inst, const int32_t *cell,
                      const double *depth, int n, double *volume_out);
int mv2_batch_label(mv2_context *ctx, const float *z_base, const mv2_avalanche_config *cfg, int32_t *labels,
                    int32_t *stats_out);
/* update_previous_and_tiles (gpu_large_avalanche.py:284-306) */
int mv2_avalanche_previous(mv2_context *ctx, const mv2_avalanche_fields *f, int32_t winner,
                           int keep_candidate, int seed_frontier, int32_t *frontier_reached,
                           int32_t *tile_flags, int tile_size, int tiles_x, int tile_count);
/* reduce_persistence (gpu_large_avalanche.py:378-451) and its host summary
 * (:769-839): lifecycle scalars of the resident avalanche history over
 * every cell.  diag_i: 0 cells ever activated, 1 activated more than once,
 * 2 wet cells (h > dry_tolerance), 3 retired candidates; diag_f: the
 * reference's avalanche_persistence_float slots 0-13 (5 and 9 maxima).
 * tile_flags (device, tile_count int32) are rewritten: 1 where a cell is
 * unstable or carries mobile material faster than activity_speed;
 * active_tile_count = their number.  initial_b_eff has the state's
 * precision; m2r_cumulative is avalanche_m2r_cumulative (double). */
int mv2_avalanche_persistence(mv2_context *ctx, const mv2_avalanche_config *cfg,
                              const mv2_avalanche_fields *f, const void *initial_b_eff,
                              const double *m2r_cumulative, int tile_size, int tiles_x,
                              int tile_count, int32_t *tile_flags, int32_t diag_i[4],
                              double diag_f[14], int32_t *active_tile_count);
/* mobilize (gpu_large_avalanche.py:309-376); diag_f[16] as there */
int mv2_avalanche_mobilize(mv2_context *ctx, const mv2_avalanche_config *cfg,
                           const mv2_avalanche_fields *f, int32_t winner,
                           double activation_time, double diag_f[16]);

/* generic canonical connected-component labelling of an int32 mask
 * (LargeAvalanche _largest_connected_region semantics, large_avalanche.py:711-778):
 * labels = min linear index of the component, -1 off-mask. */
int mv2_label_mask(mv2_context *ctx, const int32_t *mask, int32_t *labels, int ny, int nx,
                   int connectivity);

/* ---- learned avalanche warm start (no reference counterpart: the reference
 *      has no learned model; the plain relaxation above is its judge).
 *      Float/BF16 device buffers, NHWC activations; see csrc/mv2_surrogate.hip. ---- */
int mv2_nn_features(mv2_context *ctx, const float *b, const float *h, int B, int H, int W,
                    double dx, double dy, float *feat);
int mv2_nn_conv1(mv2_context *ctx, const float *feat, const float *w1, const float *b1,
                 int B, int H, int W, int dil, void *y);
int mv2_nn_conv_mfma(mv2_context *ctx, const void *x, int cx, const void *w, const float *bias,
                     const void *mask, void *y, int B, int H, int W, int dil, int relu, int mode);
int mv2_nn_pack_weights(mv2_context *ctx, const float *w, int cout, int cin, int cxf, int cxb,
                        void *fwd, void *bwd);
int mv2_nn_expand_grad(mv2_context *ctx, const float *g, int64_t npix, void *out16);
int mv2_nn_weight_grad(mv2_context *ctx, const void *dz, int zmode, const void *x, int xmode,
                       int B, int H, int W, int dil, int cout, int cin, float *dw, float *db);
int mv2_nn_mse(mv2_context *ctx, const float *pred, const float *target, int64_t n, float *grad,
               double *loss_out);
int mv2_nn_adam(mv2_context *ctx, float *param, const float *grad, float *m, float *v, int n,
                double lr, double beta1, double beta2, double eps, int step);
int mv2_nn_project_w1(mv2_context *ctx, float *w1);
/* slope-excess penalty of the warm-started surface pre + pred (no counterpart in
 * the reference; training aid of the learned warm start): adds its gradient to
 * grad in place; loss_out = weight / N * sum over edges of max(|dz| - crit, 0)^2 */
int mv2_nn_excess_loss(mv2_context *ctx, const float *pred, const float *pre, int B, int H, int W,
                       const double crit_diff[4], double weight, float *grad, double *loss_out);
int mv2_nn_apply_delta(mv2_context *ctx, float *b, const float *delta, int64_t n, double eps);
/* b <- projected(z): the change z - b, shifted to zero sum over the cells it
 * moves (|change| > eps), applied to b (heights kept >= 0).  No counterpart in
 * the reference (learned warm start). */
int mv2_nn_commit_surface(mv2_context *ctx, float *b, const float *z, int64_t n, double eps);
/* Lipschitz envelope (no counterpart in the reference; learned warm start):
 * z <- the largest surface below z (or below b + delta when delta is given)
 * with |z_a - z_b| <= crit_diff[d] across every edge of the four legacy
 * directions; device float32 H[y, x] fields.  launches_out[0] = launches used. */
int mv2_envelope(mv2_context *ctx, const float *b, const float *delta, float *z, int ny, int nx,
                 const double crit_diff[4], int max_launches, int *launches_out);
/* batched learned warm start for `count` fields of equal size stacked back to
 * back ([count][ny][nx]): the envelope of every field (no edge crosses from
 * one field to the next) and the mass projection per field.  No reference
 * counterpart (the learned warm start is this project's). */
int mv2_batch_envelope(mv2_context *ctx, const float *b, const float *delta, float *z, int count, int ny, int nx,
                       const double crit_diff[4], int max_launches, int *launches_out);
int mv2_batch_commit_surface(mv2_context *ctx, float *b, const float *z, int count, int64_t n, double eps);
/* Settled commit (no counterpart in the reference; learned warm start): the
 * commit of mv2_nn_commit_surface, then (rounds - 1) settle rounds, each the
 * Lipschitz envelope of the committed field followed by a projection whose
 * moved cells are those of the first commit and whose shift balances the mass
 * of the whole field; a round whose envelope moves nothing ends them.  z is
 * the enveloped surface on entry and scratch afterwards.  The batched form
 * works on `count` stacked fields of ny x nx, each on its own.
 * launches_out[0] = envelope launches of the settle rounds. */
int mv2_nn_commit_settled(mv2_context *ctx, float *b, float *z, int ny, int nx, const double crit_diff[4], double eps,
                          int rounds, int max_launches, int *launches_out);
int mv2_batch_commit_settled(mv2_context *ctx, float *b, float *z, int count, int ny, int nx,
                             const double crit_diff[4], double eps, int rounds, int max_launches, int *launches_out);
/* batched relaxation of independent FP32 patches (<= 4096 cells each), one
 * workgroup per patch, same schedule as mv2_relax; iterations per patch out */
int mv2_relax_patches(mv2_context *ctx, float *h, int npatch, int H, int W,
                      const double crit_diff[4], int max_iterations, double tolerance,
                      int *iterations);

#ifdef __cplusplus
}
#endif
#endif
