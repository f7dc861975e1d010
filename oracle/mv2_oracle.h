/*
 * Oracle: CPU restatement of the reference's Mobile V2 hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product links or calls this code;
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it,
 * and only as the checker / CPU baseline.
 *
 * Every function restates one reference routine op-for-op (cited per
 * function in mv2_oracle.c).  Build: oracle/Makefile (gcc -O2
 * -ffp-contract=off, no fast-math), so FP64 results are bit-reproducible.
 */
#ifndef MV2_ORACLE_H
#define MV2_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_v2_params {
    double dx, dy;
    double gravity;                 /* g */
    double earth_pressure;          /* K */
    double cfl;
    double dry_tolerance;
    double basal_friction;          /* mu */
    double tan_start;               /* tan(start angle), computed by the caller */
    double external_ax, external_ay;/* uniform external acceleration */
    int maximum_substeps;
    int limit_mode;                 /* 0: standalone (check before substep), 1: production (after) */
    int static_exit;                /* 1: apply the global static early exit */
} orc_v2_params;

typedef struct orc_v2_result {
    int status;                     /* 0 ok, 1 positivity failure, 2 maximum_substeps exceeded */
    int substeps;
    int cfl_limited;
    int static_exit;
    double maximum_cfl;
    double remaining;
    /* CPU-reference ledgers (numpy pairwise order) */
    double conservative_impulse[2];
    double friction_impulse[2];
    double external_impulse[2];
    double friction_dissipation;
    double external_work;
    /* production-kernel style per-substep sums (index order) */
    double diag[8];
} orc_v2_result;

void orc_v2_face(double bl, double br, double hl, double hr,
                 double qxl, double qyl, double qxr, double qyr,
                 int normal, double K, double g, double dry, double factor, double out[5]);

/* numpy-compatible pairwise sum of a contiguous double array */
double orc_pairwise_sum(const double *a, int64_t n, int64_t stride);

int orc_v2_step_ext(int ny, int nx, const orc_v2_params *p,
                    const double *b, double *h, double *qx, double *qy,
                    double dt, const double *ext_x, const double *ext_y, orc_v2_result *out);
int orc_v2_step(int ny, int nx, const orc_v2_params *p,
                const double *b, double *h, double *qx, double *qy,
                double dt, orc_v2_result *out);

/* max |grad(b+h)| (sqrt(gx^2+gy^2), one-sided edges) and max |q| */
void orc_v2_static(int ny, int nx, double dx, double dy,
                   const double *b, const double *h,
                   const double *qx, const double *qy, double out[2]);

/* Triangle-A-C vertex weight of cell (r,c) */
double orc_vertex_weight(int ny, int nx, double dx, double dy, int r, int c);

/* weighted summary (mass, px, py, max speed, kinetic, gravitational, internal) */
void orc_v2_summary(int ny, int nx, double dx, double dy, int weight_mode,
                    double density, double K, double g, double dry,
                    const double *b, const double *h,
                    const double *qx, const double *qy, double out[7]);

double orc_entrain(int n_cells, const double *z_base, double *b_eff, double *mobile,
                   const double *weights, const int32_t *idx, const double *req, int n);
void orc_deposit(int n_cells, double *b_eff, double *mobile, double *qx, double *qy,
                 const double *weights, const int32_t *idx, const double *req, int n,
                 double density, double out[3]);
void orc_reservoirs(int n_cells, const double *z_base, const double *b_eff,
                    const double *mobile, const double *qx, const double *qy,
                    const double *weights, double density, double out[4]);

/* Mobile -> Resting deposition step (one step_resident call) */
typedef struct orc_deposition_params {
    double dx, dy, dt;
    double settling_rate, speed_threshold;
    double density, gravity, cohesion, tan_stop;
    int settle_subcell_tail;
} orc_deposition_params;

/* FP64 storage; forcing may be NULL.  out = sum(dep w), sum(rho rx w), sum(rho ry w) */
void orc_deposition(int ny, int nx, const orc_deposition_params *p, const int32_t *forcing,
                    double *b_eff, double *mobile, double *qx, double *qy,
                    double *deposited, double *removed_x, double *removed_y,
                    double *m2r_cumulative, const double *weights, double out[3]);
/* FP32 storage: b/momentum float, mobile and deposited in int32 quanta of 2^-24 m */
void orc_deposition32(int ny, int nx, const orc_deposition_params *p, const int32_t *forcing,
                      float *b_eff, int32_t *mobile, float *qx, float *qy,
                      int32_t *deposited, float *removed_x, float *removed_y,
                      double *m2r_cumulative, const double *weights, double out[3]);

/* bilinear free-surface samples (b + m) at n fractional (row, col) pairs */
void orc_surface_sample(int ny, int nx, const double *b, const double *m, const double *rows_cols,
                        int n, double *out);

/* relax_critical_slope on an array in the legacy [axis0][axis1] layout */
int orc_relax_legacy_f64(int n0, int n1, double d0, double d1, double crit_slope,
                         int max_iterations, double tolerance, double *h,
                         int *converged, double *max_excess_last);
int orc_relax_legacy_f32(int n0, int n1, double d0, double d1, double crit_slope,
                         int max_iterations, double tolerance, float *h,
                         int *converged, double *max_excess_last);

/* compact-frontier colour batch (weighted transfer) and scan */
int orc_frontier_relax(int rows, int cols, int tile_size, int tiles_x,
                       double *height, const double *weights, int32_t *reached,
                       const int32_t *a_i, const int32_t *a_j,
                       const int32_t *b_i, const int32_t *b_j, int n_edges,
                       double critical_difference, int32_t *moved_tile_flags);
void orc_frontier_scan(int cols, const double *height, const int32_t *reached,
                       const int32_t *tile_ids,
                       const int32_t *a_i, const int32_t *a_j,
                       const int32_t *b_i, const int32_t *b_j, int n_edges,
                       double critical_difference, double tolerance,
                       int32_t *unstable_tile_flags);

typedef struct orc_avalanche_params {
    double dx, dy;
    double tan_start, tan_stop;
    double density, gravity, cohesion;
    double mobilization_depth, dry_tolerance, activity_speed;
    int release_settled_latches;
} orc_avalanche_params;

/* per-cell failure mask, slope, gradient, severity and latch release */
void orc_avalanche_initialize(int ny, int nx, const orc_avalanche_params *p,
                              const double *z_base, const double *resting,
                              const double *mobile, const double *qx, const double *qy,
                              const double *weights, const double *owned_export_baseline,
                              const double *export_cumulative,
                              int32_t *unstable, int32_t *latch,
                              double *slope, double *grad_x, double *grad_y,
                              double *severity);

/* canonical connected-component labels: label = min linear index, -1 background */
void orc_label(int ny, int nx, int connectivity, const int32_t *mask, int32_t *labels);

/* component statistics (see implementation) */
void orc_components(int ny, int nx, const int32_t *labels,
                    const double *z_base, const double *resting, const int32_t *latch,
                    const double *weights, const double *slope, const double *severity,
                    const int32_t *previous, double mobilization_depth,
                    double start_rad, double stop_rad,
                    int32_t diag_i[8], double diag_f[8]);

/* reduce_persistence over every cell: diag_i[4], diag_f[14] as the reference's
 * avalanche_persistence_int/float; tile_flags[tile_count] set to 1 where a cell
 * is unstable or carries moving mobile material */
void orc_persistence(int ny, int nx, int tile_size, int tiles_x, const double *initial_resting,
                     const double *resting, const double *mobile, const double *qx, const double *qy,
                     const double *weights, const double *slope, const int32_t *unstable,
                     const int32_t *latch, const int32_t *activation_count, const double *r2m,
                     const double *m2r, double start_rad, double stop_rad, double dry,
                     double activity_speed, int32_t diag_i[4], double diag_f[14], int32_t *tile_flags);
void orc_mobilize(int n_cells, const int32_t *labels, int32_t winner,
                  const double *z_base, double *resting, double *mobile,
                  const double *weights, int32_t *latch, double *owned_surface,
                  double *owned_export_baseline, const double *export_cumulative,
                  const double *severity, int32_t *activation_count,
                  double *first_time, double *last_time, double *r2m_cumulative,
                  double activation_time, double mobilization_depth, double dry,
                  double diag_f[16]);

#ifdef __cplusplus
}
#endif
#endif
