/*
 * Oracle: CPU restatement of the reference's Mobile V2 hot path.
 *
 * TEST INFRASTRUCTURE ONLY -- see mv2_oracle.h.  The product never links or
 * calls this file.  Compiled with -ffp-contract=off and no fast-math so that
 * every FP64 expression rounds exactly like the NumPy reference it restates.
 *
 * Reference files (paths relative to the reference repository root; pkg/ is
 * src/isaac_bulk_pipeline/):
 *   pkg/experimental/mobile_v2_reference.py   Mobile V2 FV step (CPU)
 *   pkg/experimental/mobile_v2_warp.py        Mobile V2 device kernels
 *   pkg/bulk_interaction/warp_mobile_v2.py    production adapter kernels
 *   pkg/runtime/bulk_state_authority.py       entrain/deposit/reservoirs/weights
 *   slope_model.py                            relax_critical_slope
 *   pkg/solvers/warp_compact_frontier.py      weighted colour-batch sweep
 *   pkg/runtime/gpu_large_avalanche.py        failure mask, union-find, mobilize
 */
#include "mv2_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------ */
/* numpy pairwise summation (numpy/_core/src/umath/loops_utils.h.src), used  */
/* so that scalar ledgers reproduce np.sum bit-for-bit on contiguous data.   */
/* ------------------------------------------------------------------------ */
#define PW_BLOCKSIZE 128
static double pairwise_rec(const double *a, int64_t n, int64_t s)
{
    if (n < 8) {
        double res = 0.0;
        for (int64_t i = 0; i < n; i++) res += a[i * s];
        return res;
    } else if (n <= PW_BLOCKSIZE) {
        double r[8];
        for (int k = 0; k < 8; k++) r[k] = a[k * s];
        int64_t i;
        for (i = 8; i < n - (n % 8); i += 8)
            for (int k = 0; k < 8; k++) r[k] += a[(i + k) * s];
        double res = ((r[0] + r[1]) + (r[2] + r[3])) + ((r[4] + r[5]) + (r[6] + r[7]));
        for (; i < n; i++) res += a[i * s];
        return res;
    } else {
        int64_t n2 = n / 2;
        n2 -= n2 % 8;
        return pairwise_rec(a, n2, s) + pairwise_rec(a + n2 * s, n - n2, s);
    }
}

double orc_pairwise_sum(const double *a, int64_t n, int64_t stride)
{
    return pairwise_rec(a, n, stride);
}

/* np.sum(q, axis=(0, 1)) over an (ny, nx, 2) array accumulates each kept
 * component sequentially in row-major order (probed against numpy 2.2). */
static double seq_sum(const double *a, int64_t n)
{
    double acc = 0.0;
    for (int64_t i = 0; i < n; i++) acc += a[i];
    return acc;
}

static inline double dmax(double a, double b) { return a >= b ? a : (b > a ? b : a); }
static inline double dmin(double a, double b) { return a <= b ? a : (b < a ? b : a); }
/* np.maximum / np.minimum for non-NaN operands */
static inline double npmax(double a, double b) { return a >= b ? a : b; }
static inline double npmin(double a, double b) { return a <= b ? a : b; }

/* ------------------------------------------------------------------------ */
/* Mobile V2 step, CPU reference semantics                                   */
/* mobile_v2_reference.py:191-314 (loop), :365-431 (_face_updates),          */
/* :433-449 (_physical_flux), :353-363 (static early exit).                 */
/* ------------------------------------------------------------------------ */
typedef struct face_out { double h, qxl, qyl, qxr, qyr; } face_out;

/* One shared face: hydrostatic reconstruction, Rusanov flux and the
 * two-state pressure correction (mobile_v2_reference.py:391-431). The
 * returned values are already multiplied by factor = dt/spacing. */
static face_out v2_face(double bl, double br, double hl, double hr,
                        double qxl, double qyl, double qxr, double qyr,
                        int normal, double K, double g, double dry, double factor)
{
    const double Kg = K * g;
    const double c = 0.5 * K * g;          /* (0.5*K)*g as Python evaluates it */
    double head_l = bl + K * hl;
    double head_r = br + K * hr;
    double b_face = npmax(bl, br);
    double hs_l = npmax((head_l - b_face) / K, 0.0);
    double hs_r = npmax((head_r - b_face) / K, 0.0);
    double scale_l = hl > dry ? hs_l / hl : 0.0;
    double scale_r = hr > dry ? hs_r / hr : 0.0;
    double qsx_l = qxl * scale_l, qsy_l = qyl * scale_l;
    double qsx_r = qxr * scale_r, qsy_r = qyr * scale_r;
    /* _physical_flux */
    double vx_l = hs_l > dry ? qsx_l / hs_l : 0.0;
    double vy_l = hs_l > dry ? qsy_l / hs_l : 0.0;
    double vx_r = hs_r > dry ? qsx_r / hs_r : 0.0;
    double vy_r = hs_r > dry ? qsy_r / hs_r : 0.0;
    double qn_l = normal == 0 ? qsx_l : qsy_l;
    double qn_r = normal == 0 ? qsx_r : qsy_r;
    double fl[3] = {qn_l, qn_l * vx_l, qn_l * vy_l};
    double fr[3] = {qn_r, qn_r * vx_r, qn_r * vy_r};
    fl[1 + normal] = fl[1 + normal] + (c * hs_l) * hs_l;
    fr[1 + normal] = fr[1 + normal] + (c * hs_r) * hs_r;
    double ul = hs_l > dry ? qn_l / hs_l : 0.0;
    double ur = hs_r > dry ? qn_r / hs_r : 0.0;
    double speed = npmax(fabs(ul) + sqrt(Kg * hs_l), fabs(ur) + sqrt(Kg * hs_r));
    double Ul[3] = {hs_l, qsx_l, qsy_l};
    double Ur[3] = {hs_r, qsx_r, qsy_r};
    double flux[3];
    for (int k = 0; k < 3; k++)
        flux[k] = 0.5 * (fl[k] + fr[k]) - (0.5 * speed) * (Ur[k] - Ul[k]);
    double p_l = (c * hl) * hl, p_r = (c * hr) * hr;
    double ps_l = (c * hs_l) * hs_l, ps_r = (c * hs_r) * hs_r;
    double fm[3] = {flux[0], flux[1], flux[2]};
    double fp[3] = {flux[0], flux[1], flux[2]};
    fm[1 + normal] = fm[1 + normal] + (p_l - ps_l);
    fp[1 + normal] = fp[1 + normal] + (p_r - ps_r);
    face_out o;
    o.h = factor * flux[0];
    o.qxl = factor * fm[1];
    o.qyl = factor * fm[2];
    o.qxr = factor * fp[1];
    o.qyr = factor * fp[2];
    return o;
}

static int static_below_start(int ny, int nx, double dx, double dy,
                              const double *b, const double *h,
                              const double *qx, const double *qy, double tan_start)
{
    int64_t n = (int64_t)ny * nx;
    double qmax = 0.0;
    for (int64_t i = 0; i < n; i++) {
        qmax = npmax(qmax, fabs(qx[i]));
        qmax = npmax(qmax, fabs(qy[i]));
    }
    if (qmax > 1.0e-14) return 0;
    double gmax = 0.0;
    for (int r = 0; r < ny; r++) {
        for (int c = 0; c < nx; c++) {
            /* np.gradient(free, dy, dx), edge_order=1 */
            double gx, gy;
#define FREE(rr, cc) (b[(int64_t)(rr) * nx + (cc)] + h[(int64_t)(rr) * nx + (cc)])
            if (c == 0) gx = (FREE(r, 1) - FREE(r, 0)) / dx;
            else if (c == nx - 1) gx = (FREE(r, nx - 1) - FREE(r, nx - 2)) / dx;
            else gx = (FREE(r, c + 1) - FREE(r, c - 1)) / (2.0 * dx);
            if (r == 0) gy = (FREE(1, c) - FREE(0, c)) / dy;
            else if (r == ny - 1) gy = (FREE(ny - 1, c) - FREE(ny - 2, c)) / dy;
            else gy = (FREE(r + 1, c) - FREE(r - 1, c)) / (2.0 * dy);
#undef FREE
            gmax = npmax(gmax, hypot(gx, gy));
        }
    }
    return gmax <= tan_start + 1.0e-14;
}

int orc_v2_step(int ny, int nx, const orc_v2_params *p,
                const double *b, double *h, double *qx, double *qy,
                double dt, orc_v2_result *out)
{
    return orc_v2_step_ext(ny, nx, p, b, h, qx, qy, dt, NULL, NULL, out);
}

/* The step with a per-cell external acceleration (ext_x/ext_y of ny*nx, or
 * NULL for the uniform p->external_ax/ay): the reference broadcasts
 * external_acceleration_m_s2 to q.shape (mobile_v2_reference.py:204-210) and
 * skips the static exit unless max |external| <= 1e-14 (:226-229). */
int orc_v2_step_ext(int ny, int nx, const orc_v2_params *p,
                    const double *b, double *h, double *qx, double *qy,
                    double dt, const double *ext_x, const double *ext_y, orc_v2_result *out)
{
    memset(out, 0, sizeof(*out));
    const int64_t n = (int64_t)ny * nx;
    const double dx = p->dx, dy = p->dy;
    const double K = p->earth_pressure, g = p->gravity, dry = p->dry_tolerance;
#define AX(i) (ext_x ? ext_x[i] : p->external_ax)
#define AY(i) (ext_y ? ext_y[i] : p->external_ay)
    double ext_max = 0.0;
    if (ext_x || ext_y) {
        for (int64_t i = 0; i < n; i++) ext_max = npmax(ext_max, npmax(fabs(AX(i)), fabs(AY(i))));
    } else {
        ext_max = npmax(0.0, npmax(fabs(p->external_ax), fabs(p->external_ay)));
    }
    if (p->static_exit && ext_max <= 1.0e-14 &&
        static_below_start(ny, nx, dx, dy, b, h, qx, qy, p->tan_start)) {
        out->static_exit = 1;
        out->remaining = dt;
        return 0;
    }
    const int64_t nfx = (int64_t)ny * (nx - 1), nfy = (int64_t)(ny - 1) * nx;
    face_out *fx = (face_out *)malloc(sizeof(face_out) * (nfx > 0 ? nfx : 1));
    face_out *fy = (face_out *)malloc(sizeof(face_out) * (nfy > 0 ? nfy : 1));
    double *hn = (double *)malloc(sizeof(double) * n);
    double *qxn = (double *)malloc(sizeof(double) * n);
    double *qyn = (double *)malloc(sizeof(double) * n);
    double *ext_work = (double *)malloc(sizeof(double) * n);
    double *eix = (double *)malloc(sizeof(double) * n);
    double *eiy = (double *)malloc(sizeof(double) * n);
    double *fix = (double *)malloc(sizeof(double) * n);
    double *fiy = (double *)malloc(sizeof(double) * n);
    double *fdis = (double *)malloc(sizeof(double) * n);
#ifdef _OPENMP
    double *diag_terms = (double *)malloc(sizeof(double) * 8 * n);
#endif
    const double mind = dx <= dy ? dx : dy;
    const double Kg = K * g;
    const double mug = p->basal_friction * g;
    const double area = dx * dy;
    double remaining = dt;
    int substeps = 0;
    double max_cfl = 0.0;
    int status = 0;
    while (remaining > 1.0e-14) {
        if (p->limit_mode == 0 && substeps >= p->maximum_substeps) { status = 2; break; }
        double maxw = 0.0;
        /* OMP: the maximum does not depend on the order of the cells */
#pragma omp parallel for reduction(max : maxw)
        for (int64_t i = 0; i < n; i++) {
            double vx = h[i] > dry ? qx[i] / h[i] : 0.0;
            double vy = h[i] > dry ? qy[i] / h[i] : 0.0;
            double wave = sqrt(vx * vx + vy * vy) + sqrt(Kg * npmax(h[i], 0.0));
            maxw = npmax(maxw, wave);
        }
        double stable = maxw <= 1.0e-14 ? remaining : p->cfl * mind / maxw;
        double sub_dt = npmin(remaining, stable);
        max_cfl = npmax(max_cfl, maxw * sub_dt / mind);
        if (sub_dt < remaining - 1.0e-14) out->cfl_limited = 1;
        double before_x = seq_sum(qx, n);
        double before_y = seq_sum(qy, n);
        /* faces: x then y (mobile_v2_reference.py:261-262) */
        const double fxf = sub_dt / dx, fyf = sub_dt / dy;
#pragma omp parallel for
        for (int r = 0; r < ny; r++)
            for (int c = 0; c + 1 < nx; c++) {
                int64_t L = (int64_t)r * nx + c, R = L + 1;
                fx[(int64_t)r * (nx - 1) + c] = v2_face(b[L], b[R], h[L], h[R], qx[L], qy[L],
                                                       qx[R], qy[R], 0, K, g, dry, fxf);
            }
#pragma omp parallel for
        for (int r = 0; r < ny - 1; r++)
            for (int c = 0; c < nx; c++) {
                int64_t L = (int64_t)r * nx + c, R = L + nx;
                fy[(int64_t)r * nx + c] = v2_face(b[L], b[R], h[L], h[R], qx[L], qy[L],
                                                 qx[R], qy[R], 1, K, g, dry, fyf);
            }
        double hmin = 0.0;
#pragma omp parallel for reduction(min : hmin)
        for (int r = 0; r < ny; r++)
            for (int c = 0; c < nx; c++) {
                int64_t i = (int64_t)r * nx + c;
                /* dh/dq accumulate in NumPy's order: x minus, x plus, y minus, y plus */
                double dh = 0.0, dqx = 0.0, dqy = 0.0;
                if (c + 1 < nx) {
                    face_out f = fx[(int64_t)r * (nx - 1) + c];
                    dh = dh - f.h; dqx = dqx - f.qxl; dqy = dqy - f.qyl;
                }
                if (c > 0) {
                    face_out f = fx[(int64_t)r * (nx - 1) + c - 1];
                    dh = dh + f.h; dqx = dqx + f.qxr; dqy = dqy + f.qyr;
                }
                if (r + 1 < ny) {
                    face_out f = fy[(int64_t)r * nx + c];
                    dh = dh - f.h; dqx = dqx - f.qxl; dqy = dqy - f.qyl;
                }
                if (r > 0) {
                    face_out f = fy[(int64_t)(r - 1) * nx + c];
                    dh = dh + f.h; dqx = dqx + f.qxr; dqy = dqy + f.qyr;
                }
                hn[i] = h[i] + dh;
                qxn[i] = qx[i] + dqx;
                qyn[i] = qy[i] + dqy;
                hmin = npmin(hmin, hn[i]);
            }
        if (hmin < -2.0e-13) { status = 1; break; }
#ifdef _OPENMP
        /* production-kernel style sums (mobile_v2_warp.py:193-231): the terms
         * of every cell first (independent), then the sums in cell order */
#pragma omp parallel for
        for (int64_t i = 0; i < n; i++) {
            double *t = diag_terms + (size_t)8 * i;
            double depth = npmax(hn[i], 0.0);
            t[0] = qxn[i] - qx[i];
            t[1] = qyn[i] - qy[i];
            for (int k = 2; k < 8; k++) t[k] = 0.0;
            if (depth <= dry) continue;
            double vx0 = qxn[i] / depth, vy0 = qyn[i] / depth;
            double vx1 = vx0 + AX(i) * sub_dt, vy1 = vy0 + AY(i) * sub_dt;
            t[2] = depth * (vx1 - vx0);
            t[3] = depth * (vy1 - vy0);
            t[6] = 0.5 * depth * (vx1 * vx1 + vy1 * vy1 - vx0 * vx0 - vy0 * vy0);
            double speed = sqrt(vx1 * vx1 + vy1 * vy1);
            double factor = 0.0;
            if (speed > 1.0e-15) factor = npmax(0.0, 1.0 - mug * sub_dt / speed);
            double vx2 = vx1 * factor, vy2 = vy1 * factor;
            t[4] = depth * (vx2 - vx1);
            t[5] = depth * (vy2 - vy1);
            t[7] = 0.5 * depth * (vx1 * vx1 + vy1 * vy1 - vx2 * vx2 - vy2 * vy2);
        }
        /* adding the +0 terms of dry cells leaves every sum unchanged */
        for (int64_t i = 0; i < n; i++)
            for (int k = 0; k < 8; k++) out->diag[k] += diag_terms[(size_t)8 * i + k];
#else
        /* production-kernel style sums (mobile_v2_warp.py:193-231) */
        for (int64_t i = 0; i < n; i++) {
            double depth = npmax(hn[i], 0.0);
            out->diag[0] += qxn[i] - qx[i];
            out->diag[1] += qyn[i] - qy[i];
            if (depth <= dry) continue;
            double vx0 = qxn[i] / depth, vy0 = qyn[i] / depth;
            double vx1 = vx0 + AX(i) * sub_dt, vy1 = vy0 + AY(i) * sub_dt;
            out->diag[2] += depth * (vx1 - vx0);
            out->diag[3] += depth * (vy1 - vy0);
            out->diag[6] += 0.5 * depth * (vx1 * vx1 + vy1 * vy1 - vx0 * vx0 - vy0 * vy0);
            double speed = sqrt(vx1 * vx1 + vy1 * vy1);
            double factor = 0.0;
            if (speed > 1.0e-15) factor = npmax(0.0, 1.0 - mug * sub_dt / speed);
            double vx2 = vx1 * factor, vy2 = vy1 * factor;
            out->diag[4] += depth * (vx2 - vx1);
            out->diag[5] += depth * (vy2 - vy1);
            out->diag[7] += 0.5 * depth * (vx1 * vx1 + vy1 * vy1 - vx2 * vx2 - vy2 * vy2);
        }
#endif
#pragma omp parallel for
        for (int64_t i = 0; i < n; i++) {
            hn[i] = npmax(hn[i], 0.0);
            if (hn[i] <= dry) { qxn[i] = 0.0; qyn[i] = 0.0; }
        }
        double after_x = seq_sum(qxn, n);
        double after_y = seq_sum(qyn, n);
        out->conservative_impulse[0] += (after_x - before_x) * dx * dy;
        out->conservative_impulse[1] += (after_y - before_y) * dx * dy;
        /* external source (midpoint work) and Coulomb friction (:274-304) */
#pragma omp parallel for
        for (int64_t i = 0; i < n; i++) {
            int wet = hn[i] > dry;
            double vbx = wet ? qxn[i] / hn[i] : 0.0, vby = wet ? qyn[i] / hn[i] : 0.0;
            const double ax = AX(i), ay = AY(i);
            double qex = qxn[i] + hn[i] * ax * sub_dt;
            double qey = qyn[i] + hn[i] * ay * sub_dt;
            double vex = wet ? qex / hn[i] : 0.0, vey = wet ? qey / hn[i] : 0.0;
            ext_work[i] = hn[i] * ((0.5 * (vbx + vex)) * ax + (0.5 * (vby + vey)) * ay);
            eix[i] = qex - qxn[i];
            eiy[i] = qey - qyn[i];
            double speed = sqrt(vex * vex + vey * vey);
            double decrement = mug * sub_dt;
            double factor = npmax(0.0, 1.0 - decrement / npmax(speed, 1.0e-15));
            double vax = vex * factor, vay = vey * factor;
            double qax = hn[i] * vax, qay = hn[i] * vay;
            fix[i] = qax - qex;
            fiy[i] = qay - qey;
            fdis[i] = hn[i] * (speed * speed - (vax * vax + vay * vay));
            qx[i] = qax;
            qy[i] = qay;
        }
        out->external_impulse[0] += seq_sum(eix, n) * area;
        out->external_impulse[1] += seq_sum(eiy, n) * area;
        out->external_work += orc_pairwise_sum(ext_work, n, 1) * area * sub_dt;
        out->friction_impulse[0] += seq_sum(fix, n) * area;
        out->friction_impulse[1] += seq_sum(fiy, n) * area;
        out->friction_dissipation += 0.5 * orc_pairwise_sum(fdis, n, 1) * area;
        memcpy(h, hn, sizeof(double) * n);
        remaining -= sub_dt;
        substeps += 1;
        if (p->limit_mode == 1 && substeps > p->maximum_substeps) { status = 2; break; }
    }
    free(fx); free(fy); free(hn); free(qxn); free(qyn);
    free(ext_work); free(eix); free(eiy); free(fix); free(fiy); free(fdis);
#ifdef _OPENMP
    free(diag_terms);
#endif
    out->status = status;
    out->substeps = substeps;
    out->maximum_cfl = max_cfl;
    out->remaining = remaining;
    return status;
#undef AX
#undef AY
}

void orc_v2_static(int ny, int nx, double dx, double dy,
                   const double *b, const double *h,
                   const double *qx, const double *qy, double out[2])
{
    int64_t n = (int64_t)ny * nx;
    double qmax = 0.0, gmax = 0.0;
    for (int64_t i = 0; i < n; i++) {
        qmax = npmax(qmax, fabs(qx[i]));
        qmax = npmax(qmax, fabs(qy[i]));
    }
    for (int r = 0; r < ny; r++)
        for (int c = 0; c < nx; c++) {
            int cl = c > 0 ? c - 1 : 0, cr = c + 1 < nx ? c + 1 : nx - 1;
            int rd = r > 0 ? r - 1 : 0, ru = r + 1 < ny ? r + 1 : ny - 1;
            double denx = (c > 0 && c + 1 < nx) ? 2.0 * dx : dx;
            double deny = (r > 0 && r + 1 < ny) ? 2.0 * dy : dy;
            double fr = b[(int64_t)r * nx + cr] + h[(int64_t)r * nx + cr];
            double fl = b[(int64_t)r * nx + cl] + h[(int64_t)r * nx + cl];
            double fu = b[(int64_t)ru * nx + c] + h[(int64_t)ru * nx + c];
            double fd = b[(int64_t)rd * nx + c] + h[(int64_t)rd * nx + c];
            double gx = (fr - fl) / denx, gy = (fu - fd) / deny;
            gmax = npmax(gmax, sqrt(gx * gx + gy * gy));
        }
    out[0] = gmax;
    out[1] = qmax;
}

/* bulk_state_authority.py:624-634, evaluated per cell in the same order */
double orc_vertex_weight(int ny, int nx, double dx, double dy, int r, int c)
{
    double f = dx * dy / 6.0;
    double w = 0.0;
    if (r < ny - 1 && c < nx - 1) w += 2.0 * f;
    if (r < ny - 1 && c > 0) w += f;
    if (r > 0 && c > 0) w += 2.0 * f;
    if (r > 0 && c < nx - 1) w += f;
    return w;
}

/* warp_mobile_v2.py:28-54 (summarize), index-order sums */
void orc_v2_summary(int ny, int nx, double dx, double dy, int weight_mode,
                    double density, double K, double g, double dry,
                    const double *b, const double *h,
                    const double *qx, const double *qy, double out[7])
{
    for (int k = 0; k < 7; k++) out[k] = 0.0;
    for (int r = 0; r < ny; r++)
        for (int c = 0; c < nx; c++) {
            int64_t i = (int64_t)r * nx + c;
            double w = weight_mode ? orc_vertex_weight(ny, nx, dx, dy, r, c) : dx * dy;
            double depth = h[i];
            out[0] += depth * w;
            out[1] += density * qx[i] * w;
            out[2] += density * qy[i] * w;
            if (depth > dry) {
                double vx = qx[i] / depth, vy = qy[i] / depth;
                double speed = sqrt(vx * vx + vy * vy);
                out[3] = npmax(out[3], speed);
                out[4] += 0.5 * density * depth * speed * speed * w;
            }
            out[5] += density * g * depth * b[i] * w;
            out[6] += density * 0.5 * K * g * depth * depth * w;
        }
}

/* bulk_state_authority.py:104-119 (entrain), sequential item order */
double orc_entrain(int n_cells, const double *z_base, double *b_eff, double *mobile,
                   const double *weights, const int32_t *idx, const double *req, int n)
{
    (void)n_cells;
    double total = 0.0;
    for (int k = 0; k < n; k++) {
        int i = idx[k];
        double available = npmax(b_eff[i] - z_base[i], 0.0);
        double depth = npmin(npmax(req[k], 0.0), available);
        b_eff[i] = b_eff[i] - depth;
        mobile[i] = mobile[i] + depth;
        total += depth * weights[i];
    }
    return total;
}

/* bulk_state_authority.py:121-149 (deposit) */
void orc_deposit(int n_cells, double *b_eff, double *mobile, double *qx, double *qy,
                 const double *weights, const int32_t *idx, const double *req, int n,
                 double density, double out[3])
{
    (void)n_cells;
    out[0] = out[1] = out[2] = 0.0;
    for (int k = 0; k < n; k++) {
        int i = idx[k];
        double before = npmax(mobile[i], 0.0);
        double depth = npmin(npmax(req[k], 0.0), before);
        double fraction = before > 0.0 ? depth / before : 0.0;
        double rx = qx[i] * fraction, ry = qy[i] * fraction;
        b_eff[i] = b_eff[i] + depth;
        mobile[i] = before - depth;
        qx[i] = qx[i] - rx;
        qy[i] = qy[i] - ry;
        double w = weights[i];
        out[0] += depth * w;
        out[1] += density * rx * w;
        out[2] += density * ry * w;
    }
}

/* bulk_state_authority.py:60-75 (sum_reservoirs) */
void orc_reservoirs(int n_cells, const double *z_base, const double *b_eff,
                    const double *mobile, const double *qx, const double *qy,
                    const double *weights, double density, double out[4])
{
    out[0] = out[1] = out[2] = out[3] = 0.0;
    for (int i = 0; i < n_cells; i++) {
        double w = weights[i];
        out[0] += npmax(b_eff[i] - z_base[i], 0.0) * w;
        out[1] += mobile[i] * w;
        out[2] += density * qx[i] * w;
        out[3] += density * qy[i] * w;
    }
}

/* bilinear_surface (bulk_state_authority.py:196-228) */
void orc_surface_sample(int ny, int nx, const double *b, const double *m, const double *rows_cols,
                        int n, double *out)
{
    for (int k = 0; k < n; k++) {
        double row = npmin(npmax(rows_cols[2 * k], 0.0), (double)(ny - 1));
        double col = npmin(npmax(rows_cols[2 * k + 1], 0.0), (double)(nx - 1));
        int r0 = (int)floor(row), c0 = (int)floor(col);
        int r1 = r0 + 1 < ny - 1 ? r0 + 1 : ny - 1;
        int c1 = c0 + 1 < nx - 1 ? c0 + 1 : nx - 1;
        double fr = row - (double)r0, fc = col - (double)c0;
        double v00 = b[(size_t)r0 * nx + c0] + m[(size_t)r0 * nx + c0];
        double v01 = b[(size_t)r0 * nx + c1] + m[(size_t)r0 * nx + c1];
        double v10 = b[(size_t)r1 * nx + c0] + m[(size_t)r1 * nx + c0];
        double v11 = b[(size_t)r1 * nx + c1] + m[(size_t)r1 * nx + c1];
        out[k] = (1.0 - fr) * (1.0 - fc) * v00 + (1.0 - fr) * fc * v01 + fr * (1.0 - fc) * v10
                 + fr * fc * v11;
    }
}

/* ------------------------------------------------------------------------ */
/* Mobile -> Resting deposition: warp_deposition.py:23-115 (deposit) and     */
/* :117-129 (weighted_sum), one step_resident call (:180-235).  Every cell   */
/* is judged on the field as it was before the call (the CPU operator        */
/* deposition.py:102-153 also works on a copy); the neighbour free surface   */
/* comes from copies of b_eff and mobile taken before any cell is changed.   */
/* ------------------------------------------------------------------------ */

/* warp_deposition.py:69-102: 1 when the slow cell is within the cohesive
 * stop yield of its free surface; bl/ml are the pre-call b_eff and mobile */
static int deposition_settles(int ny, int nx, int r, int c, const double *bl, const double *ml,
                              double h, double qx, double qy, const orc_deposition_params *p)
{
    double vx = qx / h, vy = qy / h;
    if (sqrt(vx * vx + vy * vy) > p->speed_threshold) return 0;
    int i = r * nx + c;
    int left = c > 0 ? i - 1 : i, right = c + 1 < nx ? i + 1 : i;
    int down = r > 0 ? i - nx : i, up = r + 1 < ny ? i + nx : i;
    double hx = (c == 0 || c + 1 == nx) ? p->dx : 2.0 * p->dx;
    double hy = (r == 0 || r + 1 == ny) ? p->dy : 2.0 * p->dy;
    double sx = (bl[right] + ml[right] - bl[left] - ml[left]) / hx;
    double sy = (bl[up] + ml[up] - bl[down] - ml[down]) / hy;
    double slope_tan = sqrt(sx * sx + sy * sy);
    double normalization = sqrt(1.0 + slope_tan * slope_tan);
    double drive = p->density * p->gravity * h * slope_tan / normalization;
    double resistance = p->cohesion + p->density * p->gravity * h * p->tan_stop / normalization;
    return !(drive > resistance);
}

void orc_deposition(int ny, int nx, const orc_deposition_params *p, const int32_t *forcing,
                    double *b_eff, double *mobile, double *qx, double *qy,
                    double *deposited, double *removed_x, double *removed_y,
                    double *m2r_cumulative, const double *weights, double out[3])
{
    const size_t n = (size_t)ny * nx;
    double *b0 = malloc(n * sizeof(double)), *m0 = malloc(n * sizeof(double));
    memcpy(b0, b_eff, n * sizeof(double));
    memcpy(m0, mobile, n * sizeof(double));
    for (int r = 0; r < ny; r++)
        for (int c = 0; c < nx; c++) {
            size_t i = (size_t)r * nx + c;
            double h = m0[i];
            deposited[i] = 0.0;
            removed_x[i] = 0.0;
            removed_y[i] = 0.0;
            if (h <= 0.0 || (forcing && forcing[i] != 0)) continue;
            if (p->settle_subcell_tail) {
                removed_x[i] = qx[i];
                removed_y[i] = qy[i];
                b_eff[i] = b_eff[i] + h;
                mobile[i] = 0.0;
                qx[i] = 0.0;
                qy[i] = 0.0;
                deposited[i] = h;
                m2r_cumulative[i] = m2r_cumulative[i] + h * weights[i];
                continue;
            }
            if (!deposition_settles(ny, nx, r, c, b0, m0, h, qx[i], qy[i], p)) continue;
            double amount = npmin(h, p->settling_rate * p->dt);
            double fraction_removed = amount / h;
            removed_x[i] = qx[i] * fraction_removed;
            removed_y[i] = qy[i] * fraction_removed;
            b_eff[i] = b_eff[i] + amount;
            mobile[i] = h - amount;
            double fraction = (h - amount) / h;
            qx[i] = qx[i] * fraction;
            qy[i] = qy[i] * fraction;
            deposited[i] = amount;
            m2r_cumulative[i] = m2r_cumulative[i] + amount * weights[i];
        }
    out[0] = out[1] = out[2] = 0.0;
    for (size_t i = 0; i < n; i++) {
        out[0] += deposited[i] * weights[i];
        out[1] += p->density * removed_x[i] * weights[i];
        out[2] += p->density * removed_y[i] * weights[i];
    }
    free(b0);
    free(m0);
}

/* The same step on FP32 storage.  The decision is evaluated in double on the
 * widened values; the amount is whole quanta, min(h, floor(rate dt / q)),
 * so b_eff + mobile moves exactly what the ledger reports. */
void orc_deposition32(int ny, int nx, const orc_deposition_params *p, const int32_t *forcing,
                      float *b_eff, int32_t *mobile, float *qx, float *qy,
                      int32_t *deposited, float *removed_x, float *removed_y,
                      double *m2r_cumulative, const double *weights, double out[3])
{
    const double Q = 1.0 / 16777216.0;
    const size_t n = (size_t)ny * nx;
    double *b0 = malloc(n * sizeof(double)), *m0 = malloc(n * sizeof(double));
    for (size_t i = 0; i < n; i++) {
        b0[i] = (double)b_eff[i];
        m0[i] = (double)mobile[i] * Q;
    }
    double cap_d = floor(p->settling_rate * p->dt * 16777216.0);
    int32_t cap = cap_d >= 2147483647.0 ? 2147483647 : (int32_t)cap_d;
    for (int r = 0; r < ny; r++)
        for (int c = 0; c < nx; c++) {
            size_t i = (size_t)r * nx + c;
            int32_t hq = mobile[i];
            deposited[i] = 0;
            removed_x[i] = 0.0f;
            removed_y[i] = 0.0f;
            if (hq <= 0 || (forcing && forcing[i] != 0)) continue;
            double h = m0[i];
            if (p->settle_subcell_tail) {
                removed_x[i] = qx[i];
                removed_y[i] = qy[i];
                b_eff[i] = (float)((double)b_eff[i] + h);
                mobile[i] = 0;
                qx[i] = 0.0f;
                qy[i] = 0.0f;
                deposited[i] = hq;
                m2r_cumulative[i] = m2r_cumulative[i] + h * weights[i];
                continue;
            }
            if (!deposition_settles(ny, nx, r, c, b0, m0, h, (double)qx[i], (double)qy[i], p)) continue;
            int32_t aq = hq < cap ? hq : cap;
            if (aq == 0) continue;
            double amount = (double)aq * Q;
            double fraction_removed = (double)aq / (double)hq;
            double fraction = (double)(hq - aq) / (double)hq;
            removed_x[i] = (float)((double)qx[i] * fraction_removed);
            removed_y[i] = (float)((double)qy[i] * fraction_removed);
            b_eff[i] = (float)((double)b_eff[i] + amount);
            mobile[i] = hq - aq;
            qx[i] = (float)((double)qx[i] * fraction);
            qy[i] = (float)((double)qy[i] * fraction);
            deposited[i] = aq;
            m2r_cumulative[i] = m2r_cumulative[i] + amount * weights[i];
        }
    out[0] = out[1] = out[2] = 0.0;
    for (size_t i = 0; i < n; i++) {
        out[0] += (double)deposited[i] * Q * weights[i];
        out[1] += p->density * (double)removed_x[i] * weights[i];
        out[2] += p->density * (double)removed_y[i] * weights[i];
    }
    free(b0);
    free(m0);
}

/* ------------------------------------------------------------------------ */
/* relax_critical_slope (slope_model.py:273-343) on a legacy-layout array    */
/* A[n0][n1]; directions (1,0),(0,1),(1,1),(1,-1) of _neighbor_pairs         */
/* (:744-750) and the views of _paired_views (:753-764).  The caller passes  */
/* crit_slope; the per-direction critical differences are formed here as     */
/* crit_slope * distance exactly like the reference (one product per edge    */
/* set, distance hypot(d0, d1) for the diagonals).                           */
/* ------------------------------------------------------------------------ */
#define RELAX_BODY(REAL)                                                          \
    const int dirs[4][2] = {{1, 0}, {0, 1}, {1, 1}, {1, -1}};                     \
    double dist[4];                                                               \
    dist[0] = d0; dist[1] = d1; dist[2] = hypot(d0, d1); dist[3] = dist[2];        \
    int iteration = 0;                                                            \
    *converged = 0;                                                               \
    double last = 0.0;                                                            \
    for (iteration = 1; iteration <= max_iterations; iteration++) {              \
        double max_excess = 0.0;                                                  \
        for (int d = 0; d < 4; d++) {                                             \
            int di = dirs[d][0], dj = dirs[d][1];                                 \
            REAL crit = (REAL)(crit_slope * dist[d]);                             \
            /* a = h[ia0.., ja0..], b = a shifted by (di,dj) over the view */    \
            int rows = n0 - (di ? 1 : 0);                                         \
            int cols = n1 - (dj ? 1 : 0);                                         \
            int ja0 = dj < 0 ? 1 : 0; /* (1,-1): a = h[:-1, 1:] */               \
            for (int phase = 0; phase < 2; phase++) {                             \
                REAL pmax = 0;                                                    \
                for (int i = 0; i < rows; i++)                                    \
                    for (int j = 0; j < cols; j++) {                              \
                        REAL a = h[(int64_t)i * n1 + j + ja0];                   \
                        REAL bb = h[(int64_t)(i + di) * n1 + j + ja0 + dj];       \
                        REAL diff = a - bb;                                       \
                        REAL ex = (REAL)fabs((double)diff) - crit;                \
                        ex = ex > 0 ? ex : 0;                                     \
                        pmax = pmax >= ex ? pmax : ex;                            \
                    }                                                             \
                for (int i = 0; i < rows; i++)                                    \
                    for (int j = 0; j < cols; j++) {                              \
                        int coord = di ? i : j;                                   \
                        if ((coord & 1) != phase) continue;                       \
                        int64_t ia = (int64_t)i * n1 + j + ja0;                   \
                        int64_t ib = (int64_t)(i + di) * n1 + j + ja0 + dj;       \
                        REAL diff = h[ia] - h[ib];                                \
                        REAL ex = (REAL)fabs((double)diff) - crit;                \
                        ex = ex > 0 ? ex : 0;                                     \
                        REAL sgn = diff > 0 ? (REAL)1 : (diff < 0 ? (REAL)-1 : (REAL)0); \
                        REAL t = ((REAL)0.5 * ex) * sgn;                          \
                        h[ia] = h[ia] - t;                                        \
                        h[ib] = h[ib] + t;                                        \
                    }                                                             \
                max_excess = max_excess >= (double)pmax ? max_excess : (double)pmax; \
            }                                                                     \
        }                                                                         \
        last = max_excess;                                                        \
        if (max_excess <= tolerance) { *converged = 1; break; }                   \
    }                                                                             \
    if (iteration > max_iterations) iteration = max_iterations;                   \
    if (max_excess_last) *max_excess_last = last;                                 \
    return iteration;

int orc_relax_legacy_f64(int n0, int n1, double d0, double d1, double crit_slope,
                         int max_iterations, double tolerance, double *h,
                         int *converged, double *max_excess_last)
{
    RELAX_BODY(double)
}

int orc_relax_legacy_f32(int n0, int n1, double d0, double d1, double crit_slope,
                         int max_iterations, double tolerance, float *h,
                         int *converged, double *max_excess_last)
{
    RELAX_BODY(float)
}

/* ------------------------------------------------------------------------ */
/* compact colour batch: warp_compact_frontier.py:29-130                     */
/* ------------------------------------------------------------------------ */
static void mark_tile_halo(int row, int col, int rows, int cols, int tile_size,
                           int tiles_x, int32_t *flags)
{
    for (int orow = -1; orow < 2; orow++)
        for (int ocol = -1; ocol < 2; ocol++) {
            int nr = row + orow, nc = col + ocol;
            if (nr >= 0 && nr < rows && nc >= 0 && nc < cols)
                flags[(nr / tile_size) * tiles_x + nc / tile_size] = 1;
        }
}

int orc_frontier_relax(int rows, int cols, int tile_size, int tiles_x,
                       double *height, const double *weights, int32_t *reached,
                       const int32_t *a_i, const int32_t *a_j,
                       const int32_t *b_i, const int32_t *b_j, int n_edges,
                       double critical_difference, int32_t *moved_tile_flags)
{
    int moved = 0;
    for (int e = 0; e < n_edges; e++) {
        int64_t a = (int64_t)a_i[e] * cols + a_j[e];
        int64_t b = (int64_t)b_i[e] * cols + b_j[e];
        if (reached[a] == 0 && reached[b] == 0) continue;
        double difference = height[a] - height[b];
        double excess = fabs(difference) - critical_difference;
        if (excess > 0.0) {
            double signed_excess = difference < 0.0 ? -excess : excess;
            double wa = weights[a], wb = weights[b];
            double ws = wa + wb;
            double ta = signed_excess * wb / ws;
            double tb = signed_excess * wa / ws;
            height[a] = height[a] - ta;
            height[b] = height[b] + tb;
            reached[a] = 1;
            reached[b] = 1;
            moved++;
            mark_tile_halo(a_i[e], a_j[e], rows, cols, tile_size, tiles_x, moved_tile_flags);
            mark_tile_halo(b_i[e], b_j[e], rows, cols, tile_size, tiles_x, moved_tile_flags);
        }
    }
    return moved;
}

void orc_frontier_scan(int cols, const double *height, const int32_t *reached,
                       const int32_t *tile_ids,
                       const int32_t *a_i, const int32_t *a_j,
                       const int32_t *b_i, const int32_t *b_j, int n_edges,
                       double critical_difference, double tolerance,
                       int32_t *unstable_tile_flags)
{
    for (int e = 0; e < n_edges; e++) {
        int64_t a = (int64_t)a_i[e] * cols + a_j[e];
        int64_t b = (int64_t)b_i[e] * cols + b_j[e];
        if (reached[a] != 0 || reached[b] != 0) {
            double excess = fabs(height[a] - height[b]) - critical_difference;
            if (excess > tolerance) unstable_tile_flags[tile_ids[e]] = 1;
        }
    }
}

/* ------------------------------------------------------------------------ */
/* LargeAvalanche device semantics: gpu_large_avalanche.py:59-162            */
/* ------------------------------------------------------------------------ */
void orc_avalanche_initialize(int ny, int nx, const orc_avalanche_params *p,
                              const double *z_base, const double *resting,
                              const double *mobile, const double *qx, const double *qy,
                              const double *weights, const double *owned_export_baseline,
                              const double *export_cumulative,
                              int32_t *unstable, int32_t *latch,
                              double *slope, double *grad_x, double *grad_y,
                              double *severity)
{
    for (int row = 0; row < ny; row++)
        for (int col = 0; col < nx; col++) {
            int64_t index = (int64_t)row * nx + col;
            int64_t left = index, right = index, down = index, up = index;
            double denom_x = p->dx, denom_y = p->dy;
            if (col > 0) left = index - 1;
            if (col + 1 < nx) right = index + 1;
            if (col > 0 && col + 1 < nx) denom_x = 2.0 * p->dx;
            if (row > 0) down = index - nx;
            if (row + 1 < ny) up = index + nx;
            if (row > 0 && row + 1 < ny) denom_y = 2.0 * p->dy;
            double gx = (resting[right] + mobile[right] - resting[left] - mobile[left]) / denom_x;
            double gy = (resting[up] + mobile[up] - resting[down] - mobile[down]) / denom_y;
            double magnitude = sqrt(gx * gx + gy * gy);
            grad_x[index] = gx;
            grad_y[index] = gy;
            slope[index] = atan(magnitude);
            double layer_depth = dmin(dmax(resting[index] - z_base[index], 0.0), p->mobilization_depth);
            double normalization = sqrt(1.0 + magnitude * magnitude);
            double drive = p->density * p->gravity * layer_depth * magnitude / normalization;
            double normal = p->density * p->gravity * layer_depth / normalization;
            double start_margin = drive - (p->cohesion + normal * p->tan_start);
            double stop_margin = drive - (p->cohesion + normal * p->tan_stop);
            double s = stop_margin / dmax(normal * (p->tan_start - p->tan_stop), 1.0e-12);
            severity[index] = s < 0.0 ? 0.0 : (s > 1.0 ? 1.0 : s);
            int active = layer_depth > p->dry_tolerance && start_margin > 0.0;
            unstable[index] = active ? 1 : 0;
            double speed = 0.0;
            if (mobile[index] > p->dry_tolerance) {
                double vx = qx[index] / mobile[index], vy = qy[index] / mobile[index];
                speed = sqrt(vx * vx + vy * vy);
            }
            double exported = export_cumulative[index] - owned_export_baseline[index];
            int resolved = exported > p->dry_tolerance * weights[index];
            if (p->release_settled_latches && speed <= p->activity_speed &&
                (resolved || stop_margin <= 0.0))
                latch[index] = 0;
        }
}

static int32_t uf_find(int32_t *parent, int32_t x)
{
    while (parent[x] != x) x = parent[x];
    return x;
}

/* canonical labels via union-find that always hooks the larger root onto
 * the smaller one (gpu_large_avalanche.py:40-56 semantics) */
void orc_label(int ny, int nx, int connectivity, const int32_t *mask, int32_t *labels)
{
    int64_t n = (int64_t)ny * nx;
    int32_t *parent = labels;
    for (int64_t i = 0; i < n; i++) parent[i] = mask[i] ? (int32_t)i : -1;
    for (int r = 0; r < ny; r++)
        for (int c = 0; c < nx; c++) {
            int32_t i = r * nx + c;
            if (parent[i] < 0) continue;
            int32_t nb[4];
            int k = 0;
            if (c + 1 < nx) nb[k++] = i + 1;
            if (r + 1 < ny) nb[k++] = i + nx;
            if (connectivity == 8 && r + 1 < ny) {
                if (c + 1 < nx) nb[k++] = i + nx + 1;
                if (c > 0) nb[k++] = i + nx - 1;
            }
            for (int t = 0; t < k; t++) {
                int32_t j = nb[t];
                if (parent[j] < 0) continue;
                int32_t ra = uf_find(parent, i), rb = uf_find(parent, j);
                if (ra == rb) continue;
                if (ra < rb) parent[rb] = ra; else parent[ra] = rb;
            }
        }
    for (int64_t i = 0; i < n; i++)
        if (parent[i] >= 0) parent[i] = uf_find(parent, (int32_t)i);
}

/* aggregate_components/reduce_components/select_winner/winner_diagnostics
 * (gpu_large_avalanche.py:201-281) with index-order sums.
 * diag_i: [regions, largest, winner, unstable, previous-overlap]
 * diag_f: [sum excess_start, max excess_start, sum excess_stop,
 *          max excess_stop, area, mobilizable volume, root excess sum]   */
void orc_components(int ny, int nx, const int32_t *labels,
                    const double *z_base, const double *resting, const int32_t *latch,
                    const double *weights, const double *slope, const double *severity,
                    const int32_t *previous, double mobilization_depth,
                    double start_rad, double stop_rad,
                    int32_t diag_i[8], double diag_f[8])
{
    int64_t n = (int64_t)ny * nx;
    int32_t *counts = (int32_t *)calloc((size_t)n, sizeof(int32_t));
    for (int k = 0; k < 8; k++) { diag_i[k] = 0; diag_f[k] = 0.0; }
    diag_i[2] = (int32_t)(n + 1);
    for (int64_t i = 0; i < n; i++)
        if (labels[i] >= 0) counts[labels[i]]++;
    for (int64_t i = 0; i < n; i++) {
        if (counts[i] > 0) {
            diag_i[0]++;
            diag_i[3] += counts[i];
            if (counts[i] > diag_i[1]) diag_i[1] = counts[i];
        }
    }
    for (int64_t i = 0; i < n; i++)
        if (diag_i[1] > 0 && counts[i] == diag_i[1] && i < diag_i[2]) diag_i[2] = (int32_t)i;
    int32_t winner = diag_i[2];
    for (int64_t i = 0; i < n; i++) {
        if (labels[i] < 0 || labels[i] != winner) continue;
        diag_i[4] += previous ? previous[i] : 0;
        double es = dmax(slope[i] - start_rad, 0.0);
        double et = dmax(slope[i] - stop_rad, 0.0);
        diag_f[0] += es;
        diag_f[1] = dmax(diag_f[1], es);
        diag_f[2] += et;
        diag_f[3] = dmax(diag_f[3], et);
        double transfer = 0.0;
        if (latch[i] == 0)
            transfer = dmin(dmax(resting[i] - z_base[i], 0.0), mobilization_depth * severity[i]);
        diag_f[4] += weights[i];
        diag_f[5] += transfer * weights[i];
        diag_f[6] += es;
    }
    free(counts);
}

/* reduce_persistence (gpu_large_avalanche.py:378-451) in cell order; the
 * maxima start at zero like the reference's zeroed accumulators */
void orc_persistence(int ny, int nx, int tile_size, int tiles_x, const double *initial_resting,
                     const double *resting, const double *mobile, const double *qx, const double *qy,
                     const double *weights, const double *slope, const int32_t *unstable,
                     const int32_t *latch, const int32_t *activation_count, const double *r2m,
                     const double *m2r, double start_rad, double stop_rad, double dry,
                     double activity_speed, int32_t diag_i[4], double diag_f[14], int32_t *tile_flags)
{
    for (int k = 0; k < 4; k++) diag_i[k] = 0;
    for (int k = 0; k < 14; k++) diag_f[k] = 0.0;
    int tiles_y = (ny + tile_size - 1) / tile_size;
    for (int t = 0; t < tiles_y * tiles_x; t++) tile_flags[t] = 0;
    for (int row = 0; row < ny; row++)
        for (int col = 0; col < nx; col++) {
            size_t i = (size_t)row * nx + col;
            int count = activation_count[i];
            double h = mobile[i];
            double speed = 0.0;
            double w = weights[i];
            if (h > dry) {
                double vx = qx[i] / h, vy = qy[i] / h;
                speed = sqrt(vx * vx + vy * vy);
                diag_i[2] += 1;
                diag_f[4] += h * w;
                diag_f[5] = npmax(diag_f[5], speed);
                if (speed > activity_speed) diag_f[13] += h * w;
            }
            if (unstable[i] != 0 || (h > dry && speed > activity_speed))
                tile_flags[(row / tile_size) * tiles_x + col / tile_size] = 1;
            double delta = resting[i] + h - initial_resting[i];
            diag_f[6] += delta * w;
            diag_f[7] += fabs(delta) * w * 0.5;
            diag_f[0] += r2m[i];
            diag_f[1] += m2r[i];
            if (count <= 0) continue;
            double s = slope[i];
            diag_i[0] += 1;
            diag_f[2] += w;
            diag_f[8] += s;
            diag_f[9] = npmax(diag_f[9], s);
            diag_f[10] += s - stop_rad;
            diag_f[11] += start_rad - s;
            if (count > 1) {
                diag_i[1] += 1;
                diag_f[3] += w;
            }
            if (h <= dry && s <= stop_rad && latch[i] == 0) {
                diag_i[3] += 1;
                diag_f[12] += w;
            }
        }
}

/* mobilize (gpu_large_avalanche.py:309-376); diag_f indices as there */
void orc_mobilize(int n_cells, const int32_t *labels, int32_t winner,
                  const double *z_base, double *resting, double *mobile,
                  const double *weights, int32_t *latch, double *owned_surface,
                  double *owned_export_baseline, const double *export_cumulative,
                  const double *severity, int32_t *activation_count,
                  double *first_time, double *last_time, double *r2m_cumulative,
                  double activation_time, double mobilization_depth, double dry,
                  double diag_f[16])
{
    for (int i = 0; i < n_cells; i++) {
        if (labels[i] < 0 || labels[i] != winner || latch[i] != 0) continue;
        double transfer = dmin(dmax(resting[i] - z_base[i], 0.0), mobilization_depth * severity[i]);
        if (transfer <= dry) continue;
        owned_surface[i] = resting[i] + mobile[i];
        owned_export_baseline[i] = export_cumulative[i];
        resting[i] = resting[i] - transfer;
        mobile[i] = mobile[i] + transfer;
        latch[i] = 1;
        double weighted = transfer * weights[i];
        int32_t prev = activation_count[i];
        activation_count[i] = prev + 1;
        if (prev == 0) {
            first_time[i] = activation_time;
            diag_f[10] += weighted;
            diag_f[12] += weights[i];
        } else {
            diag_f[11] += weighted;
            diag_f[13] += weights[i];
        }
        last_time[i] = activation_time;
        r2m_cumulative[i] = r2m_cumulative[i] + weighted;
        diag_f[7] += weighted;
    }
}

/* single shared face, exposed for the one-face contract check
 * (tools/run_mobile_v2_reference_contracts.py:106-112) */
void orc_v2_face(double bl, double br, double hl, double hr,
                 double qxl, double qyl, double qxr, double qyr,
                 int normal, double K, double g, double dry, double factor, double out[5])
{
    face_out f = v2_face(bl, br, hl, hr, qxl, qyl, qxr, qyr, normal, K, g, dry, factor);
    out[0] = f.h; out[1] = f.qxl; out[2] = f.qyl; out[3] = f.qxr; out[4] = f.qyr;
}
