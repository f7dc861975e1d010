// Mobile V2 finite-volume step on the device (gfx950).
//
// One cooperative (co-resident) launch runs a whole step_resident call:
// ledger + static diagnostic + first CFL reduction, then the data-dependent
// substep loop with ONE grid barrier per substep, then the final ledger.
// The global CFL maximum of the state produced by substep n is reduced while
// that state is written (wave shuffles + one atomicMax per workgroup), so the
// host never sees the loop.  Fields ping-pong between the caller's arrays and
// library scratch; an odd substep count ends with a copy back.
//
// Numerics follow pkg/experimental/mobile_v2_reference.py (CPU reference):
//   FP64 mode: every face and cell update repeats the reference's operation
//              order (mobile_v2_reference.py:236-307, :365-449) with
//              -ffp-contract=off, so fields match the CPU oracle bit for bit.
//   FP32 mode: float fluxes; mobile thickness in int32 quanta of 2^-28 m.
//              Each face's mass flux is rounded once to whole quanta and added
//              to one cell / removed from the other, so total mass is exact;
//              outflow through one face is capped at a quarter of the donor's
//              quanta (never active in exact arithmetic for CFL <= 0.25), so
//              thickness stays non-negative without a mass-creating clamp.
#include "mv2_internal.hpp"

#include <cmath>

namespace mv2 {

struct SyncBlock {
    unsigned counter;
    unsigned pad0;
    unsigned long long wave_slot[3];     // double bits of the CFL maximum
    unsigned long long static_grad;      // double bits
    unsigned long long static_q;         // double bits
    int positivity;                      // 1 once any cell failed positivity
    int overflow;                        // 1 once a thickness left the int32 range
    int pad1[2];
};

// per-workgroup partial record
constexpr int kSum = 8;                  // summary entries
constexpr int kDiag = 10;
constexpr int kClasses = 9;              // weight classes of the exact FP32 mass ledger
struct Partial {
    double before[kSum];
    double after[kSum];
    double diag[kDiag];
    long long mass_before[kClasses];
    long long mass_after[kClasses];
};

struct F64P {
    using real = double;
    using hst = double;
    static constexpr bool kQuant = false;
    __device__ static double depth(double v) { return v; }
};
struct F32P {
    using real = float;
    using hst = int32_t;
    static constexpr bool kQuant = true;
    __device__ static float depth(int32_t v) { return (float)v * (float)kQuantum; }
};

template <class P>
struct V2Args {
    int ny, nx;
    double dx, dy;
    const typename P::real *b;
    typename P::hst *h[2];
    typename P::real *qx[2];
    typename P::real *qy[2];
    mv2_step_config cfg;
    SyncBlock *sync;
    Partial *partials;
    mv2_step_record *rec;
    int clamp_quarter;                   // FP32: cap outflow per face at h/4
};

__device__ inline double npmax(double a, double b) { return a >= b ? a : b; }
__device__ inline double npmin(double a, double b) { return a <= b ? a : b; }

// ---------------------------------------------------------------------------
// FP64 face: mobile_v2_reference.py:391-431 (+ _physical_flux :433-449).
// Returns values already multiplied by factor = dt/spacing.
// ---------------------------------------------------------------------------
struct Face64 { double h, qxl, qyl, qxr, qyr; };

template <int NORMAL>
__device__ inline Face64 face64(double bl, double br, double hl, double hr,
                                double qxl, double qyl, double qxr, double qyr,
                                double K, double g, double dry, double factor)
{
    const double Kg = K * g;
    const double c = 0.5 * K * g;
    double head_l = bl + K * hl;
    double head_r = br + K * hr;
    double b_face = npmax(bl, br);
    double hs_l = npmax((head_l - b_face) / K, 0.0);
    double hs_r = npmax((head_r - b_face) / K, 0.0);
    double scale_l = hl > dry ? hs_l / hl : 0.0;
    double scale_r = hr > dry ? hs_r / hr : 0.0;
    double qsx_l = qxl * scale_l, qsy_l = qyl * scale_l;
    double qsx_r = qxr * scale_r, qsy_r = qyr * scale_r;
    double vx_l = hs_l > dry ? qsx_l / hs_l : 0.0;
    double vy_l = hs_l > dry ? qsy_l / hs_l : 0.0;
    double vx_r = hs_r > dry ? qsx_r / hs_r : 0.0;
    double vy_r = hs_r > dry ? qsy_r / hs_r : 0.0;
    double qn_l = NORMAL == 0 ? qsx_l : qsy_l;
    double qn_r = NORMAL == 0 ? qsx_r : qsy_r;
    double fl0 = qn_l, fl1 = qn_l * vx_l, fl2 = qn_l * vy_l;
    double fr0 = qn_r, fr1 = qn_r * vx_r, fr2 = qn_r * vy_r;
    if (NORMAL == 0) {
        fl1 = fl1 + (c * hs_l) * hs_l;
        fr1 = fr1 + (c * hs_r) * hs_r;
    } else {
        fl2 = fl2 + (c * hs_l) * hs_l;
        fr2 = fr2 + (c * hs_r) * hs_r;
    }
    double ul = hs_l > dry ? qn_l / hs_l : 0.0;
    double ur = hs_r > dry ? qn_r / hs_r : 0.0;
    double speed = npmax(fabs(ul) + sqrt(Kg * hs_l), fabs(ur) + sqrt(Kg * hs_r));
    double hs = 0.5 * speed;
    double flux0 = 0.5 * (fl0 + fr0) - hs * (hs_r - hs_l);
    double flux1 = 0.5 * (fl1 + fr1) - hs * (qsx_r - qsx_l);
    double flux2 = 0.5 * (fl2 + fr2) - hs * (qsy_r - qsy_l);
    double p_l = (c * hl) * hl, p_r = (c * hr) * hr;
    double ps_l = (c * hs_l) * hs_l, ps_r = (c * hs_r) * hs_r;
    double m1 = flux1, m2 = flux2, p1 = flux1, p2 = flux2;
    if (NORMAL == 0) {
        m1 = m1 + (p_l - ps_l);
        p1 = p1 + (p_r - ps_r);
    } else {
        m2 = m2 + (p_l - ps_l);
        p2 = p2 + (p_r - ps_r);
    }
    Face64 o;
    o.h = factor * flux0;
    o.qxl = factor * m1;
    o.qyl = factor * m2;
    o.qxr = factor * p1;
    o.qyr = factor * p2;
    return o;
}

// ---------------------------------------------------------------------------
// FP32 face: same equations in float, velocities from q/h via the hardware
// reciprocal, pressure correction through the shared p - ps form, mass flux
// rounded once to whole quanta.  Deterministic: both cells that share the
// face evaluate it from identical operands with identical instructions.
// ---------------------------------------------------------------------------
struct Face32 { int h; float qxl, qyl, qxr, qyr; };

struct Consts32 {
    float K, invK, Kg, c, dry;
};

__device__ inline float rcpf(float x) { return __builtin_amdgcn_rcpf(x); }

template <int NORMAL>
__device__ inline Face32 face32(float bl, float br, int hql, int hqr, float hl, float hr,
                                float qxl, float qyl, float qxr, float qyr,
                                const Consts32 &k, float factor, int clamp_quarter)
{
    float head_l = bl + k.K * hl;
    float head_r = br + k.K * hr;
    float b_face = fmaxf(bl, br);
    float hs_l = fmaxf((head_l - b_face) * k.invK, 0.0f);
    float hs_r = fmaxf((head_r - b_face) * k.invK, 0.0f);
    // reconstructed velocity equals the cell velocity; zero when either
    // the cell or its reconstruction is dry (reference: scale=0 / v=0)
    float il = rcpf(hl), ir = rcpf(hr);
    bool wl = (hl > k.dry) && (hs_l > k.dry);
    bool wr = (hr > k.dry) && (hs_r > k.dry);
    float vxl = wl ? qxl * il : 0.0f, vyl = wl ? qyl * il : 0.0f;
    float vxr = wr ? qxr * ir : 0.0f, vyr = wr ? qyr * ir : 0.0f;
    float vnl = NORMAL == 0 ? vxl : vyl, vtl = NORMAL == 0 ? vyl : vxl;
    float vnr = NORMAL == 0 ? vxr : vyr, vtr = NORMAL == 0 ? vyr : vxr;
    float qnl = hs_l * vnl, qtl = hs_l * vtl;
    float qnr = hs_r * vnr, qtr = hs_r * vtr;
    float psl = (k.c * hs_l) * hs_l, psr = (k.c * hs_r) * hs_r;
    float fln = qnl * vnl + psl, frn = qnr * vnr + psr;
    float flt = qnl * vtl, frt = qnr * vtr;
    float speed = fmaxf(fabsf(vnl) + __builtin_sqrtf(k.Kg * hs_l),
                        fabsf(vnr) + __builtin_sqrtf(k.Kg * hs_r));
    float hsp = 0.5f * speed;
    float fh = 0.5f * (qnl + qnr) - hsp * (hs_r - hs_l);
    float fn = 0.5f * (fln + frn) - hsp * (qnr - qnl);
    float ft = 0.5f * (flt + frt) - hsp * (qtr - qtl);
    float pl = (k.c * hl) * hl, pr = (k.c * hr) * hr;
    Face32 o;
    float mq = rintf(fh * factor * (float)kInvQuantum);
    int q = (int)mq;
    if (clamp_quarter) {
        if (q > 0) q = min(q, hql >> 2);
        else q = max(q, -(hqr >> 2));
    }
    o.h = q;
    float nl = (fn + (pl - psl)) * factor;
    float nr = (fn + (pr - psr)) * factor;
    float tt = ft * factor;
    if (NORMAL == 0) {
        o.qxl = nl; o.qyl = tt; o.qxr = nr; o.qyr = tt;
    } else {
        o.qxl = tt; o.qyl = nl; o.qxr = tt; o.qyr = nr;
    }
    return o;
}

// ---------------------------------------------------------------------------
// per-cell work of one substep
// ---------------------------------------------------------------------------
struct CellOut {
    double wave;      // CFL wave speed of the new state
    double hmin;      // new thickness (m) before clamping, for positivity
    int overflow;
};

template <class P>
struct Acc {
    double diag[kDiag];
};

// FP64: reference operation order, gathering the four faces of the cell.
__device__ inline CellOut cell64(const V2Args<F64P> &a, int src, int r, int c, double sub_dt,
                                 Acc<F64P> &acc)
{
    const int nx = a.nx, ny = a.ny;
    const int64_t i = (int64_t)r * nx + c;
    const double K = a.cfg.earth_pressure, g = a.cfg.gravity, dry = a.cfg.dry_tolerance;
    const double *b = a.b;
    const double *h = a.h[src];
    const double *qx = a.qx[src];
    const double *qy = a.qy[src];
    const double fxf = sub_dt / a.dx, fyf = sub_dt / a.dy;
    double bi = b[i], hi = h[i], qxi = qx[i], qyi = qy[i];
    double dh = 0.0, dqx = 0.0, dqy = 0.0;
    if (c + 1 < nx) {
        Face64 f = face64<0>(bi, b[i + 1], hi, h[i + 1], qxi, qyi, qx[i + 1], qy[i + 1], K, g, dry, fxf);
        dh = dh - f.h; dqx = dqx - f.qxl; dqy = dqy - f.qyl;
    }
    if (c > 0) {
        Face64 f = face64<0>(b[i - 1], bi, h[i - 1], hi, qx[i - 1], qy[i - 1], qxi, qyi, K, g, dry, fxf);
        dh = dh + f.h; dqx = dqx + f.qxr; dqy = dqy + f.qyr;
    }
    if (r + 1 < ny) {
        Face64 f = face64<1>(bi, b[i + nx], hi, h[i + nx], qxi, qyi, qx[i + nx], qy[i + nx], K, g, dry, fyf);
        dh = dh - f.h; dqx = dqx - f.qxl; dqy = dqy - f.qyl;
    }
    if (r > 0) {
        Face64 f = face64<1>(b[i - nx], bi, h[i - nx], hi, qx[i - nx], qy[i - nx], qxi, qyi, K, g, dry, fyf);
        dh = dh + f.h; dqx = dqx + f.qxr; dqy = dqy + f.qyr;
    }
    double hn = hi + dh;
    double qxn = qxi + dqx, qyn = qyi + dqy;
    CellOut out;
    out.hmin = hn;
    out.overflow = 0;
    acc.diag[0] += qxn - qxi;
    acc.diag[1] += qyn - qyi;
    double depth = npmax(hn, 0.0);
    const double ax = a.cfg.external_ax, ay = a.cfg.external_ay;
    const double mug = a.cfg.basal_friction * g;
    double qax = 0.0, qay = 0.0;
    if (depth > dry) {
        double vbx = qxn / depth, vby = qyn / depth;
        double qex = qxn + depth * ax * sub_dt;
        double qey = qyn + depth * ay * sub_dt;
        double vex = qex / depth, vey = qey / depth;
        acc.diag[2] += qex - qxn;
        acc.diag[3] += qey - qyn;
        acc.diag[6] += depth * ((0.5 * (vbx + vex)) * ax + (0.5 * (vby + vey)) * ay) * sub_dt;
        double speed = sqrt(vex * vex + vey * vey);
        double decrement = mug * sub_dt;
        double factor = npmax(0.0, 1.0 - decrement / npmax(speed, 1.0e-15));
        double vax = vex * factor, vay = vey * factor;
        qax = depth * vax;
        qay = depth * vay;
        acc.diag[4] += qax - qex;
        acc.diag[5] += qay - qey;
        acc.diag[7] += depth * (speed * speed - (vax * vax + vay * vay));
    }
    // sum(q_new) - sum(q_old) with dry cells zeroed (mobile_v2_reference.py:258-272)
    acc.diag[8] += (depth > dry ? qxn : 0.0) - qxi;
    acc.diag[9] += (depth > dry ? qyn : 0.0) - qyi;
    const int dst = src ^ 1;
    a.h[dst][i] = depth;
    a.qx[dst][i] = qax;
    a.qy[dst][i] = qay;
    // CFL wave speed of the new state (mobile_v2_reference.py:239-247)
    double vx = depth > dry ? qax / depth : 0.0;
    double vy = depth > dry ? qay / depth : 0.0;
    out.wave = sqrt(vx * vx + vy * vy) + sqrt((K * g) * depth);
    return out;
}

__device__ inline CellOut cell32(const V2Args<F32P> &a, int src, int r, int c, double sub_dt,
                                 Acc<F32P> &acc)
{
    const int nx = a.nx, ny = a.ny;
    const int64_t i = (int64_t)r * nx + c;
    Consts32 k;
    k.K = (float)a.cfg.earth_pressure;
    k.invK = (float)(1.0 / a.cfg.earth_pressure);
    k.Kg = (float)(a.cfg.earth_pressure * a.cfg.gravity);
    k.c = (float)(0.5 * a.cfg.earth_pressure * a.cfg.gravity);
    k.dry = (float)a.cfg.dry_tolerance;
    const float *b = a.b;
    const int32_t *h = a.h[src];
    const float *qx = a.qx[src];
    const float *qy = a.qy[src];
    const float fxf = (float)(sub_dt / a.dx), fyf = (float)(sub_dt / a.dy);
    const float Qf = (float)kQuantum;
    float bi = b[i];
    int hqi = h[i];
    float hi = (float)hqi * Qf;
    float qxi = qx[i], qyi = qy[i];
    long long dq = 0;
    float dqx = 0.0f, dqy = 0.0f;
    const int cq = a.clamp_quarter;
    if (c + 1 < nx) {
        int hq2 = h[i + 1];
        Face32 f = face32<0>(bi, b[i + 1], hqi, hq2, hi, (float)hq2 * Qf, qxi, qyi, qx[i + 1], qy[i + 1], k, fxf, cq);
        dq -= f.h; dqx = dqx - f.qxl; dqy = dqy - f.qyl;
    }
    if (c > 0) {
        int hq2 = h[i - 1];
        Face32 f = face32<0>(b[i - 1], bi, hq2, hqi, (float)hq2 * Qf, hi, qx[i - 1], qy[i - 1], qxi, qyi, k, fxf, cq);
        dq += f.h; dqx = dqx + f.qxr; dqy = dqy + f.qyr;
    }
    if (r + 1 < ny) {
        int hq2 = h[i + nx];
        Face32 f = face32<1>(bi, b[i + nx], hqi, hq2, hi, (float)hq2 * Qf, qxi, qyi, qx[i + nx], qy[i + nx], k, fyf, cq);
        dq -= f.h; dqx = dqx - f.qxl; dqy = dqy - f.qyl;
    }
    if (r > 0) {
        int hq2 = h[i - nx];
        Face32 f = face32<1>(b[i - nx], bi, hq2, hqi, (float)hq2 * Qf, hi, qx[i - nx], qy[i - nx], qxi, qyi, k, fyf, cq);
        dq += f.h; dqx = dqx + f.qxr; dqy = dqy + f.qyr;
    }
    long long hn = (long long)hqi + dq;
    CellOut out;
    out.overflow = hn > 2147483647LL;
    out.hmin = (double)hn * kQuantum;
    int hq = (int)(hn < 0 ? 0 : (hn > 2147483647LL ? 2147483647LL : hn));
    float qxn = qxi + dqx, qyn = qyi + dqy;
    acc.diag[0] += (double)(qxn - qxi);
    acc.diag[1] += (double)(qyn - qyi);
    float depth = (float)hq * Qf;
    float qax = 0.0f, qay = 0.0f, wave_v = 0.0f;
    if (depth > k.dry) {
        const float dt = (float)sub_dt;
        const float ax = (float)a.cfg.external_ax, ay = (float)a.cfg.external_ay;
        float inv = rcpf(depth);
        float vbx = qxn * inv, vby = qyn * inv;
        float vex = vbx + ax * dt, vey = vby + ay * dt;
        float speed = __builtin_sqrtf(vex * vex + vey * vey);
        float decrement = (float)(a.cfg.basal_friction * a.cfg.gravity) * dt;
        float factor = speed > 1.0e-15f ? fmaxf(0.0f, 1.0f - decrement * rcpf(speed)) : 0.0f;
        float vax = vex * factor, vay = vey * factor;
        qax = depth * vax;
        qay = depth * vay;
        wave_v = speed * factor;
        double dd = (double)depth;
        acc.diag[2] += dd * (double)(vex - vbx);
        acc.diag[3] += dd * (double)(vey - vby);
        acc.diag[6] += dd * (0.5 * (double)(vbx + vex) * a.cfg.external_ax +
                             0.5 * (double)(vby + vey) * a.cfg.external_ay) * sub_dt;
        acc.diag[4] += (double)qax - dd * (double)vex;
        acc.diag[5] += (double)qay - dd * (double)vey;
        acc.diag[7] += dd * ((double)speed * speed - ((double)vax * vax + (double)vay * vay));
    }
    acc.diag[8] += (double)((depth > k.dry ? qxn : 0.0f) - qxi);
    acc.diag[9] += (double)((depth > k.dry ? qyn : 0.0f) - qyi);
    const int dst = src ^ 1;
    a.h[dst][i] = hq;
    a.qx[dst][i] = qax;
    a.qy[dst][i] = qay;
    out.wave = (double)(wave_v + __builtin_sqrtf(k.Kg * depth));
    return out;
}

// ---------------------------------------------------------------------------
// FP32 tile substep: each shared face is evaluated once per tile from LDS.
// A 16 x 32 tile stages an (18 x 34) halo block of cells (depth, quanta,
// momentum, wet-masked velocity), computes its 16 x 33 x-faces and 17 x 32
// y-faces into LDS, then updates its cells.  Faces on a tile border are
// computed by both neighbouring tiles from bit-identical operands, so both
// sides move the same integer number of quanta.
// ---------------------------------------------------------------------------
constexpr int kTH = 16, kTW = 32;
constexpr int kLH = kTH + 2, kLW = kTW + 2;
constexpr int kFX = kTH * (kTW + 1);
constexpr int kFY = (kTH + 1) * kTW;

struct Tile32 {
    float b[kLH * kLW];
    float h[kLH * kLW];
    int hq[kLH * kLW];
    float qx[kLH * kLW];
    float qy[kLH * kLW];
    float vx[kLH * kLW];
    float vy[kLH * kLW];
    int fxq[kFX];
    float fxl[kFX], fxr[kFX], fxt[kFX];
    int fyq[kFY];
    float fyl[kFY], fyr[kFY], fyt[kFY];
};

// face from cell-centred velocities (zero for dry cells); the reconstructed
// velocity is the cell velocity, zeroed where the reconstruction is dry
template <int NORMAL>
__device__ inline Face32 face32v(float bl, float br, int hql, int hqr, float hl, float hr,
                                 float vxl, float vyl, float vxr, float vyr,
                                 const Consts32 &k, float factor, int clamp_quarter)
{
    float head_l = bl + k.K * hl;
    float head_r = br + k.K * hr;
    float b_face = fmaxf(bl, br);
    float hs_l = fmaxf((head_l - b_face) * k.invK, 0.0f);
    float hs_r = fmaxf((head_r - b_face) * k.invK, 0.0f);
    bool wl = hs_l > k.dry, wr = hs_r > k.dry;
    float vnl = wl ? (NORMAL == 0 ? vxl : vyl) : 0.0f, vtl = wl ? (NORMAL == 0 ? vyl : vxl) : 0.0f;
    float vnr = wr ? (NORMAL == 0 ? vxr : vyr) : 0.0f, vtr = wr ? (NORMAL == 0 ? vyr : vxr) : 0.0f;
    float qnl = hs_l * vnl, qtl = hs_l * vtl;
    float qnr = hs_r * vnr, qtr = hs_r * vtr;
    float psl = (k.c * hs_l) * hs_l, psr = (k.c * hs_r) * hs_r;
    float fln = qnl * vnl + psl, frn = qnr * vnr + psr;
    float flt = qnl * vtl, frt = qnr * vtr;
    float speed = fmaxf(fabsf(vnl) + __builtin_sqrtf(k.Kg * hs_l),
                        fabsf(vnr) + __builtin_sqrtf(k.Kg * hs_r));
    float hsp = 0.5f * speed;
    float fh = 0.5f * (qnl + qnr) - hsp * (hs_r - hs_l);
    float fn = 0.5f * (fln + frn) - hsp * (qnr - qnl);
    float ft = 0.5f * (flt + frt) - hsp * (qtr - qtl);
    float pl = (k.c * hl) * hl, pr = (k.c * hr) * hr;
    Face32 o;
    int q = (int)rintf(fh * factor * (float)kInvQuantum);
    if (clamp_quarter) {
        if (q > 0) q = min(q, hql >> 2);
        else q = max(q, -(hqr >> 2));
    }
    o.h = q;
    float nl = (fn + (pl - psl)) * factor;
    float nr = (fn + (pr - psr)) * factor;
    float tt = ft * factor;
    if (NORMAL == 0) {
        o.qxl = nl; o.qyl = tt; o.qxr = nr; o.qyr = tt;
    } else {
        o.qxl = tt; o.qyl = nl; o.qxr = tt; o.qyr = nr;
    }
    return o;
}

__device__ inline Consts32 consts32(const mv2_step_config &cfg)
{
    Consts32 k;
    k.K = (float)cfg.earth_pressure;
    k.invK = (float)(1.0 / cfg.earth_pressure);
    k.Kg = (float)(cfg.earth_pressure * cfg.gravity);
    k.c = (float)(0.5 * cfg.earth_pressure * cfg.gravity);
    k.dry = (float)cfg.dry_tolerance;
    return k;
}

// tile t (row-major over tiles_x) -> substep update of its cells
__device__ void tile_substep32(const V2Args<F32P> &a, Tile32 &T, int src, int t, int tiles_x,
                               double sub_dt, Acc<F32P> &acc, double &wmax, double &hmin, int &ovf)
{
    const int ny = a.ny, nx = a.nx;
    const int r0 = (t / tiles_x) * kTH, c0 = (t % tiles_x) * kTW;
    const Consts32 k = consts32(a.cfg);
    const float Qf = (float)kQuantum;
    const float *gb = a.b;
    const int32_t *gh = a.h[src];
    const float *gqx = a.qx[src];
    const float *gqy = a.qy[src];
    for (int e = threadIdx.x; e < kLH * kLW; e += blockDim.x) {
        const int lr = e / kLW, lc = e - lr * kLW;
        const int gr = r0 - 1 + lr, gc = c0 - 1 + lc;
        float b = 0.0f, h = 0.0f, qx = 0.0f, qy = 0.0f, vx = 0.0f, vy = 0.0f;
        int hq = 0;
        if (gr >= 0 && gr < ny && gc >= 0 && gc < nx) {
            const int64_t i = (int64_t)gr * nx + gc;
            b = gb[i];
            hq = gh[i];
            qx = gqx[i];
            qy = gqy[i];
            h = (float)hq * Qf;
            if (h > k.dry) {
                float inv = rcpf(h);
                vx = qx * inv;
                vy = qy * inv;
            }
        }
        T.b[e] = b; T.h[e] = h; T.hq[e] = hq; T.qx[e] = qx; T.qy[e] = qy; T.vx[e] = vx; T.vy[e] = vy;
    }
    __syncthreads();
    const float fxf = (float)(sub_dt / a.dx), fyf = (float)(sub_dt / a.dy);
    const int cq = a.clamp_quarter;
    for (int e = threadIdx.x; e < kFX + kFY; e += blockDim.x) {
        if (e < kFX) {
            const int lr = e / (kTW + 1), kk = e - lr * (kTW + 1);
            const int gr = r0 + lr, gcl = c0 - 1 + kk;
            if (gr >= ny || gcl < 0 || gcl + 1 >= nx) continue;
            const int L = (lr + 1) * kLW + kk, R = L + 1;
            Face32 f = face32v<0>(T.b[L], T.b[R], T.hq[L], T.hq[R], T.h[L], T.h[R], T.vx[L], T.vy[L],
                                  T.vx[R], T.vy[R], k, fxf, cq);
            T.fxq[e] = f.h; T.fxl[e] = f.qxl; T.fxr[e] = f.qxr; T.fxt[e] = f.qyl;
        } else {
            const int ey = e - kFX;
            const int kk = ey / kTW, lc = ey - kk * kTW;
            const int grl = r0 - 1 + kk, gc = c0 + lc;
            if (gc >= nx || grl < 0 || grl + 1 >= ny) continue;
            const int L = kk * kLW + lc + 1, R = L + kLW;
            Face32 f = face32v<1>(T.b[L], T.b[R], T.hq[L], T.hq[R], T.h[L], T.h[R], T.vx[L], T.vy[L],
                                  T.vx[R], T.vy[R], k, fyf, cq);
            T.fyq[ey] = f.h; T.fyl[ey] = f.qyl; T.fyr[ey] = f.qyr; T.fyt[ey] = f.qxl;
        }
    }
    __syncthreads();
    const int dst = src ^ 1;
    for (int e = threadIdx.x; e < kTH * kTW; e += blockDim.x) {
        const int ir = e / kTW, ic = e - ir * kTW;
        const int gr = r0 + ir, gc = c0 + ic;
        if (gr >= ny || gc >= nx) continue;
        const int L = (ir + 1) * kLW + ic + 1;
        long long dq = 0;
        float dqx = 0.0f, dqy = 0.0f;
        if (gc + 1 < nx) {
            const int f = ir * (kTW + 1) + ic + 1;
            dq -= T.fxq[f]; dqx = dqx - T.fxl[f]; dqy = dqy - T.fxt[f];
        }
        if (gc > 0) {
            const int f = ir * (kTW + 1) + ic;
            dq += T.fxq[f]; dqx = dqx + T.fxr[f]; dqy = dqy + T.fxt[f];
        }
        if (gr + 1 < ny) {
            const int f = (ir + 1) * kTW + ic;
            dq -= T.fyq[f]; dqx = dqx - T.fyt[f]; dqy = dqy - T.fyl[f];
        }
        if (gr > 0) {
            const int f = ir * kTW + ic;
            dq += T.fyq[f]; dqx = dqx + T.fyt[f]; dqy = dqy + T.fyr[f];
        }
        const int hqi = T.hq[L];
        const float qxi = T.qx[L], qyi = T.qy[L];
        long long hn = (long long)hqi + dq;
        ovf |= hn > 2147483647LL;
        hmin = npmin(hmin, (double)hn * kQuantum);
        int hq = (int)(hn < 0 ? 0 : (hn > 2147483647LL ? 2147483647LL : hn));
        float qxn = qxi + dqx, qyn = qyi + dqy;
        acc.diag[0] += (double)(qxn - qxi);
        acc.diag[1] += (double)(qyn - qyi);
        float depth = (float)hq * Qf;
        float qax = 0.0f, qay = 0.0f, wave_v = 0.0f;
        if (depth > k.dry) {
            const float dt = (float)sub_dt;
            const float ax = (float)a.cfg.external_ax, ay = (float)a.cfg.external_ay;
            float inv = rcpf(depth);
            float vbx = qxn * inv, vby = qyn * inv;
            float vex = vbx + ax * dt, vey = vby + ay * dt;
            float speed = __builtin_sqrtf(vex * vex + vey * vey);
            float decrement = (float)(a.cfg.basal_friction * a.cfg.gravity) * dt;
            float factor = speed > 1.0e-15f ? fmaxf(0.0f, 1.0f - decrement * rcpf(speed)) : 0.0f;
            float vax = vex * factor, vay = vey * factor;
            qax = depth * vax;
            qay = depth * vay;
            wave_v = speed * factor;
            double dd = (double)depth;
            acc.diag[2] += dd * (double)(vex - vbx);
            acc.diag[3] += dd * (double)(vey - vby);
            acc.diag[6] += dd * (0.5 * (double)(vbx + vex) * a.cfg.external_ax +
                                 0.5 * (double)(vby + vey) * a.cfg.external_ay) * sub_dt;
            acc.diag[4] += (double)qax - dd * (double)vex;
            acc.diag[5] += (double)qay - dd * (double)vey;
            acc.diag[7] += dd * ((double)speed * speed - ((double)vax * vax + (double)vay * vay));
        }
        acc.diag[8] += (double)((depth > k.dry ? qxn : 0.0f) - qxi);
        acc.diag[9] += (double)((depth > k.dry ? qyn : 0.0f) - qyi);
        const int64_t i = (int64_t)gr * nx + gc;
        a.h[dst][i] = hq;
        a.qx[dst][i] = qax;
        a.qy[dst][i] = qay;
        wmax = npmax(wmax, (double)(wave_v + __builtin_sqrtf(k.Kg * depth)));
    }
    __syncthreads();
}

// Workgroups are dispatched round-robin over the 8 XCDs.  Renumber them so
// that the workgroups of one XCD own one contiguous band of tiles: halos
// then mostly hit that XCD's L2.
__device__ inline void xcd_tile_range(int tiles, int &first, int &last)
{
    const int G = gridDim.x, b = blockIdx.x;
    const int x = b % 8, j = b / 8;
    const int pb = x * (G / 8) + min(x, G % 8) + j;
    first = (int)((long long)pb * tiles / G);
    last = (int)((long long)(pb + 1) * tiles / G);
}

template <class P>
__device__ inline CellOut cell_update(const V2Args<P> &a, int src, int r, int c, double sub_dt, Acc<P> &acc);
template <>
__device__ inline CellOut cell_update<F64P>(const V2Args<F64P> &a, int src, int r, int c, double sub_dt, Acc<F64P> &acc)
{
    return cell64(a, src, r, c, sub_dt, acc);
}
template <>
__device__ inline CellOut cell_update<F32P>(const V2Args<F32P> &a, int src, int r, int c, double sub_dt, Acc<F32P> &acc)
{
    return cell32(a, src, r, c, sub_dt, acc);
}

// wave speed of an existing state (CFL of the first substep)
template <class P>
__device__ inline double cell_wave(const V2Args<P> &a, int64_t i)
{
    const double K = a.cfg.earth_pressure, g = a.cfg.gravity, dry = a.cfg.dry_tolerance;
    if (P::kQuant) {
        float d = P::depth(a.h[0][i]);
        float vx = 0.0f, vy = 0.0f;
        if (d > (float)dry) {
            float inv = rcpf(d);
            vx = (float)a.qx[0][i] * inv;
            vy = (float)a.qy[0][i] * inv;
        }
        return (double)(__builtin_sqrtf(vx * vx + vy * vy) + __builtin_sqrtf((float)(K * g) * d));
    } else {
        double d = (double)a.h[0][i];
        double vx = d > dry ? (double)a.qx[0][i] / d : 0.0;
        double vy = d > dry ? (double)a.qy[0][i] / d : 0.0;
        return sqrt(vx * vx + vy * vy) + sqrt((K * g) * npmax(d, 0.0));
    }
}

// ledger terms of one cell (warp_mobile_v2.py:41-54 with weight w, density rho)
template <class P>
__device__ inline void cell_summary(const V2Args<P> &a, int buf, int64_t i, double w, double (&s)[kSum])
{
    const double dry = a.cfg.dry_tolerance, rho = a.cfg.density;
    double depth = (double)P::depth(a.h[buf][i]);
    double qx = (double)a.qx[buf][i], qy = (double)a.qy[buf][i];
    double b = (double)a.b[i];
    s[0] += depth * w;
    s[1] += rho * qx * w;
    s[2] += rho * qy * w;
    if (depth > dry) {
        double vx = qx / depth, vy = qy / depth;
        double speed = sqrt(vx * vx + vy * vy);
        s[3] = npmax(s[3], speed);
        s[4] += 0.5 * rho * depth * speed * speed * w;
    }
    s[5] += rho * a.cfg.gravity * depth * b * w;
    s[6] += rho * 0.5 * a.cfg.earth_pressure * a.cfg.gravity * depth * depth * w;
    s[7] = npmin(s[7], depth);
}

__device__ inline int weight_class(int ny, int nx, int r, int c)
{
    int rc = r == 0 ? 0 : (r == ny - 1 ? 2 : 1);
    int cc = c == 0 ? 0 : (c == nx - 1 ? 2 : 1);
    return rc * 3 + cc;
}

// reduce per-workgroup partials in fixed order (workgroup 0 only)
template <class P>
__device__ void finalize_record(const V2Args<P> &a, double *smem, bool after_valid)
{
    const int nparts = gridDim.x;
    double v[kSum * 2 + kDiag];
    for (int k = 0; k < kSum * 2 + kDiag; k++) v[k] = 0.0;
    double mx[4] = {0.0, 0.0, 0.0, 0.0};          // max speed before/after; -min h before/after
    mx[2] = -1e300; mx[3] = -1e300;
    long long mq[2 * kClasses];
    for (int k = 0; k < 2 * kClasses; k++) mq[k] = 0;
    for (int p = threadIdx.x; p < nparts; p += blockDim.x) {
        const Partial &pp = a.partials[p];
        for (int k = 0; k < kSum; k++) {
            v[k] += pp.before[k];
            v[kSum + k] += pp.after[k];
        }
        for (int k = 0; k < kDiag; k++) v[2 * kSum + k] += pp.diag[k];
        mx[0] = npmax(mx[0], pp.before[3]);
        mx[1] = npmax(mx[1], pp.after[3]);
        mx[2] = npmax(mx[2], -pp.before[7]);
        mx[3] = npmax(mx[3], -pp.after[7]);
        for (int k = 0; k < kClasses; k++) {
            mq[k] += pp.mass_before[k];
            mq[kClasses + k] += pp.mass_after[k];
        }
    }
    block_reduce<double, kSum * 2 + kDiag, 0>(v, smem);
    block_reduce<double, 4, 1>(mx, smem);
    // exact integer sums: reduce through LDS as 64-bit integers
    long long *ls = reinterpret_cast<long long *>(smem);
    for (int k = 0; k < 2 * kClasses; k++) {
        long long x = mq[k];
        for (int off = kWave / 2; off > 0; off >>= 1) x += __shfl_down(x, off, kWave);
        if ((threadIdx.x & (kWave - 1)) == 0) ls[k * (blockDim.x / kWave) + threadIdx.x / kWave] = x;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        long long tot[2 * kClasses];
        for (int k = 0; k < 2 * kClasses; k++) {
            tot[k] = 0;
            for (int w = 0; w < (int)(blockDim.x / kWave); w++) tot[k] += ls[k * (blockDim.x / kWave) + w];
        }
        mv2_step_record *rec = a.rec;
        for (int k = 0; k < kSum; k++) {
            rec->summary_before[k] = v[k];
            rec->summary_after[k] = after_valid ? v[kSum + k] : v[k];
        }
        rec->summary_before[3] = mx[0];
        rec->summary_after[3] = after_valid ? mx[1] : mx[0];
        rec->summary_before[7] = -mx[2];
        rec->summary_after[7] = after_valid ? -mx[3] : -mx[2];
        for (int k = 0; k < kDiag; k++) rec->diag[k] = after_valid ? v[2 * kSum + k] : 0.0;
        if (P::kQuant) {
            // exact mobile volume: integer quanta per weight class times its weight
            const double area = a.dx * a.dy, f6 = a.dx * a.dy / 6.0;
            const int rep_r[3] = {0, 1, a.ny - 1}, rep_c[3] = {0, 1, a.nx - 1};
            double mb = 0.0, ma = 0.0, qb = 0.0, qa = 0.0;
            for (int k = 0; k < kClasses; k++) {
                double w = a.cfg.weight_mode ? vertex_weight(a.ny, a.nx, f6, rep_r[k / 3], rep_c[k % 3]) : area;
                mb += ((double)tot[k] * kQuantum) * w;
                ma += ((double)tot[kClasses + k] * kQuantum) * w;
                qb += (double)tot[k];
                qa += (double)tot[kClasses + k];
            }
            rec->summary_before[0] = mb;
            rec->summary_after[0] = after_valid ? ma : mb;
            rec->mass_quanta_before = qb;
            rec->mass_quanta_after = after_valid ? qa : qb;
        }
    }
    __syncthreads();
}

template <class P>
__device__ void summary_pass(const V2Args<P> &a, int buf, bool before, double *smem)
{
    double s[kSum];
    for (int k = 0; k < kSum; k++) s[k] = 0.0;
    s[7] = 1e300;
    long long mq[kClasses];
    for (int k = 0; k < kClasses; k++) mq[k] = 0;
    const int64_t n = (int64_t)a.ny * a.nx;
    const double area = a.dx * a.dy, f6 = a.dx * a.dy / 6.0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        int r = (int)(i / a.nx), c = (int)(i - (int64_t)r * a.nx);
        double w = cell_weight(a.cfg.weight_mode, a.ny, a.nx, area, f6, r, c);
        cell_summary<P>(a, buf, i, w, s);
        if (P::kQuant) mq[weight_class(a.ny, a.nx, r, c)] += (long long)a.h[buf][i];
    }
    double sums[6] = {s[0], s[1], s[2], s[4], s[5], s[6]};
    double mxs[2] = {s[3], -s[7]};
    block_reduce<double, 6, 0>(sums, smem);
    block_reduce<double, 2, 1>(mxs, smem);
    long long *ls = reinterpret_cast<long long *>(smem);
    long long tot[kClasses];
    for (int k = 0; k < kClasses; k++) {
        long long x = mq[k];
        for (int off = kWave / 2; off > 0; off >>= 1) x += __shfl_down(x, off, kWave);
        if ((threadIdx.x & (kWave - 1)) == 0) ls[k * (blockDim.x / kWave) + threadIdx.x / kWave] = x;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 0; k < kClasses; k++) {
            tot[k] = 0;
            for (int w = 0; w < (int)(blockDim.x / kWave); w++) tot[k] += ls[k * (blockDim.x / kWave) + w];
        }
        Partial &pp = a.partials[blockIdx.x];
        double *dst = before ? pp.before : pp.after;
        dst[0] = sums[0]; dst[1] = sums[1]; dst[2] = sums[2]; dst[3] = mxs[0];
        dst[4] = sums[3]; dst[5] = sums[4]; dst[6] = sums[5]; dst[7] = -mxs[1];
        for (int k = 0; k < kClasses; k++) (before ? pp.mass_before : pp.mass_after)[k] = tot[k];
    }
    __syncthreads();
}

// static diagnostic (warp_mobile_v2.py:56-85, gradient of H = b + h in the
// CPU form: free surface first, then the centred/one-sided difference)
template <class P>
__device__ void static_pass(const V2Args<P> &a, double *smem)
{
    double m[2] = {0.0, 0.0};
    const int ny = a.ny, nx = a.nx;
    const int64_t n = (int64_t)ny * nx;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        int r = (int)(i / nx), c = (int)(i - (int64_t)r * nx);
        int cl = c > 0 ? c - 1 : 0, cr = c + 1 < nx ? c + 1 : nx - 1;
        int rd = r > 0 ? r - 1 : 0, ru = r + 1 < ny ? r + 1 : ny - 1;
        double denx = (c > 0 && c + 1 < nx) ? 2.0 * a.dx : a.dx;
        double deny = (r > 0 && r + 1 < ny) ? 2.0 * a.dy : a.dy;
        auto H = [&](int rr, int cc) {
            int64_t j = (int64_t)rr * nx + cc;
            return (double)a.b[j] + (double)P::depth(a.h[0][j]);
        };
        double gx = (H(r, cr) - H(r, cl)) / denx;
        double gy = (H(ru, c) - H(rd, c)) / deny;
        m[0] = npmax(m[0], sqrt(gx * gx + gy * gy));
        m[1] = npmax(m[1], fabs((double)a.qx[0][i]));
        m[1] = npmax(m[1], fabs((double)a.qy[0][i]));
    }
    block_reduce<double, 2, 1>(m, smem);
    if (threadIdx.x == 0) {
        atomic_max_nonneg(&a.sync->static_grad, m[0]);
        atomic_max_nonneg(&a.sync->static_q, m[1]);
    }
}

template <class P>
__global__ void __launch_bounds__(kBlock) v2_step_kernel(V2Args<P> a)
{
    __shared__ double smem[64 * 4];
    unsigned epoch = 0;
    const int ny = a.ny, nx = a.nx;
    const int64_t n = (int64_t)ny * nx;
    const mv2_step_config &cfg = a.cfg;
    // ---- phase 0: ledger before, static diagnostic, first CFL maximum ----
    summary_pass<P>(a, 0, true, smem);
    if (cfg.static_exit) static_pass<P>(a, smem);
    {
        double wmax[1] = {0.0};
        for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
             i += (int64_t)gridDim.x * blockDim.x)
            wmax[0] = npmax(wmax[0], cell_wave<P>(a, i));
        block_reduce<double, 1, 1>(wmax, smem);
        if (threadIdx.x == 0) atomic_max_nonneg(&a.sync->wave_slot[0], wmax[0]);
    }
    grid_barrier(&a.sync->counter, epoch);
    const double ext_max = npmax(0.0, npmax(fabs(cfg.external_ax), fabs(cfg.external_ay)));
    if (cfg.static_exit && ext_max <= 1.0e-14) {
        double gmax = __longlong_as_double((long long)__hip_atomic_load(&a.sync->static_grad, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
        double qmax = __longlong_as_double((long long)__hip_atomic_load(&a.sync->static_q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
        if (qmax <= 1.0e-14 && gmax <= cfg.tan_start + 1.0e-14) {
            if (blockIdx.x == 0) {
                finalize_record<P>(a, smem, false);
                if (threadIdx.x == 0) {
                    a.rec->static_exit = 1;
                    a.rec->substeps = 0;
                    a.rec->remaining = cfg.dt;
                    a.rec->maximum_cfl = 0.0;
                }
            }
            return;
        }
    }
    // ---- substep loop (mobile_v2_reference.py:236-307) ----
    const double mind = a.dx <= a.dy ? a.dx : a.dy;
    double remaining = cfg.dt;
    double max_cfl = 0.0;
    int substeps = 0;
    int cfl_limited = 0;
    int numeric = 0;
    int src = 0;
    Acc<P> acc;
    for (int k = 0; k < kDiag; k++) acc.diag[k] = 0.0;
    while (remaining > 1.0e-14) {
        if (cfg.limit_mode == 0 && substeps >= cfg.maximum_substeps) { numeric = 2; break; }
        const int slot = substeps % 3;
        const double maxw = __longlong_as_double((long long)__hip_atomic_load(&a.sync->wave_slot[slot], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
        const double stable = maxw <= 1.0e-14 ? remaining : cfg.cfl * mind / maxw;
        const double sub_dt = npmin(remaining, stable);
        max_cfl = npmax(max_cfl, maxw * sub_dt / mind);
        if (sub_dt < remaining - 1.0e-14) cfl_limited = 1;
        if (blockIdx.x == 0 && threadIdx.x == 0)
            __hip_atomic_store(&a.sync->wave_slot[(substeps + 2) % 3], 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        double wmax[1] = {0.0};
        double hmin = 0.0;
        int ovf = 0;
        if constexpr (P::kQuant) {
            __shared__ Tile32 tile;
            const int tiles_x = (nx + kTW - 1) / kTW;
            const int tiles = tiles_x * ((ny + kTH - 1) / kTH);
            int first, last;
            xcd_tile_range(tiles, first, last);
            for (int t = first; t < last; t++)
                tile_substep32(a, tile, src, t, tiles_x, sub_dt, acc, wmax[0], hmin, ovf);
        } else {
            for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
                 i += (int64_t)gridDim.x * blockDim.x) {
                int r = (int)(i / nx), c = (int)(i - (int64_t)r * nx);
                CellOut o = cell_update<P>(a, src, r, c, sub_dt, acc);
                wmax[0] = npmax(wmax[0], o.wave);
                hmin = npmin(hmin, o.hmin);
                ovf |= o.overflow;
            }
        }
        block_reduce<double, 1, 1>(wmax, smem);
        double hm[1] = {hmin};
        block_reduce<double, 1, 2>(hm, smem);
        int anyovf = __syncthreads_or(ovf);
        if (threadIdx.x == 0) {
            atomic_max_nonneg(&a.sync->wave_slot[(substeps + 1) % 3], wmax[0]);
            const double tol = P::kQuant ? -1.0e-300 : -2.0e-13;
            if (hm[0] < tol) atomicOr(&a.sync->positivity, 1);
            if (anyovf) atomicOr(&a.sync->overflow, 1);
        }
        grid_barrier(&a.sync->counter, epoch);
        if (__hip_atomic_load(&a.sync->overflow, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) { numeric = 3; break; }
        if (__hip_atomic_load(&a.sync->positivity, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) { numeric = 1; break; }
        src ^= 1;
        remaining -= sub_dt;
        substeps += 1;
        if (cfg.limit_mode == 1 && substeps > cfg.maximum_substeps) { numeric = 2; break; }
    }
    // ---- copy back when the final state sits in scratch ----
    if (src == 1) {
        for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
             i += (int64_t)gridDim.x * blockDim.x) {
            a.h[0][i] = a.h[1][i];
            a.qx[0][i] = a.qx[1][i];
            a.qy[0][i] = a.qy[1][i];
        }
        grid_barrier(&a.sync->counter, epoch);
    }
    // ---- ledger after + per-workgroup diagnostics ----
    summary_pass<P>(a, 0, false, smem);
    block_reduce<double, kDiag, 0>(acc.diag, smem);
    if (threadIdx.x == 0)
        for (int k = 0; k < kDiag; k++) a.partials[blockIdx.x].diag[k] = acc.diag[k];
    grid_barrier(&a.sync->counter, epoch);
    if (blockIdx.x == 0) {
        finalize_record<P>(a, smem, true);
        if (threadIdx.x == 0) {
            a.rec->substeps = substeps;
            a.rec->cfl_limited = cfl_limited;
            a.rec->static_exit = 0;
            a.rec->maximum_cfl = max_cfl;
            a.rec->remaining = remaining;
            a.rec->numeric_error = numeric;
            a.rec->status = numeric ? MV2_ENUMERIC : MV2_OK;
        }
    }
}

__global__ void zero_words(unsigned *p, int n)
{
    for (int i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0u;
}

__global__ void zero_record(mv2_step_record *rec)
{
    unsigned *p = reinterpret_cast<unsigned *>(rec);
    for (int i = threadIdx.x; i < (int)(sizeof(mv2_step_record) / 4); i += blockDim.x) p[i] = 0u;
}

template <class P>
static int launch_step(mv2_context *ctx, const mv2_step_config *cfg, mv2_step_record *out)
{
    const int64_t n = (int64_t)ctx->ny * ctx->nx;
    int per_cu = 0;
    MV2_HIP(ctx, hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, v2_step_kernel<P>, kBlock, 0));
    if (per_cu < 1) return set_error(ctx, MV2_EDEVICE, "[MobileV2] kernel cannot be resident");
    int grid = per_cu * ctx->cu_count;
    int need = grid_for(n);
    if (P::kQuant) need = ((ctx->nx + kTW - 1) / kTW) * ((ctx->ny + kTH - 1) / kTH);
    if (grid > need) grid = need;
    if (grid < 1) grid = 1;
    if (int rc = ensure_scratch(ctx, ctx->partials, sizeof(Partial) * (size_t)grid)) return rc;
    V2Args<P> a;
    a.ny = ctx->ny;
    a.nx = ctx->nx;
    a.dx = ctx->dx;
    a.dy = ctx->dy;
    a.b = (const typename P::real *)ctx->b_eff;
    a.h[0] = (typename P::hst *)ctx->mobile;
    a.h[1] = (typename P::hst *)ctx->alt_h.ptr;
    a.qx[0] = (typename P::real *)ctx->qx;
    a.qx[1] = (typename P::real *)ctx->alt_qx.ptr;
    a.qy[0] = (typename P::real *)ctx->qy;
    a.qy[1] = (typename P::real *)ctx->alt_qy.ptr;
    a.cfg = *cfg;
    a.sync = (SyncBlock *)ctx->sync.ptr;
    a.partials = (Partial *)ctx->partials.ptr;
    a.rec = (mv2_step_record *)ctx->result.ptr;
    a.clamp_quarter = cfg->cfl <= 0.25 ? 1 : 0;
    hipLaunchKernelGGL(zero_words, dim3(1), dim3(64), 0, ctx->stream, (unsigned *)ctx->sync.ptr,
                       (int)(sizeof(SyncBlock) / 4));
    hipLaunchKernelGGL(zero_record, dim3(1), dim3(64), 0, ctx->stream, a.rec);
    void *args[] = {&a};
    MV2_HIP(ctx, hipEventRecord(ctx->ev_start, ctx->stream));
    MV2_HIP(ctx, hipLaunchCooperativeKernel((const void *)v2_step_kernel<P>, dim3(grid), dim3(kBlock),
                                            args, 0, ctx->stream));
    MV2_HIP(ctx, hipEventRecord(ctx->ev_stop, ctx->stream));
    MV2_HIP(ctx, hipMemcpyAsync(ctx->host_result.ptr, a.rec, sizeof(mv2_step_record),
                                hipMemcpyDeviceToHost, ctx->stream));
    MV2_HIP(ctx, hipStreamSynchronize(ctx->stream));
    *out = *(const mv2_step_record *)ctx->host_result.ptr;
    float ms = 0.0f;
    MV2_HIP(ctx, hipEventElapsedTime(&ms, ctx->ev_start, ctx->ev_stop));
    out->kernel_ms = ms;
    return MV2_OK;
}

template <class P>
static int launch_summary(mv2_context *ctx, const mv2_step_config *cfg, double out[8])
{
    mv2_step_config c = *cfg;
    c.static_exit = 1;
    c.external_ax = 0.0;
    c.external_ay = 0.0;
    c.tan_start = 1e300;     // forces the static exit right after the ledger
    c.maximum_substeps = 0;
    mv2_step_record rec;
    // A state with |q| > 1e-14 does not take the static exit; in that case the
    // zero-substep limit stops the loop before any update (limit_mode 0).
    c.limit_mode = 0;
    if (int rc = launch_step<P>(ctx, &c, &rec)) return rc;
    for (int k = 0; k < 8; k++) out[k] = rec.summary_before[k];
    return MV2_OK;
}

}  // namespace mv2

using namespace mv2;

extern "C" int mv2_step(mv2_context *ctx, const mv2_step_config *cfg, mv2_step_record *out)
{
    if (!ctx) return set_error(nullptr, MV2_EINVAL, "[MobileV2] null context");
    if (!ctx->bound) return set_error(ctx, MV2_ESTATE, "[MobileV2Production] bind state first");
    if (!cfg || !out) return set_error(ctx, MV2_EINVAL, "[MobileV2] null config/record");
    if (!(cfg->dt > 0.0) || !std::isfinite(cfg->dt))
        return set_error(ctx, MV2_EINVAL, "dt must be finite/positive");
    if (!(cfg->cfl > 0.0) || !(cfg->earth_pressure > 0.0) || !(cfg->gravity > 0.0) ||
        !(cfg->dry_tolerance > 0.0) || !(cfg->basal_friction > 0.0))
        return set_error(ctx, MV2_EINVAL, "MobileV2 configuration must be finite/positive");
    if (hipSetDevice(ctx->device) != hipSuccess) return set_error(ctx, MV2_EDEVICE, "hipSetDevice failed");
    int rc = ctx->precision == MV2_FP64 ? launch_step<F64P>(ctx, cfg, out)
                                        : launch_step<F32P>(ctx, cfg, out);
    if (rc) return rc;
    if (out->numeric_error == 1)
        return set_error(ctx, MV2_ENUMERIC, "MobileV2 positivity failure");
    if (out->numeric_error == 2)
        return set_error(ctx, MV2_ENUMERIC, "MobileV2 maximum_substeps exceeded");
    if (out->numeric_error == 3)
        return set_error(ctx, MV2_ENUMERIC, "MobileV2 mobile thickness exceeds the FP32 storage range");
    return MV2_OK;
}

extern "C" int mv2_summary(mv2_context *ctx, const mv2_step_config *cfg, double out[8])
{
    if (!ctx) return set_error(nullptr, MV2_EINVAL, "[MobileV2] null context");
    if (!ctx->bound) return set_error(ctx, MV2_ESTATE, "[MobileV2Production] bind state first");
    if (hipSetDevice(ctx->device) != hipSuccess) return set_error(ctx, MV2_EDEVICE, "hipSetDevice failed");
    return ctx->precision == MV2_FP64 ? launch_summary<F64P>(ctx, cfg, out)
                                      : launch_summary<F32P>(ctx, cfg, out);
}

extern "C" int mv2_static_diagnostic(mv2_context *ctx, double out[2])
{
    if (!ctx) return set_error(nullptr, MV2_EINVAL, "[MobileV2] null context");
    if (!ctx->bound) return set_error(ctx, MV2_ESTATE, "[MobileV2Production] bind state first");
    mv2_step_config c{};
    c.dt = 1.0;
    c.gravity = 9.81;
    c.earth_pressure = 0.45;
    c.cfl = 0.18;
    c.dry_tolerance = 1e-10;
    c.basal_friction = 0.55;
    c.tan_start = -1.0;      // never static: the loop is cut by the zero-substep limit
    c.density = 1.0;
    c.maximum_substeps = 0;
    c.limit_mode = 0;
    c.static_exit = 1;
    mv2_step_record rec;
    int rc = ctx->precision == MV2_FP64 ? launch_step<F64P>(ctx, &c, &rec)
                                        : launch_step<F32P>(ctx, &c, &rec);
    if (rc) return rc;
    SyncBlock sb;
    MV2_HIP(ctx, hipMemcpy(&sb, ctx->sync.ptr, sizeof(SyncBlock), hipMemcpyDeviceToHost));
    out[0] = __builtin_bit_cast(double, sb.static_grad);
    out[1] = __builtin_bit_cast(double, sb.static_q);
    return MV2_OK;
}
