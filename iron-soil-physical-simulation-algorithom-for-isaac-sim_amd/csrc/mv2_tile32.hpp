// FP32 Mobile V2 face and tile machinery shared by the single-terrain step
// (mv2_substep.hip) and the batched step (mv2_batch.hip).
//
// A TH x TW tile stages a (TH+2) x (TW+2) halo block of cells (depth, quanta,
// momentum, wet-masked velocity) in LDS, computes its TH x (TW+1) x-faces and
// (TH+1) x TW y-faces into LDS, then updates its cells.  Faces on a tile
// border are computed by both neighbouring tiles from bit-identical operands,
// so both sides move the same integer number of quanta.  The tile functions
// are templates over the tile shape and over the argument record, which must
// provide ny, nx, dx, dy, b, h[2], qx[2], qy[2], cfg and clamp_quarter.
#pragma once

#include "mv2_internal.hpp"

namespace mv2 {

constexpr int kDiag = 10;      // substep diagnostics accumulated over a step

// numpy's maximum / minimum on non-NaN operands, in the reference's order
__device__ inline double npmax(double a, double b) { return a >= b ? a : b; }
__device__ inline double npmin(double a, double b) { return a <= b ? a : b; }

struct Face32 { int h; float qxl, qyl, qxr, qyr; };

struct Consts32 {
    float K, invK, Kg, c, dry;
};

__device__ inline float rcpf(float x) { return __builtin_amdgcn_rcpf(x); }

template <int TH, int TW>
struct TileT {
    static constexpr int LH = TH + 2, LW = TW + 2;
    static constexpr int FX = TH * (TW + 1), FY = (TH + 1) * TW;
    float b[LH * LW];
    float h[LH * LW];
    int hq[LH * LW];
    float qx[LH * LW];
    float qy[LH * LW];
    float vx[LH * LW];
    float vy[LH * LW];
    int fxq[FX];
    float fxl[FX], fxr[FX], fxt[FX];
    int fyq[FY];
    float fyl[FY], fyr[FY], fyt[FY];
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

// Stage the TH x TW tile at (r0, c0) plus its one-cell halo into LDS.  The
// loop has a fixed trip count and is unrolled, so every load of the thread is
// in flight at once (one memory round trip for the whole tile).  The caller
// synchronises before the LDS is read.
template <int TH, int TW, class A>
__device__ inline void tile_stage(const A &a, TileT<TH, TW> &T, int src, int r0, int c0)
{
    using TT = TileT<TH, TW>;
    constexpr int LH = TT::LH, LW = TT::LW;
    constexpr int kStageIters = (LH * LW + kBlock - 1) / kBlock;
    const int ny = a.ny, nx = a.nx;
    const float Qf = (float)kQuantum;
    const float dry = (float)a.cfg.dry_tolerance;
    const float *gb = a.b;
    const int32_t *gh = (src ? a.h[1] : a.h[0]);
    const float *gqx = (src ? a.qx[1] : a.qx[0]);
    const float *gqy = (src ? a.qy[1] : a.qy[0]);
    float b[kStageIters], qx[kStageIters], qy[kStageIters];
    int hq[kStageIters];
#pragma unroll
    for (int j = 0; j < kStageIters; j++) {
        const int e = threadIdx.x + j * kBlock;
        const int lr = e / LW, lc = e - lr * LW;
        const int gr = r0 - 1 + lr, gc = c0 - 1 + lc;
        b[j] = 0.0f; qx[j] = 0.0f; qy[j] = 0.0f; hq[j] = 0;
        if (e < LH * LW && gr >= 0 && gr < ny && gc >= 0 && gc < nx) {
            const int i = gr * nx + gc;
            b[j] = gb[i];
            hq[j] = gh[i];
            qx[j] = gqx[i];
            qy[j] = gqy[i];
        }
    }
#pragma unroll
    for (int j = 0; j < kStageIters; j++) {
        const int e = threadIdx.x + j * kBlock;
        if (e >= LH * LW) continue;
        const float h = (float)hq[j] * Qf;
        float vx = 0.0f, vy = 0.0f;
        if (h > dry) {
            float inv = rcpf(h);
            vx = qx[j] * inv;
            vy = qy[j] * inv;
        }
        T.b[e] = b[j]; T.h[e] = h; T.hq[e] = hq[j]; T.qx[e] = qx[j]; T.qy[e] = qy[j]; T.vx[e] = vx; T.vy[e] = vy;
    }
}

// faces and cell update of the staged tile t (after a barrier)
template <int TH, int TW, class A>
__device__ inline void tile_update(const A &a, TileT<TH, TW> &T, int src, int r0, int c0, double sub_dt,
                                   double (&acc_diag)[10], double &wmax, double &hmin, int &ovf)
{
    using TT = TileT<TH, TW>;
    constexpr int LW = TT::LW, FX = TT::FX, FY = TT::FY;
    static_assert(TH * TW % kBlock == 0, "tile cells must split evenly over the block");
    const int ny = a.ny, nx = a.nx;
    const Consts32 k = consts32(a.cfg);
    const float Qf = (float)kQuantum;
    const float fxf = (float)(sub_dt / a.dx), fyf = (float)(sub_dt / a.dy);
    const int cq = a.clamp_quarter;
#pragma unroll
    for (int j = 0; j < (FX + FY + kBlock - 1) / kBlock; j++) {
        const int e = threadIdx.x + j * kBlock;
        if (e >= FX + FY) continue;
        if (e < FX) {
            const int lr = e / (TW + 1), kk = e - lr * (TW + 1);
            const int gr = r0 + lr, gcl = c0 - 1 + kk;
            if (gr >= ny || gcl < 0 || gcl + 1 >= nx) continue;
            const int L = (lr + 1) * LW + kk, R = L + 1;
            Face32 f = face32v<0>(T.b[L], T.b[R], T.hq[L], T.hq[R], T.h[L], T.h[R], T.vx[L], T.vy[L],
                                  T.vx[R], T.vy[R], k, fxf, cq);
            T.fxq[e] = f.h; T.fxl[e] = f.qxl; T.fxr[e] = f.qxr; T.fxt[e] = f.qyl;
        } else {
            const int ey = e - FX;
            const int kk = ey / TW, lc = ey - kk * TW;
            const int grl = r0 - 1 + kk, gc = c0 + lc;
            if (gc >= nx || grl < 0 || grl + 1 >= ny) continue;
            const int L = kk * LW + lc + 1, R = L + LW;
            Face32 f = face32v<1>(T.b[L], T.b[R], T.hq[L], T.hq[R], T.h[L], T.h[R], T.vx[L], T.vy[L],
                                  T.vx[R], T.vy[R], k, fyf, cq);
            T.fyq[ey] = f.h; T.fyl[ey] = f.qyl; T.fyr[ey] = f.qyr; T.fyt[ey] = f.qxl;
        }
    }
    __syncthreads();
    const int dst = src ^ 1;
    // per-thread sums of the substep diagnostics over this thread's cells, in
    // FP32 (two cells per thread), promoted to FP64 for the block reduction
    float dg[kDiag];
#pragma unroll
    for (int d = 0; d < kDiag; d++) dg[d] = 0.0f;
    long long hn_min = 0;
    float wv = 0.0f;
    const float dt = (float)sub_dt;
    const float ax = (float)a.cfg.external_ax, ay = (float)a.cfg.external_ay;
    const float decrement = (float)(a.cfg.basal_friction * a.cfg.gravity) * dt;
#pragma unroll
    for (int j = 0; j < TH * TW / kBlock; j++) {
        const int e = threadIdx.x + j * kBlock;
        const int ir = e / TW, ic = e - ir * TW;
        const int gr = r0 + ir, gc = c0 + ic;
        if (gr >= ny || gc >= nx) continue;
        const int L = (ir + 1) * LW + ic + 1;
        long long dq = 0;
        float dqx = 0.0f, dqy = 0.0f;
        if (gc + 1 < nx) {
            const int f = ir * (TW + 1) + ic + 1;
            dq -= T.fxq[f]; dqx = dqx - T.fxl[f]; dqy = dqy - T.fxt[f];
        }
        if (gc > 0) {
            const int f = ir * (TW + 1) + ic;
            dq += T.fxq[f]; dqx = dqx + T.fxr[f]; dqy = dqy + T.fxt[f];
        }
        if (gr + 1 < ny) {
            const int f = (ir + 1) * TW + ic;
            dq -= T.fyq[f]; dqx = dqx - T.fyt[f]; dqy = dqy - T.fyl[f];
        }
        if (gr > 0) {
            const int f = ir * TW + ic;
            dq += T.fyq[f]; dqx = dqx + T.fyt[f]; dqy = dqy + T.fyr[f];
        }
        const int hqi = T.hq[L];
        const float qxi = T.qx[L], qyi = T.qy[L];
        const long long hn = (long long)hqi + dq;
        ovf |= hn > 2147483647LL;
        hn_min = hn < hn_min ? hn : hn_min;
        const int hq = (int)(hn < 0 ? 0 : (hn > 2147483647LL ? 2147483647LL : hn));
        const float qxn = qxi + dqx, qyn = qyi + dqy;
        dg[0] += qxn - qxi;
        dg[1] += qyn - qyi;
        const float depth = (float)hq * Qf;
        float qax = 0.0f, qay = 0.0f, wave_v = 0.0f;
        const bool wet = depth > k.dry;
        if (wet) {
            const float inv = rcpf(depth);
            const float vbx = qxn * inv, vby = qyn * inv;
            const float vex = vbx + ax * dt, vey = vby + ay * dt;
            const float speed = __builtin_sqrtf(vex * vex + vey * vey);
            const float factor = speed > 1.0e-15f ? fmaxf(0.0f, 1.0f - decrement * rcpf(speed)) : 0.0f;
            const float vax = vex * factor, vay = vey * factor;
            qax = depth * vax;
            qay = depth * vay;
            wave_v = speed * factor;
            dg[2] += depth * (vex - vbx);
            dg[3] += depth * (vey - vby);
            dg[6] += depth * (0.5f * (vbx + vex) * ax + 0.5f * (vby + vey) * ay) * dt;
            dg[4] += qax - depth * vex;
            dg[5] += qay - depth * vey;
            dg[7] += depth * (speed * speed - (vax * vax + vay * vay));
        }
        dg[8] += (wet ? qxn : 0.0f) - qxi;
        dg[9] += (wet ? qyn : 0.0f) - qyi;
        const int i = gr * nx + gc;
        (dst ? a.h[1] : a.h[0])[i] = hq;
        (dst ? a.qx[1] : a.qx[0])[i] = qax;
        (dst ? a.qy[1] : a.qy[0])[i] = qay;
        wv = fmaxf(wv, wave_v + __builtin_sqrtf(k.Kg * depth));
    }
#pragma unroll
    for (int d = 0; d < kDiag; d++) acc_diag[d] += (double)dg[d];
    hmin = npmin(hmin, (double)hn_min * kQuantum);
    wmax = npmax(wmax, (double)wv);
    __syncthreads();
}


}  // namespace mv2
