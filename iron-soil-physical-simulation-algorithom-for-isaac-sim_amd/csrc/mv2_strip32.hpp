// FP32 Mobile V2 substep in wave strips, shared by the single-terrain step
// (mv2_substep.hip) and the batched step (mv2_batch.hip).
//
// One wave owns a band of SH rows of 62 columns.  Its 64 lanes hold 64
// consecutive columns: lanes 1..62 are the band's columns, lanes 0 and 63 the
// halo columns on either side.  Every lane keeps its column's SH + 2 rows
// (the band plus one halo row above and below) in registers, so:
//   - the y-faces of a column are computed by its own lane from consecutive
//     rows, each once;
//   - the x-face between lanes l and l + 1 is computed by lane l from its
//     right neighbour's values, moved across by a DPP wave shift, and handed
//     to lane l + 1 by the opposite shift;
//   - nothing goes through LDS and no barrier is needed.
// Faces on a band or strip border are computed by both neighbouring waves
// from bit-identical operands with identical instructions, so both sides
// move the same integer number of quanta and the total is exact.
//
// The argument record A must provide ny, nx, dx, dy, b, h[2], qx[2], qy[2],
// cfg and clamp_quarter.
#pragma once

#include "mv2_internal.hpp"

namespace mv2 {

constexpr int kDiag = 10;               // substep diagnostics accumulated over a step
constexpr int kStripW = kWave - 2;      // columns a wave updates

// numpy's maximum / minimum on non-NaN operands, in the reference's order
__device__ inline double npmax(double a, double b) { return a >= b ? a : b; }
__device__ inline double npmin(double a, double b) { return a <= b ? a : b; }

__device__ inline float rcpf(float x) { return __builtin_amdgcn_rcpf(x); }
// hardware square root (about 1 ulp); every evaluation of a given face uses it
__device__ inline float sqrtf_hw(float x) { return __builtin_amdgcn_sqrtf(x); }

// lane l receives lane l + 1's value (wave_shl:1); lane 63 receives 0
__device__ inline int from_next_lane(int v) { return __builtin_amdgcn_mov_dpp(v, 0x130, 0xf, 0xf, true); }
// lane l receives lane l - 1's value (wave_shr:1); lane 0 receives 0
__device__ inline int from_prev_lane(int v) { return __builtin_amdgcn_mov_dpp(v, 0x138, 0xf, 0xf, true); }
// the wave's index in its workgroup, as a scalar
__device__ inline int wave_index() { return __builtin_amdgcn_readfirstlane((int)(threadIdx.x / kWave)); }
// a wave-uniform value, held in a scalar register
__device__ inline float uniform(float v)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_readfirstlane(__builtin_bit_cast(int, v)));
}
__device__ inline float from_next_lane(float v)
{
    return __builtin_bit_cast(float, from_next_lane(__builtin_bit_cast(int, v)));
}
__device__ inline float from_prev_lane(float v)
{
    return __builtin_bit_cast(float, from_prev_lane(__builtin_bit_cast(int, v)));
}

// double-precision value of lane l + 1 / lane l - 1 (two 32-bit DPP moves)
__device__ inline double from_next_lane_d(double v)
{
    const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
    const int lo = from_next_lane((int)(unsigned)u), hi = from_next_lane((int)(unsigned)(u >> 32));
    return __builtin_bit_cast(double, ((unsigned long long)(unsigned)hi << 32) | (unsigned)lo);
}
__device__ inline double from_prev_lane_d(double v)
{
    const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
    const int lo = from_prev_lane((int)(unsigned)u), hi = from_prev_lane((int)(unsigned)(u >> 32));
    return __builtin_bit_cast(double, ((unsigned long long)(unsigned)hi << 32) | (unsigned)lo);
}

struct Face32 { int h; float qxl, qyl, qxr, qyr; };

// Element at a 32-bit byte offset from a wave-uniform base: the address is
// the scalar base plus the unsigned per-lane offset, one global access with
// no 64-bit address arithmetic per lane (fields hold fewer than 2^30 cells)
template <class T>
__device__ inline T &at_bytes(T *base, unsigned off)
{
    return *reinterpret_cast<T *>(reinterpret_cast<char *>(base) + off);
}
template <class T>
__device__ inline const T &at_bytes(const T *base, unsigned off)
{
    return *reinterpret_cast<const T *>(reinterpret_cast<const char *>(base) + off);
}

// FP32 constants of one substep, rounded once from the FP64 configuration and
// kept in scalar registers
struct Consts32 {
    float K, invK, Kg, c, dry;
    float fxf, fyf;                     // sub_dt / dx, sub_dt / dy
    float dt, ax, ay, decrement;        // sub_dt, external acceleration, mu g sub_dt
    int clamp_quarter;
};

__device__ inline Consts32 consts32(const mv2_step_config &cfg, double sub_dt, double dx, double dy,
                                    int clamp_quarter)
{
    Consts32 k;
    k.K = uniform((float)cfg.earth_pressure);
    k.invK = uniform((float)(1.0 / cfg.earth_pressure));
    k.Kg = uniform((float)(cfg.earth_pressure * cfg.gravity));
    k.c = uniform((float)(0.5 * cfg.earth_pressure * cfg.gravity));
    k.dry = uniform((float)cfg.dry_tolerance);
    k.fxf = uniform((float)(sub_dt / dx));
    k.fyf = uniform((float)(sub_dt / dy));
    k.dt = uniform((float)sub_dt);
    k.ax = uniform((float)cfg.external_ax);
    k.ay = uniform((float)cfg.external_ay);
    k.decrement = uniform((float)(cfg.basal_friction * cfg.gravity) * k.dt);
    k.clamp_quarter = clamp_quarter;
    return k;
}

// One side of a face: the cell's terrain b, hydrostatic head b + K h,
// pressure (K g / 2) h^2, a quarter of its quanta (outflow cap) and its
// wet-masked velocity.  Computed once per cell and shared by its four faces.
struct Side32 {
    float b, head, p, vx, vy;
    int q4;
};

__device__ inline Side32 side32(float b, int hq, float qx, float qy, const Consts32 &k)
{
    Side32 s;
    const float h = (float)hq * (float)kQuantum;
    s.b = b;
    s.head = fmaf(k.K, h, b);
    s.p = (k.c * h) * h;
    s.q4 = hq >> 2;
    s.vx = 0.0f;
    s.vy = 0.0f;
    if (h > k.dry) {
        const float inv = rcpf(h);
        s.vx = qx * inv;
        s.vy = qy * inv;
    }
    return s;
}

// the right neighbour's side, moved across from lane l + 1
__device__ inline Side32 side_from_next_lane(const Side32 &s)
{
    Side32 r;
    r.b = from_next_lane(s.b);
    r.head = from_next_lane(s.head);
    r.p = from_next_lane(s.p);
    r.vx = from_next_lane(s.vx);
    r.vy = from_next_lane(s.vy);
    r.q4 = from_next_lane(s.q4);
    return r;
}

// Face of mobile_v2_reference.py:391-449 in float: hydrostatic
// reconstruction at b_face = max(b_l, b_r), the reconstructed velocity is the
// cell velocity (zero where the reconstruction is dry), Rusanov flux with the
// larger of the two wave speeds, pressure correction p - ps on each side.  The
// mass flux is rounded once to whole quanta; returned momentum fluxes are
// multiplied by factor = dt / spacing.  Every product-sum is an explicit fused
// multiply-add, so each evaluation of a face gives the same bits.
template <int NORMAL>
__device__ inline Face32 face32v(const Side32 &L, const Side32 &R, const Consts32 &k, float factor, float hfactor,
                                 float qhalf)
{
    const float b_face = fmaxf(L.b, R.b);
    const float hs_l = fmaxf((L.head - b_face) * k.invK, 0.0f);
    const float hs_r = fmaxf((R.head - b_face) * k.invK, 0.0f);
    const bool wl = hs_l > k.dry, wr = hs_r > k.dry;
    const float vnl = wl ? (NORMAL == 0 ? L.vx : L.vy) : 0.0f, vtl = wl ? (NORMAL == 0 ? L.vy : L.vx) : 0.0f;
    const float vnr = wr ? (NORMAL == 0 ? R.vx : R.vy) : 0.0f, vtr = wr ? (NORMAL == 0 ? R.vy : R.vx) : 0.0f;
    const float qnl = hs_l * vnl, qtl = hs_l * vtl;
    const float qnr = hs_r * vnr, qtr = hs_r * vtr;
    const float psl = (k.c * hs_l) * hs_l, psr = (k.c * hs_r) * hs_r;
    const float fln = fmaf(qnl, vnl, psl), frn = fmaf(qnr, vnr, psr);
    const float flt = qnl * vtl, frt = qnr * vtr;
    const float speed = fmaxf(fabsf(vnl) + sqrtf_hw(k.Kg * hs_l), fabsf(vnr) + sqrtf_hw(k.Kg * hs_r));
    // twice the Rusanov fluxes: F = (f_l + f_r) - speed (u_r - u_l); the
    // halving is exact and is folded into the factors below
    const float Fh = fmaf(-speed, hs_r - hs_l, qnl + qnr);
    const float Fn = fmaf(-speed, qnr - qnl, fln + frn);
    const float Ft = fmaf(-speed, qtr - qtl, flt + frt);
    Face32 o;
    int q = (int)rintf(Fh * qhalf);
    // outflow of a side is capped at a quarter of its quanta (q > 0 leaves
    // the left cell, q < 0 the right one); both caps are >= 0, so applying
    // both is the same as applying the one of the flow's direction
    if (k.clamp_quarter) q = max(min(q, L.q4), -R.q4);
    o.h = q;
    const float nl = fmaf(0.5f, Fn, L.p - psl) * factor;
    const float nr = fmaf(0.5f, Fn, R.p - psr) * factor;
    const float tt = Ft * hfactor;
    if (NORMAL == 0) {
        o.qxl = nl; o.qyl = tt; o.qxr = nr; o.qyr = tt;
    } else {
        o.qxl = tt; o.qyl = nl; o.qxr = tt; o.qyr = nr;
    }
    return o;
}

// Squared gradient of H = b + h at one cell for the static test
// max |grad H| <= tan_start + 1e-14 (warp_mobile_v2.py:56-85: centred
// differences inside, one-sided at the terrain edge).  Taken with reciprocal
// multiplies (a few ulps from the quotients); a cell within a relative 1e-12
// of the bound is redone with the quotients, so the test decides exactly as
// with quotients everywhere.
struct StaticGrad {
    double denx, inv_denx, dy, inv_dy1, inv_dy2, band_lo, band_hi;
    __device__ StaticGrad(double tan_start, double dx, double dy_, bool inner_x)
    {
        const double bound = tan_start + 1.0e-14;
        band_lo = bound * bound * (1.0 - 1.0e-12);
        band_hi = bound * bound * (1.0 + 1.0e-12);
        denx = inner_x ? 2.0 * dx : dx;
        inv_denx = 1.0 / denx;
        dy = dy_;
        inv_dy1 = 1.0 / dy_;
        inv_dy2 = 1.0 / (2.0 * dy_);
    }
    __device__ double grad2(double hl, double hr, double hd, double hu, bool inner_y) const
    {
        const double gx = (hr - hl) * inv_denx, gy = (hu - hd) * (inner_y ? inv_dy2 : inv_dy1);
        double g2 = gx * gx + gy * gy;
        if (g2 >= band_lo && g2 <= band_hi) {
            const double ex = (hr - hl) / denx, ey = (hu - hd) / (inner_y ? 2.0 * dy : dy);
            g2 = ex * ex + ey * ey;
        }
        return g2;
    }
};

// the rows of one lane's column: band rows 1..SH, halo rows 0 and SH + 1
template <int SH>
struct StripRows {
    float b[SH + 2], qx[SH + 2], qy[SH + 2];
    int hq[SH + 2];
};

// Load the wave's band (first row r0, first column c0) with its halo.  Every
// load of the lane is issued before any is used (one memory round trip);
// cells outside the terrain read as zero and feed no face that is used.  The
// loads are unconditional (indices clamped into the terrain, the value then
// replaced by zero): a bounds branch around each row would split the loads
// into blocks, across which the compiler's waits can no longer count on
// loads returning in issue order and fall back to waiting for all of them.
template <int SH, class A>
__device__ inline void strip_load(const A &a, int src, int r0, int c0, StripRows<SH> &R)
{
    const int lane = threadIdx.x & (kWave - 1);
    const int ny = a.ny, nx = a.nx;
    const int gc = c0 - 1 + lane;
    const bool col_in = gc >= 0 && gc < nx;
    const int cc = min(max(gc, 0), nx - 1);
    const float *gb = a.b;
    const int32_t *gh = (src ? a.h[1] : a.h[0]);
    const float *gqx = (src ? a.qx[1] : a.qx[0]);
    const float *gqy = (src ? a.qy[1] : a.qy[0]);
#pragma unroll
    for (int j = 0; j < SH + 2; j++) {
        const int gr = r0 - 1 + j;
        const bool in = col_in && gr >= 0 && gr < ny;
        const unsigned o = 4u * (unsigned)(min(max(gr, 0), ny - 1) * nx + cc);
        const float bv = at_bytes(gb, o), qxv = at_bytes(gqx, o), qyv = at_bytes(gqy, o);
        const int hv = at_bytes(gh, o);
        R.b[j] = in ? bv : 0.0f;
        R.hq[j] = in ? hv : 0;
        R.qx[j] = in ? qxv : 0.0f;
        R.qy[j] = in ? qyv : 0.0f;
    }
}

// One substep of the wave's band into buffer src ^ 1.  Per lane: dg gathers
// the substep diagnostics of its cells, wv the CFL wave speed of the new
// state, hn_min the smallest new thickness (quanta) before clamping, ovf the
// int32 range failure.  DIAGF 1 gathers only the sums the production record
// reads (0, 1, 4, 5, 7); the others stay 0.  INTERIOR: the band touches no
// edge of the terrain (strip_interior), so every cell has all four faces and
// no bounds test is needed.
template <int SH>
__device__ inline bool strip_interior(int ny, int nx, int r0, int c0)
{
    return r0 > 0 && r0 + SH < ny && c0 > 0 && c0 + kStripW < nx;
}

// per-cell consumer of the new state that does nothing (the default)
struct NoCellSink {
    __device__ void operator()(int, int, int, float, float, float) const {}
};

// `sink(row, column, quanta, qx, qy, b)` sees every owned cell's new state
// as it is stored (a step's last substep forms the closing ledger there).
template <int SH, class A, int DIAGF = 0, bool INTERIOR = false, class Sink = NoCellSink>
__device__ inline void strip_update(const A &a, const StripRows<SH> &R, int src, int r0, int c0, double sub_dt,
                                    float (&dg)[kDiag], float &wv, long long &hn_min, int &ovf,
                                    Sink &&sink = NoCellSink{})
{
    const int lane = threadIdx.x & (kWave - 1);
    const int ny = a.ny, nx = a.nx;
    const int gc = c0 - 1 + lane;
    const bool owner = lane >= 1 && lane <= kStripW && (INTERIOR || gc < nx);
    const Consts32 k = consts32(a.cfg, sub_dt, a.dx, a.dy, a.clamp_quarter);
    const float dt = k.dt, decrement = k.decrement;
    const int dst = src ^ 1;
    int32_t *oh = (dst ? a.h[1] : a.h[0]);
    float *oqx = (dst ? a.qx[1] : a.qx[0]);
    float *oqy = (dst ? a.qy[1] : a.qy[0]);

    // per direction: dt / spacing, half of it, half of it in quanta per metre
    const float hfx = 0.5f * k.fxf, hfy = 0.5f * k.fyf;
    const float qfx = uniform(hfx * (float)kInvQuantum), qfy = uniform(hfy * (float)kInvQuantum);
    // The sides of the rows are formed as the loop reaches them (row j + 1
    // in step j), so the rows' loads are waited for one by one and the work
    // of the first rows overlaps the arrival of the later ones.
    Side32 cur = side32(R.b[1], R.hq[1], R.qx[1], R.qy[1], k);
    // y-face between halo row 0 and band row 1
    Face32 fdown = face32v<1>(side32(R.b[0], R.hq[0], R.qx[0], R.qy[0], k), cur, k, k.fyf, hfy, qfy);
#pragma unroll
    for (int j = 1; j <= SH; j++) {
        const int gr = r0 - 1 + j;
        const Side32 next = side32(R.b[j + 1], R.hq[j + 1], R.qx[j + 1], R.qy[j + 1], k);
        const Face32 fup = face32v<1>(cur, next, k, k.fyf, hfy, qfy);
        // x-face between this lane's column and the next
        const Face32 fx = face32v<0>(cur, side_from_next_lane(cur), k, k.fxf, hfx, qfx);
        // the x-face on the left of this lane's column, from the lane before
        const int lq = from_prev_lane(fx.h);
        const float lqx = from_prev_lane(fx.qxr), lqy = from_prev_lane(fx.qyr);
        if (owner && (INTERIOR || gr < ny)) {
            const bool has_r = INTERIOR || gc + 1 < nx, has_l = INTERIOR || gc > 0;
            const bool has_u = INTERIOR || gr + 1 < ny, has_d = INTERIOR || gr > 0;
            float dqx = 0.0f, dqy = 0.0f;
            if (has_r) {
                dqx = dqx - fx.qxl; dqy = dqy - fx.qyl;
            }
            if (has_l) {
                dqx = dqx + lqx; dqy = dqy + lqy;
            }
            if (has_u) {
                dqx = dqx - fup.qxl; dqy = dqy - fup.qyl;
            }
            if (has_d) {
                dqx = dqx + fdown.qxr; dqy = dqy + fdown.qyr;
            }
            const int hqi = R.hq[j];
            const float qxi = R.qx[j], qyi = R.qy[j];
            int hq;
            if (k.clamp_quarter) {
                // every face carries at most a quarter of a cell's quanta
                // (< 2^29), so the four fit int32; outflows total at most hqi,
                // so the new count lies in [0, 2^32) and is exact as uint32
                const int d32 = ((has_l ? lq : 0) - (has_r ? fx.h : 0)) + ((has_d ? fdown.h : 0) - (has_u ? fup.h : 0));
                const unsigned hn = (unsigned)hqi + (unsigned)d32;
                ovf |= hn > 2147483647u;
                hq = (int)min(hn, 2147483647u);
            } else {
                long long dq = 0;
                if (has_r) dq -= fx.h;
                if (has_l) dq += lq;
                if (has_u) dq -= fup.h;
                if (has_d) dq += fdown.h;
                const long long hn = (long long)hqi + dq;
                ovf |= hn > 2147483647LL;
                hn_min = hn < hn_min ? hn : hn_min;
                hq = (int)(hn < 0 ? 0 : (hn > 2147483647LL ? 2147483647LL : hn));
                // an empty barrier: converting the clamped value from 32 bits
                // is exact and far cheaper than the 64-bit conversion it would
                // otherwise fold into
                __asm__ volatile("" : "+v"(hq));
            }
            const float qxn = qxi + dqx, qyn = qyi + dqy;
            dg[0] += qxn - qxi;
            dg[1] += qyn - qyi;
            const float depth = (float)hq * (float)kQuantum;
            float qax = 0.0f, qay = 0.0f, wave_v = 0.0f;
            const bool wet = depth > k.dry;
            if (wet) {
                float ax = k.ax, ay = k.ay;
                if (a.ext_x) {                          // per-cell external acceleration
                    ax = at_bytes(a.ext_x, 4u * (unsigned)(gr * nx + gc));
                    ay = at_bytes(a.ext_y, 4u * (unsigned)(gr * nx + gc));
                }
                const float inv = rcpf(depth);
                const float vbx = qxn * inv, vby = qyn * inv;
                const float vex = vbx + ax * dt, vey = vby + ay * dt;
                const float speed = sqrtf_hw(vex * vex + vey * vey);
                const float factor = speed > 1.0e-15f ? fmaxf(0.0f, 1.0f - decrement * rcpf(speed)) : 0.0f;
                const float vax = vex * factor, vay = vey * factor;
                qax = depth * vax;
                qay = depth * vay;
                wave_v = speed * factor;
                if constexpr (DIAGF == 0) {
                    dg[2] += depth * (vex - vbx);
                    dg[3] += depth * (vey - vby);
                    dg[6] += depth * (0.5f * (vbx + vex) * ax + 0.5f * (vby + vey) * ay) * dt;
                }
                dg[4] += qax - depth * vex;
                dg[5] += qay - depth * vey;
                dg[7] += depth * (speed * speed - (vax * vax + vay * vay));
            }
            if constexpr (DIAGF == 0) {
                dg[8] += (wet ? qxn : 0.0f) - qxi;
                dg[9] += (wet ? qyn : 0.0f) - qyi;
            }
            const unsigned o = 4u * (unsigned)(gr * nx + gc);
            at_bytes(oh, o) = hq;
            at_bytes(oqx, o) = qax;
            at_bytes(oqy, o) = qay;
            wv = fmaxf(wv, wave_v + sqrtf_hw(k.Kg * depth));
            sink(gr, gc, hq, qax, qay, R.b[j]);
        }
        fdown = fup;
        cur = next;
    }
}

}  // namespace mv2
