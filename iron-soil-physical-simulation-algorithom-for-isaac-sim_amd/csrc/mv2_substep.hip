// Mobile V2 finite-volume step on the device (gfx950).
//
// A step_resident call is a short stream of ordinary launches:
//   reset     zero the loop state and the diagnostic partials
//   begin     ledger before, static diagnostic, CFL maximum of the input
//   substep k one launch per substep; the loop state lives on the device
//             (Control below), so launches past the end of the loop return
//             at once and the host never reads a per-substep value
//   end       copy back when the final state sits in scratch, ledger after
//             (FP32: done by the first substep launch that finds the loop
//             ended, so it costs no launch of its own)
//   fold      one wave per ledger / diagnostic field folds the workgroup
//             partials in fixed order into pinned host memory; the host
//             assembles the record after its one stream sync
// The host enqueues as many substep launches as the previous call needed
// (plus one) and reads the record once; only a call that needs more substeps
// than enqueued pays a second round trip.  A kernel boundary costs a few
// microseconds, a grid-wide barrier inside a persistent launch of ~1000
// workgroups costs tens, which is why the loop is not one cooperative kernel.
// The CFL maximum of the state produced by substep k is reduced while that
// state is written (wave shuffles + one atomicMax per workgroup) into a
// rotating slot that substep k+1 reads.
//
// Numerics follow pkg/experimental/mobile_v2_reference.py (CPU reference):
//   FP64 mode: every face and cell update repeats the reference's operation
//              order (mobile_v2_reference.py:236-307, :365-449) with
//              -ffp-contract=off, so fields match the CPU oracle bit for bit.
//   FP32 mode: float fluxes; mobile thickness in int32 quanta of 2^-24 m.
//              Each face's mass flux is rounded once to whole quanta and added
//              to one cell / removed from the other, so total mass is exact;
//              outflow through one face is capped at a quarter of the donor's
//              quanta (never active in exact arithmetic for CFL <= 0.25), so
//              thickness stays non-negative without a mass-creating clamp.
#include "mv2_internal.hpp"
#include "mv2_strip32.hpp"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>

namespace mv2 {

// Loop state of one step_resident call.  Substep launch k reads ctl[k & 1]
// and its workgroup 0 writes the successor state into ctl[(k + 1) & 1], so a
// launch never reads what it writes; the stream order between launches makes
// the write visible to launch k + 1.
struct Control {
    double remaining;          // simulated time still to integrate
    double max_cfl;
    double last_remaining;     // state before the most recent substep
    int substeps, src;         // substeps taken; buffer holding the current state
    int last_substeps, last_src;
    int done;                  // loop finished (normally or by an error)
    int numeric;               // 0 ok, 1 positivity, 2 substep limit, 3 int32 range
    int cfl_limited, static_exit;
    int rolled;                // 1 once a failed substep was rolled back
    int end_done;              // FP32 ledger after: 1 stored by a launch past the loop's end (state copied
                               // back), 2 stored by the last substep (the fold pass copies back)
    // FP32 speculative step: launch 0 took substep 0 with the CFL maximum
    // `pred` predicted by the previous step's end pass; launch 1 checks it
    // (restarted: it did not hold, and the step began again from its input)
    int spec, restarted;
    double pred;
};

struct SyncBlock {
    unsigned long long pad[2];
    int flags[2];                        // launch k ORs into [(k+1)&1]: bit 0 positivity, bit 1 range
    Control ctl[2];
};
// Steps alternate between two loop-state blocks (and two groups of spread
// sets) by the parity of the step count; the finalize pass of a step zeroes
// the other parity's block, so the next step needs no reset launch.
constexpr size_t kSyncParityBytes = 256;
static_assert(sizeof(SyncBlock) <= kSyncParityBytes, "two loop-state blocks fit below the relaxation's");

// spread-maximum sets of one step (kSpreadWords u64 each): the CFL maximum
// rotates over three sets (launch k reads k % 3, fills (k+1) % 3 and clears
// (k+2) % 3); two more hold the static diagnostic of the input state
constexpr int kWaveSets = 3;
constexpr int kSetStaticGrad = 3;
constexpr int kSetStaticQ = 4;
constexpr int kStepSpreadSets = 5;

// per-workgroup partial ledger of the begin / end passes
constexpr int kSum = 8;                  // summary entries
constexpr int kClasses = 9;              // weight classes of the exact FP32 mass ledger
// Partials are stored field-major (structure of arrays): field f of
// workgroup g at [f * groups + g], so the folding pass reads each field of
// consecutive workgroups with coalesced loads.
constexpr int kLedgerF = 2 * kSum;        // doubles per workgroup: before[8], after[8]
constexpr int kLedgerQ = 2 * kClasses;    // int64 per workgroup: quanta before[9], after[9]

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
    // per-cell external acceleration (state precision; null: the uniform
    // cfg.external_ax / external_ay) and max |external| for the static exit
    const typename P::real *ext_x, *ext_y;
    double ext_max;
    SyncBlock *sync;
    SyncBlock *sync_next;                // the other parity's block, zeroed by the finalize pass
    unsigned long long *spread_next;     // the other parity's spread sets, zeroed likewise
    double *ledger_f;                    // kLedgerF x groups of the begin / end passes
    long long *ledger_q;                 // kLedgerQ x groups
    int groups;                          // workgroups of the begin / end passes
    double *diag_part;                   // kDiag x (workgroups of the substep launches)
    int diag_groups;
    unsigned long long *spread;          // kStepSpreadSets spread-maximum sets
    int clamp_quarter;                   // FP32: cap outflow per face at h/4
    // FP32 speculation (spread sets outside the parity groups): the true CFL
    // maximum of the input from launch 0 (this parity's set; the other
    // parity's is zeroed by the fold pass), the predicted maximum this step
    // reads (zeroed by its fold pass) and the one its end pass writes
    unsigned long long *spec_true, *spec_true_next;
    unsigned long long *pred, *pred_next;
};


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
    const double ax = a.ext_x ? a.ext_x[i] : a.cfg.external_ax;
    const double ay = a.ext_y ? a.ext_y[i] : a.cfg.external_ay;
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

// FP32 tiles of the single-terrain step: kStripW columns by 4 x SH rows, one
// band of SH rows per wave (faces in mv2_strip32.hpp); SH is chosen per grid
// at launch (kSH when nothing else is known)
constexpr int kSH = 4;

// Workgroups are dispatched round-robin over the 8 XCDs.  Renumber them so
// that the workgroups of one XCD own one contiguous band of tiles: halos
// then mostly hit that XCD's L2.
__device__ inline void xcd_tile_range(int tiles, int &first, int &last)
{
    const int G = gridDim.x;
    const int pb = xcd_block();
    first = (int)((long long)pb * tiles / G);
    last = (int)((long long)(pb + 1) * tiles / G);
}


// ---------------------------------------------------------------------------
// Ledger passes.  Each thread takes kLedgerU cells per grid stride and loads
// all their fields (and the four neighbours the static diagnostic needs)
// before any arithmetic, so a pass over the grid costs about one memory round
// trip instead of one per cell.
// ---------------------------------------------------------------------------
constexpr int kLedgerU = 4;
constexpr int kLedgerGroupsMax = 512;      // FP64 ledger passes: at most this many workgroups

struct CellVals {
    double b, depth, qx, qy;
    long long quanta;
};

template <class P>
__device__ inline CellVals load_cell(const V2Args<P> &a, int buf, int i)
{
    CellVals v;
    v.b = (double)a.b[i];
    const typename P::hst hr = a.h[buf][i];
    v.depth = (double)P::depth(hr);
    v.quanta = P::kQuant ? (long long)hr : 0;
    v.qx = (double)a.qx[buf][i];
    v.qy = (double)a.qy[buf][i];
    return v;
}

// wave speed of an existing state (CFL of the first substep), in the
// arithmetic of the substep that follows
template <class P>
__device__ inline double cell_wave(const V2Args<P> &a, const CellVals &v)
{
    const double K = a.cfg.earth_pressure, g = a.cfg.gravity, dry = a.cfg.dry_tolerance;
    if (P::kQuant) {
        const float d = (float)v.depth;
        float vx = 0.0f, vy = 0.0f;
        if (d > (float)dry) {
            float inv = rcpf(d);
            vx = (float)v.qx * inv;
            vy = (float)v.qy * inv;
        }
        return (double)(sqrtf_hw(vx * vx + vy * vy) + sqrtf_hw((float)(K * g) * d));
    } else {
        const double d = v.depth;
        double vx = d > dry ? v.qx / d : 0.0;
        double vy = d > dry ? v.qy / d : 0.0;
        return sqrt(vx * vx + vy * vy) + sqrt((K * g) * npmax(d, 0.0));
    }
}

// ledger terms of one cell (warp_mobile_v2.py:41-54 with weight w, density rho)
template <class P>
__device__ inline void cell_summary(const V2Args<P> &a, const CellVals &v, double w, double (&s)[kSum])
{
    const double dry = a.cfg.dry_tolerance, rho = a.cfg.density;
    const double depth = v.depth, qx = v.qx, qy = v.qy, b = v.b;
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

// per-thread accumulators of one ledger pass
struct LedgerAcc {
    double s[kSum];
    long long mq[kClasses];
    __device__ void init()
    {
        for (int k = 0; k < kSum; k++) s[k] = 0.0;
        s[7] = 1e300;
        for (int k = 0; k < kClasses; k++) mq[k] = 0;
    }
    template <class P>
    __device__ void add(const V2Args<P> &a, const CellVals &v, int i)
    {
        const int r = i / a.nx, c = i - r * a.nx;
        const double area = a.dx * a.dy, f6 = a.dx * a.dy / 6.0;
        cell_summary<P>(a, v, cell_weight(a.cfg.weight_mode, a.ny, a.nx, area, f6, r, c), s);
        if (P::kQuant) {
            const int cls = weight_class(a.ny, a.nx, r, c);
#pragma unroll
            for (int q = 0; q < kClasses; q++) mq[q] += cls == q ? v.quanta : 0;   // no dynamic indexing
        }
    }
};

// reduce the pass over the block and store the workgroup's partial ledger
template <class P>
__device__ inline void store_ledger(const V2Args<P> &a, LedgerAcc &L, bool before, double *smem)
{
    double sums[6] = {L.s[0], L.s[1], L.s[2], L.s[4], L.s[5], L.s[6]};
    double mxs[2] = {L.s[3], -L.s[7]};
    block_reduce<double, 6, 0>(sums, smem);
    block_reduce<double, 2, 1>(mxs, smem);
    if (P::kQuant) block_reduce<long long, kClasses, 0>(L.mq, reinterpret_cast<long long *>(smem));
    if (threadIdx.x == 0) {
        const int G = a.groups, g = blockIdx.x;
        double *dst = a.ledger_f + (before ? 0 : kSum) * G + g;
        const double vals[kSum] = {sums[0], sums[1], sums[2], mxs[0], sums[3], sums[4], sums[5], -mxs[1]};
        for (int k = 0; k < kSum; k++) dst[k * G] = vals[k];
        long long *q = a.ledger_q + (before ? 0 : kClasses) * G + g;
        for (int k = 0; k < kClasses; k++) q[k * G] = L.mq[k];
    }
}

// Loop state that launch k acts on.  A failure flag raised by the launch that
// wrote ctl[k & 1] rolls that substep back (the state before it is kept and
// the loop ends with the error); flags of older launches were already folded
// in by the launch after them (rolled == 1).  Both words were written by
// earlier launches of the stream, so plain loads see them.
// The record and flags of launch k were written by earlier launches (launch
// k writes only the other slot): scalar reads (load_launch_record).
// A failure flagged by the previous launch (bit 0 positivity, bit 1 int32
// range) rolls the loop state back to before the failing substep and ends
// the loop.
__device__ inline void apply_failure_flags(Control &C, int f)
{
    if (f && !C.rolled) {
        C.numeric = (f & 2) ? 3 : 1;
        C.src = C.last_src;
        C.remaining = C.last_remaining;
        C.substeps = C.last_substeps;
        C.done = 1;
        C.rolled = 1;
        C.end_done = 0;    // a closing ledger of the failed substep is void
    }
}

__device__ inline Control resolve_control(const SyncBlock *s, int k)
{
    Control C = load_launch_record(&s->ctl[k & 1]);
    apply_failure_flags(C, load_launch_record(&s->flags[k & 1]));
    return C;
}

// the loop state of a step's first substep: the whole frame dt still to go
__device__ inline Control fresh_control(double dt)
{
    Control C;
    __builtin_memset(&C, 0, sizeof(C));
    C.remaining = dt;
    return C;
}

// zero one parity's loop state and spread sets (first step, or after an
// error), the true-maximum set of its launch 0 and the prediction set its end
// pass fills
__global__ void v2_reset_kernel(SyncBlock *s, unsigned long long *spread, unsigned long long *spec_true,
                                unsigned long long *pred_next)
{
    unsigned *ps = reinterpret_cast<unsigned *>(s);
    for (int i = threadIdx.x; i < (int)(sizeof(SyncBlock) / 4); i += blockDim.x) ps[i] = 0u;
    for (int i = threadIdx.x; i < kStepSpreadSets * kSpread; i += blockDim.x) spread[(int64_t)i * kSpreadStride] = 0ull;
    for (int i = threadIdx.x; i < kSpread; i += blockDim.x) {
        spec_true[(int64_t)i * kSpreadStride] = 0ull;
        pred_next[(int64_t)i * kSpreadStride] = 0ull;
    }
}

// One pass over the input state: ledger before, static diagnostic
// (warp_mobile_v2.py:56-85: gradient of H = b + h, free surface first, then the
// centred / one-sided difference) and the CFL maximum of the first substep.
template <class P>
__global__ void __launch_bounds__(kBlock) v2_begin_kernel(V2Args<P> a)
{
    __shared__ double smem[64 * 4];
    const int ny = a.ny, nx = a.nx, n = ny * nx;
    const int stride = gridDim.x * blockDim.x;
    const bool stat = a.cfg.static_exit != 0;
    LedgerAcc L;
    L.init();
    double m[3] = {0.0, 0.0, 0.0};              // max |grad H|, max |q|, max wave speed
    for (int base = blockIdx.x * blockDim.x + threadIdx.x; base < n; base += stride * kLedgerU) {
        CellVals v[kLedgerU];
        double hn[kLedgerU][4];                 // H at left, right, down, up (clamped)
#pragma unroll
        for (int u = 0; u < kLedgerU; u++) {
            const int i = base + u * stride;
            if (i >= n) continue;
            v[u] = load_cell<P>(a, 0, i);
            if (stat) {
                const int r = i / nx, c = i - r * nx;
                const int nb[4] = {c > 0 ? i - 1 : i, c + 1 < nx ? i + 1 : i, r > 0 ? i - nx : i,
                                   r + 1 < ny ? i + nx : i};
#pragma unroll
                for (int q = 0; q < 4; q++) hn[u][q] = (double)a.b[nb[q]] + (double)P::depth(a.h[0][nb[q]]);
            }
        }
#pragma unroll
        for (int u = 0; u < kLedgerU; u++) {
            const int i = base + u * stride;
            if (i >= n) continue;
            L.add<P>(a, v[u], i);
            if (stat) {
                const int r = i / nx, c = i - r * nx;
                const double denx = (c > 0 && c + 1 < nx) ? 2.0 * a.dx : a.dx;
                const double deny = (r > 0 && r + 1 < ny) ? 2.0 * a.dy : a.dy;
                const double gx = (hn[u][1] - hn[u][0]) / denx;
                const double gy = (hn[u][3] - hn[u][2]) / deny;
                m[0] = npmax(m[0], sqrt(gx * gx + gy * gy));
                m[1] = npmax(m[1], fabs(v[u].qx));
                m[1] = npmax(m[1], fabs(v[u].qy));
            }
            m[2] = npmax(m[2], cell_wave<P>(a, v[u]));
        }
    }
    store_ledger<P>(a, L, true, smem);
    block_reduce<double, 3, 1>(m, smem);
    if (threadIdx.x == 0) {
        if (stat) {
            spread_max(a.spread + kSetStaticGrad * kSpreadWords, m[0]);
            spread_max(a.spread + kSetStaticQ * kSpreadWords, m[1]);
        }
        spread_max(a.spread, m[2]);
        if (blockIdx.x == 0) a.sync->ctl[0].remaining = a.cfg.dt;
    }
}

// ---------------------------------------------------------------------------
// FP32 ledger passes in the strip layout of the substep (one tile of the
// substep grid per workgroup, one band of SH rows per wave): x-neighbours
// by DPP, y-neighbours from the rows in registers, no integer division.  Per
// cell the ledger terms of warp_mobile_v2.py:41-54 are formed in float and
// summed per lane in float over the band, then in double; the mobile volume
// itself comes exactly from the quanta per weight class.
// ---------------------------------------------------------------------------

struct Ledger32 {
    float qx, qy, ke, pot, prs, vmax, hmin;   // sums weighted by the cell weight; max speed, min depth
    // quanta of the band's cells by row class: first row, inner rows, last row
    // (three scalars, not an array: a select between array slots would be
    // turned into a dynamically indexed private array)
    long long q_first, q_inner, q_last;
    __device__ void init()
    {
        qx = qy = ke = pot = prs = vmax = 0.0f;
        hmin = 3.0e38f;
        q_first = q_inner = q_last = 0;
    }
    // One cell at row r (uniform over the wave), weight w, terrain b.
    // Returns the cell's CFL wave speed in the arithmetic of the substep:
    // |v| + sqrt(K g h), the velocity from the hardware reciprocal and zero
    // where the cell is dry.  Every pass that forms a CFL maximum from a
    // stored state calls this, so a prediction and the true maximum it is
    // checked against are the same bits.
    __device__ float add(int r, int ny, int hq, float qxv, float qyv, float b, float w, float dry, float Kg)
    {
        const float d = (float)hq * (float)kQuantum;
        const long long u = (unsigned)hq;
        // selects, not branches: this runs inside the substep's row loop,
        // where a join would make the compiler wait for every row in flight
        q_first += r == 0 ? u : 0;
        q_last += r == ny - 1 ? u : 0;
        q_inner += (r != 0 && r != ny - 1) ? u : 0;
        qx += qxv * w;
        qy += qyv * w;
        float sp = 0.0f;
        if (d > dry) {
            const float inv = rcpf(d);
            const float vx = qxv * inv, vy = qyv * inv;
            const float s2 = vx * vx + vy * vy;
            sp = sqrtf_hw(s2);
            vmax = fmaxf(vmax, sp);
            ke += d * s2 * w;
        }
        pot += d * b * w;
        prs += d * d * w;
        hmin = fminf(hmin, d);
        return sp + sqrtf_hw(Kg * d);
    }
};

// weight of cell (r, c): the cell area, or the Triangle-A-C vertex weight
// (bulk_state_authority.py:624-634) as a multiple of area / 6
__device__ inline float cell_weight32(int mode, int ny, int nx, float area, float f6, int r, int c)
{
    if (!mode) return area;
    const int n = (r < ny - 1 && c < nx - 1 ? 2 : 0) + (r < ny - 1 && c > 0 ? 1 : 0) + (r > 0 && c > 0 ? 2 : 0) +
                  (r > 0 && c < nx - 1 ? 1 : 0);
    return (float)n * f6;
}

// The weights of one column (a lane's) for the three row classes: a row's
// weight then depends only on whether it is the first, the last or an inner
// row, a uniform choice per row.
struct ColWeights {
    float first, inner, last;
    __device__ ColWeights(int mode, int ny, int nx, float area, float f6, int c)
        : first(cell_weight32(mode, ny, nx, area, f6, 0, c)),
          inner(cell_weight32(mode, ny, nx, area, f6, ny > 2 ? 1 : 0, c)),
          last(cell_weight32(mode, ny, nx, area, f6, ny - 1, c)) {}
    // bit selects on the values: a select between the members themselves
    // becomes a select of their addresses and puts the struct in scratch
    __device__ float at(int r, int ny) const
    {
        const unsigned m0 = r == 0 ? ~0u : 0u, ml = r == ny - 1 ? ~0u : 0u;
        unsigned w = (__float_as_uint(first) & m0) | (__float_as_uint(inner) & ~m0);
        w = (__float_as_uint(last) & ml) | (w & ~ml);
        return __uint_as_float(w);
    }
};

// 64-bit value of lane `src` (uniform) of the wave
__device__ inline long long readlane64(long long v, int src)
{
    const unsigned long long u = (unsigned long long)v;
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)u, src);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(u >> 32), src);
    return (long long)(((unsigned long long)hi << 32) | lo);
}

// The wave's ledger as partial gw (before or after) of a.groups: fixed-order
// DPP trees over the lanes (sums and extrema in double, the quanta of each
// row class as integers), no barrier.  Weight class = row class (first,
// inner, last row) x column class (first, inner, last column); the first and
// last columns are at most one lane each and are read back from it.
__device__ inline void store_ledger32(const V2Args<F32P> &a, const Ledger32 &L, int gc, bool owner, bool before)
{
    const double rho = a.cfg.density, g = a.cfg.gravity, K = a.cfg.earth_pressure;
    // float DPP trees (one instruction per step; the lanes' sums are float
    // already), widened per wave; the fold sums the waves in double
    const double f0 = (double)wave_sum(L.qx), f1 = (double)wave_sum(L.qy);
    const double f2 = (double)wave_max(L.vmax);
    const double f3 = (double)wave_sum(L.ke), f4 = (double)wave_sum(L.pot), f5 = (double)wave_sum(L.prs);
    const double f6 = (double)wave_min(L.hmin);
    const unsigned long long first_col = __ballot(owner && gc == 0);
    const unsigned long long last_col = __ballot(owner && gc == a.nx - 1 && gc != 0);
    const int lf = first_col ? __ffsll((long long)first_col) - 1 : 0;
    const int ll = last_col ? __ffsll((long long)last_col) - 1 : 0;
    long long mq[kClasses];
#pragma unroll
    for (int r = 0; r < 3; r++) {
        const long long v = owner ? (r == 0 ? L.q_first : (r == 1 ? L.q_inner : L.q_last)) : 0;
        const long long tot = wave_sum(v);
        const long long e0 = first_col ? readlane64(v, lf) : 0;
        const long long e2 = last_col ? readlane64(v, ll) : 0;
        mq[3 * r] = e0;
        mq[3 * r + 1] = tot - e0 - e2;
        mq[3 * r + 2] = e2;
    }
    if ((threadIdx.x & (kWave - 1)) == 0) {
        const int G = a.groups, gw = blockIdx.x * (kBlock / kWave) + wave_index();
        double *dst = a.ledger_f + (before ? 0 : kSum) * G + gw;
        const double vals[kSum] = {0.0, rho * f0, rho * f1, f2, 0.5 * rho * f3, rho * g * f4,
                                   rho * 0.5 * K * g * f5, f6};
#pragma unroll
        for (int q = 0; q < kSum; q++) dst[q * G] = vals[q];
        long long *qd = a.ledger_q + (before ? 0 : kClasses) * G + gw;
#pragma unroll
        for (int q = 0; q < kClasses; q++) qd[q * G] = mq[q];
    }
}

// Substep k of the loop of mobile_v2_reference.py:236-307: decide from the
// loop state whether the loop has ended (then return at once), otherwise take
// sub_dt from the CFL maximum of the current state, update every cell into the
// other buffer and reduce the next CFL maximum, positivity and range flags and
// the diagnostics of the substep.
//
// A working launch k always finds the current state in buffer k & 1 (every
// earlier launch worked and flipped it).  An EAGER launch stages its tile
// from that buffer before the loop state and the CFL maximum arrive, so all
// three loads share one memory round trip; launches expected to find the loop
// ended (EAGER = 0) read the loop state first and skip the staging.  EAGER is
// a template parameter: a run-time branch around the staging would join two
// paths with different loads in flight, and the compiler's wait at the join
// would hold the eager path's first use of the CFL maximum until its whole
// tile had arrived.
constexpr int kSubFold = 3 + kDiag;   // values folded per workgroup at the end of a substep

template <int SH>
__device__ inline void end32_body(const V2Args<F32P> &a, int fsrc);

// Per-cell sink of a step's last substep: the closing ledger of the new state
// and the CFL maximum the next step's first launch will use as its prediction,
// each cell's terms formed exactly as the end pass forms them from the stored
// state (end32_body), so the partials are the same bits.
struct LedgerSink {
    Ledger32 L;
    ColWeights cw;
    float pred;          // a float maximum: the same value as the end pass's double one
    int ny;
    float dry, Kg;
    __device__ void operator()(int r, int, int hq, float qxv, float qyv, float b)
    {
        pred = fmaxf(pred, L.add(r, ny, hq, qxv, qyv, b, cw.at(r, ny), dry, Kg));
    }
};

// The loop state after a substep of sub_dt taken from state C with CFL
// maximum maxw; the loop's own exits in the reference's order
// (mobile_v2_reference.py:236-307).  Uniform: every wave forms it.
__device__ inline Control next_control(const Control &C, const mv2_step_config &cfg, double maxw, double sub_dt,
                                       double mind)
{
    Control nc = C;
    nc.last_remaining = C.remaining;
    nc.last_substeps = C.substeps;
    nc.last_src = C.src;
    nc.max_cfl = npmax(C.max_cfl, maxw * sub_dt / mind);
    if (sub_dt < C.remaining - 1.0e-14) nc.cfl_limited = 1;
    nc.remaining = C.remaining - sub_dt;
    nc.substeps = C.substeps + 1;
    nc.src = C.src ^ 1;
    if (cfg.limit_mode == 1 && nc.substeps > cfg.maximum_substeps) {
        nc.numeric = 2;
        nc.done = 1;
    } else if (!(nc.remaining > 1.0e-14)) {
        nc.done = 1;
    } else if (cfg.limit_mode == 0 && nc.substeps >= cfg.maximum_substeps) {
        nc.numeric = 2;
        nc.done = 1;
    }
    return nc;
}

// One FP32 substep of the wave's band, staged from buffer C.src: dt from the
// CFL maximum maxw, the successor loop state (written by workgroup 0;
// spec_pred >= 0 marks it speculative, with that predicted maximum, for
// launch 1 to check), the update into the other buffer and the wave-level
// epilogue.  dprev is this lane's diagnostic slot before the substep (0 for
// a step's first).  A substep that ends the loop and is not speculative also
// stores the closing ledger and the next step's CFL prediction from the new
// state in its registers (end_done = 2); the fold pass then copies the state
// back when it sits in the second buffer.  A failure flagged by this substep
// rolls the loop back and voids that ledger (apply_failure_flags).
template <int SH, int DIAGF, bool CAN_CLOSE = true>
__device__ inline void substep32_work(const V2Args<F32P> &a, const StripRows<SH> &rows, const Control &C,
                                      double maxw, int k, int r0, int c0, double dprev, double spec_pred)
{
    const mv2_step_config &cfg = a.cfg;
    SyncBlock *s = a.sync;
    const int lane = threadIdx.x & (kWave - 1);
    const int gw = blockIdx.x * (kBlock / kWave) + wave_index();
    const double mind = a.dx <= a.dy ? a.dx : a.dy;
    const double stable = maxw <= 1.0e-14 ? C.remaining : cfg.cfl * mind / maxw;
    const double sub_dt = npmin(C.remaining, stable);
    Control nc = next_control(C, cfg, maxw, sub_dt, mind);
    // CAN_CLOSE false: a speculative launch 0, whose substep never closes
    // the step (the close path and its registers are left out)
    const bool last = CAN_CLOSE && spec_pred < 0.0 && nc.done;
    if (last) nc.end_done = 2;
    if (spec_pred >= 0.0) {
        nc.spec = 1;
        nc.pred = spec_pred;
    }
    if (blockIdx.x == 0) spread_clear(a.spread + ((k + 2) % kWaveSets) * kSpreadWords);
    if (blockIdx.x == 0 && threadIdx.x == 0) s->ctl[(k + 1) & 1] = nc;
    float dg[kDiag];
#pragma unroll
    for (int d = 0; d < kDiag; d++) dg[d] = 0.0f;
    float wv = 0.0f;
    long long hn_min = 0;
    int ovf = 0;
    const bool interior = strip_interior<SH>(a.ny, a.nx, r0, c0);
    if (last) {
        const int gcl = c0 - 1 + lane;
        LedgerSink sink{Ledger32{}, ColWeights(cfg.weight_mode, a.ny, a.nx, uniform((float)(a.dx * a.dy)),
                                               uniform((float)(a.dx * a.dy / 6.0)), gcl),
                        0.0f, a.ny, uniform((float)cfg.dry_tolerance),
                        uniform((float)(cfg.earth_pressure * cfg.gravity))};
        sink.L.init();
        if (interior)
            strip_update<SH, V2Args<F32P>, DIAGF, true>(a, rows, C.src, r0, c0, sub_dt, dg, wv, hn_min, ovf, sink);
        else
            strip_update<SH, V2Args<F32P>, DIAGF, false>(a, rows, C.src, r0, c0, sub_dt, dg, wv, hn_min, ovf, sink);
        const int gc = c0 - 1 + lane;
        store_ledger32(a, sink.L, gc, lane >= 1 && lane <= kStripW && gc < a.nx, false);
        const double pm = wave_max((double)sink.pred);
        if (lane == 0) atomic_max_nonneg(&a.pred_next[(gw % kSpread) * kSpreadStride], pm);
    } else if (interior) {
        strip_update<SH, V2Args<F32P>, DIAGF, true>(a, rows, C.src, r0, c0, sub_dt, dg, wv, hn_min, ovf);
    } else {
        strip_update<SH, V2Args<F32P>, DIAGF, false>(a, rows, C.src, r0, c0, sub_dt, dg, wv, hn_min, ovf);
    }
    // Wave-level epilogue, no barrier: the wave's wave-speed maximum goes to a
    // spread slot of its own, a failure (negative thickness before clamping,
    // int32 range) to the flags, and its diagnostic sums (a fixed-order DPP
    // tree) to its partials, which the fold pass sums in fixed order.
    const float wmx = wave_max(wv);
    const bool neg = __ballot(hn_min < 0) != 0ull;
    const bool rng = __ballot(ovf != 0) != 0ull;
    float ds[kDiag];
#pragma unroll
    for (int d = 0; d < kDiag; d++) {
        const bool kept = DIAGF == 0 || d == 0 || d == 1 || d == 4 || d == 5 || d == 7;
        ds[d] = kept ? wave_sum(dg[d]) : 0.0f;
    }
    if (lane == 0) {
        atomic_max_nonneg(&a.spread[((k + 1) % kWaveSets) * kSpreadWords + (gw % kSpread) * kSpreadStride],
                          (double)wmx);
        const int f = (neg ? 1 : 0) | (rng ? 2 : 0);
        if (f) atomicOr(&s->flags[(k + 1) & 1], f);
    }
    if (lane < kDiag) {
        float v = ds[0];
#pragma unroll
        for (int d = 1; d < kDiag; d++) v = lane == d ? ds[d] : v;
        const bool kept = DIAGF == 0 || lane == 0 || lane == 1 || lane == 4 || lane == 5 || lane == 7;
        a.diag_part[(int64_t)lane * a.diag_groups + gw] = kept ? dprev + (double)v : 0.0;
    }
}

// CLOSE: the launch may close the step (FP32: the last launch the host
// enqueues in a round; a step that ends in an earlier launch is closed by the
// end pass of the launch after it, which the round holds)
template <class P, int SH = kSH, int DIAGF = 0, int EAGER = 0, int CHECK = 1, bool CLOSE = true>
__global__ void __launch_bounds__(kBlock) v2_substep_kernel(V2Args<P> a, int k)
{
    constexpr int kRows = SH * (kBlock / kWave);     // rows of a tile
    __shared__ double smem[64 * 4];
    __shared__ double fold_lds[kSubFold * kBlock];
    SyncBlock *s = a.sync;
    const mv2_step_config &cfg = a.cfg;
    const bool lead = blockIdx.x == 0 && threadIdx.x == 0;
    const int src = k & 1;
    const int lane = threadIdx.x & (kWave - 1);
    const int ny = a.ny, nx = a.nx;
    if constexpr (P::kQuant) {
        // Every kernel argument the prologue uses is made ready here, before
        // the first branch, so the scalar loads go out as one batch with one
        // wait; read at first use they would each be waited for in turn
        // ahead of the tile's loads.
        __asm__ volatile("" ::"s"(ny), "s"(nx), "s"(a.b), "s"(a.h[0]), "s"(a.h[1]), "s"(a.qx[0]), "s"(a.qx[1]),
                         "s"(a.qy[0]), "s"(a.qy[1]), "s"(a.spread), "s"(a.diag_part), "s"(a.diag_groups),
                         "s"(a.sync), "s"(a.ext_x), "s"(k), "s"((int)gridDim.x), "s"(a.cfg.density),
                         "s"(a.clamp_quarter), "s"(a.spec_true));
    }
    const int tiles_x = (nx + kStripW - 1) / kStripW;
    const int tiles = tiles_x * ((ny + kRows - 1) / kRows);
    int t = 0;
    if constexpr (P::kQuant) {
        // one tile per workgroup (the launch's own grid) needs no division
        if (gridDim.x == tiles) {
            t = xcd_block();
        } else {
            int last;
            xcd_tile_range(tiles, t, last);
        }
    }
    // this wave's band of the tile
    const int r0 = (t / tiles_x) * kRows + wave_index() * SH;
    const int c0 = (t % tiles_x) * kStripW;
    // FP32: each wave keeps its own diagnostic partials (global wave index)
    const int gw = blockIdx.x * (kBlock / kWave) + wave_index();
    // Requests that do not depend on the loop state.  Loads return in issue
    // order, so the spread slots (needed first, for the substep's dt) go out
    // before the tile, whose rows are then waited for one by one as the
    // update reaches them; the diagnostic slots (needed last) go out after
    // it.  All are vector loads (lane-dependent addresses): a scalar load
    // here would be waited for before the tile's loads could issue.
    // FP32: every wave folds the 64 slots of the previous launch's maximum
    // itself (no barrier); FP64: the first wave, through LDS.  FP32 also
    // reads the true maximum of the step's input that a speculative launch 0
    // left for launch 1.
    // FP32: every wave reads and folds the slots itself (no barrier; a
    // first-wave fold shared through LDS measured slower).  CHECK launches
    // (the first two of a step) also read the sets of the input's static
    // maxima and true CFL maximum, which the loop-state test of a step's
    // first substep reads; every load is unconditional (loads under a branch
    // make the compiler wait for every load at the join).
    double slot = 0.0, tslot = 0.0, gslot = 0.0, qslot = 0.0;
    if constexpr (P::kQuant) {
        const int sl = lane * kSpreadStride;
        slot = __longlong_as_double((long long)a.spread[(k % kWaveSets) * kSpreadWords + sl]);
        if constexpr (CHECK) {
            tslot = __longlong_as_double((long long)a.spec_true[sl]);
            gslot = __longlong_as_double((long long)a.spread[kSetStaticGrad * kSpreadWords + sl]);
            qslot = __longlong_as_double((long long)a.spread[kSetStaticQ * kSpreadWords + sl]);
        }
    } else {
        if (threadIdx.x < kWave)
            slot = __longlong_as_double((long long)a.spread[(k % kWaveSets) * kSpreadWords + lane * kSpreadStride]);
    }
    StripRows<SH> rows;
    if constexpr (P::kQuant) {
        if constexpr (EAGER) strip_load<SH>(a, src, r0, c0, rows);
    }
    // the diagnostic slots accumulate over the working launches; a step's
    // first substep starts them from zero.  Lane d holds field d: of every
    // wave (FP32) or of the workgroup's first wave (FP64).
    double dprev = 0.0;
    if constexpr (P::kQuant) {
        // unconditional (clamped) for the same reason as the tile's loads
        dprev = a.diag_part[(int64_t)min(lane, kDiag - 1) * a.diag_groups + gw];
    }
    if (k != 0) {
        if constexpr (!P::kQuant) {
            if (threadIdx.x < kDiag) dprev = a.diag_part[(int64_t)threadIdx.x * a.diag_groups + blockIdx.x];
        }
    }
    Control C = load_launch_record(&s->ctl[k & 1]);
    const int flags = load_launch_record(&s->flags[k & 1]);
    // FP32: each wave's folds of the slots (the same in every wave)
    double fold_true = 0.0, fold_grad = 0.0, fold_q = 0.0;
    if constexpr (P::kQuant && CHECK) {
        fold_true = wave_max(tslot);
        fold_grad = wave_max(gslot);
        fold_q = wave_max(qslot);
    }
    // FP32 speculative step, launch 1: launch 0 took substep 0 with the
    // predicted CFL maximum.  A static input ends the step (the input in
    // buffer 0 is untouched); a true maximum other than the prediction
    // restarts the step from buffer 0 here (the failure flags of the
    // discarded substep are dropped); otherwise the loop goes on.
    bool restart = false, spec_static = false;
    double maxw_true = 0.0;
    if (P::kQuant && CHECK && C.spec) {
        const double pred = C.pred;
        bool is_static = false;
        if (cfg.static_exit) {
            // warp_mobile_v2.py:148-160 - globally static state below the start angle
            const double ext_max = a.ext_x ? a.ext_max : npmax(0.0, npmax(fabs(cfg.external_ax), fabs(cfg.external_ay)));
            is_static = ext_max <= 1.0e-14 && fold_q <= 1.0e-14 && fold_grad <= cfg.tan_start + 1.0e-14;
        }
        if (is_static) {
            C = fresh_control(cfg.dt);
            C.static_exit = 1;
            C.done = 1;
            spec_static = true;
        } else {
            maxw_true = fold_true;
            if (maxw_true != pred) {
                C = fresh_control(cfg.dt);
                C.restarted = 1;
                restart = true;
                // the flags launch 0's discarded substep left (read above by
                // every workgroup of this launch and dropped); launch 2 ORs
                // its own into this word
                if (lead) s->flags[k & 1] = 0;
            } else {
                C.spec = 0;
                apply_failure_flags(C, flags);
            }
        }
    } else {
        apply_failure_flags(C, flags);
        if (!C.done && k == 0 && cfg.static_exit && (!P::kQuant || CHECK)) {
            // warp_mobile_v2.py:148-160 - globally static state below the start angle
            const double ext_max = a.ext_x ? a.ext_max : npmax(0.0, npmax(fabs(cfg.external_ax), fabs(cfg.external_ay)));
            const double qmax = P::kQuant ? fold_q : spread_read_max(a.spread + kSetStaticQ * kSpreadWords, smem);
            const double gmax = P::kQuant ? fold_grad : spread_read_max(a.spread + kSetStaticGrad * kSpreadWords, smem);
            if (ext_max <= 1.0e-14 && qmax <= 1.0e-14 && gmax <= cfg.tan_start + 1.0e-14) {
                C.static_exit = 1;
                C.done = 1;
            }
        }
    }
    if (!C.done && !(C.remaining > 1.0e-14)) C.done = 1;
    if (!C.done && cfg.limit_mode == 0 && C.substeps >= cfg.maximum_substeps) {
        C.numeric = 2;
        C.done = 1;
    }
    if (C.done) {
        if constexpr (P::kQuant) {
            // the first launch past the loop's end: copy back + ledger after,
            // unless the last substep stored the ledger (end_done 2: the fold
            // pass copies back)
            if (C.end_done == 0) {
                if (!C.static_exit) end32_body<SH>(a, C.src);
                C.end_done = 1;
            }
            // a static input after a speculative substep 0: its diagnostics are void
            if (spec_static && lane < kDiag) a.diag_part[(int64_t)lane * a.diag_groups + gw] = 0.0;
        }
        if (lead) s->ctl[(k + 1) & 1] = C;
        return;
    }
    if constexpr (P::kQuant) {
        // CFL maximum of the current state (mobile_v2_reference.py:239-251)
        const double maxw = restart ? maxw_true : wave_max(slot);
        if (EAGER && C.src != src) {
            // The state is in the other buffer (a restarted speculative
            // step): pass the loop state on unchanged, with the maximum in the
            // next launch's set, so that launch k + 1 finds the state in
            // buffer (k + 1) & 1 as its early staging assumes.  (A reload
            // here would join two paths with different loads in flight, and
            // the compiler's wait at the join would hold every launch's first
            // row until its whole tile had arrived.)
            if (blockIdx.x == 0) spread_clear(a.spread + ((k + 2) % kWaveSets) * kSpreadWords);
            if (lead) {
                atomic_max_nonneg(&a.spread[((k + 1) % kWaveSets) * kSpreadWords], maxw);
                s->ctl[(k + 1) & 1] = C;
            }
            return;
        }
        // launches expected to find the loop ended stage the state only now
        if constexpr (!EAGER) strip_load<SH>(a, C.src, r0, c0, rows);
        // a step's first substep starts its diagnostic slots from zero
        substep32_work<SH, DIAGF, CLOSE>(a, rows, C, maxw, k, r0, c0, C.substeps > 0 && lane < kDiag ? dprev : 0.0,
                                         -1.0);
    } else {
        // CFL maximum of the current state (mobile_v2_reference.py:239-251)
        if (threadIdx.x < kWave) {
            const double v = wave_max(slot);
            if (threadIdx.x == 0) smem[0] = v;
        }
        __syncthreads();
        const double maxw = smem[0];
        const double mind = a.dx <= a.dy ? a.dx : a.dy;
        const double stable = maxw <= 1.0e-14 ? C.remaining : cfg.cfl * mind / maxw;
        const double sub_dt = npmin(C.remaining, stable);
        if (blockIdx.x == 0) spread_clear(a.spread + ((k + 2) % kWaveSets) * kSpreadWords);
        if (lead) {
            Control nc = C;
            nc.last_remaining = C.remaining;
            nc.last_substeps = C.substeps;
            nc.last_src = C.src;
            nc.max_cfl = npmax(C.max_cfl, maxw * sub_dt / mind);
            if (sub_dt < C.remaining - 1.0e-14) nc.cfl_limited = 1;
            nc.remaining = C.remaining - sub_dt;
            nc.substeps = C.substeps + 1;
            nc.src = C.src ^ 1;
            // the loop's own exits, in the reference's order
            if (cfg.limit_mode == 1 && nc.substeps > cfg.maximum_substeps) {
                nc.numeric = 2;
                nc.done = 1;
            } else if (!(nc.remaining > 1.0e-14)) {
                nc.done = 1;
            } else if (cfg.limit_mode == 0 && nc.substeps >= cfg.maximum_substeps) {
                nc.numeric = 2;
                nc.done = 1;
            }
            s->ctl[(k + 1) & 1] = nc;
        }
        Acc<P> acc;
        for (int d = 0; d < kDiag; d++) acc.diag[d] = 0.0;
        double wmax[1] = {0.0};
        double hmin = 0.0;
        int ovf = 0;
        {
            const int n = ny * nx;
            for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
                const int r = i / nx, c = i - r * nx;
                CellOut o = cell64(a, src, r, c, sub_dt, acc);
                wmax[0] = npmax(wmax[0], o.wave);
                hmin = npmin(hmin, o.hmin);
                ovf |= o.overflow;
            }
        }
        // one field-parallel fold of the workgroup's values: 0 wave-speed
        // maximum, 1 overflow flag (maximum), 2 minimum new depth, 3..
        // diagnostic sums
        double fv[kSubFold];
        fv[0] = wmax[0];
        fv[1] = ovf ? 1.0 : 0.0;
        fv[2] = hmin;
#pragma unroll
        for (int d = 0; d < kDiag; d++) fv[3 + d] = acc.diag[d];
        block_fold<double, kSubFold, 0x3u, 0x4u>(fv, fold_lds);
        if (threadIdx.x == 0) {
            spread_max(a.spread + ((k + 1) % kWaveSets) * kSpreadWords, fv[0]);
            int f = 0;
            if (fv[2] < -2.0e-13) f |= 1;
            if (fv[1] != 0.0) f |= 2;
            if (f) atomicOr(&s->flags[(k + 1) & 1], f);
        }
        // each workgroup owns its diagnostic slot across the substep
        // launches; the folded sums stay in the fold's LDS (field f at f x kBlock)
        if (threadIdx.x < kDiag)
            a.diag_part[(int64_t)threadIdx.x * a.diag_groups + blockIdx.x] =
                dprev + fold_lds[(3 + threadIdx.x) * kBlock];
    }
}

// copy back when the final state sits in scratch; ledger after
template <class P>
__global__ void __launch_bounds__(kBlock) v2_end_kernel(V2Args<P> a, int k)
{
    __shared__ double smem[64 * 4];
    const Control C = resolve_control(a.sync, k);
    if (C.static_exit) return;
    const int fsrc = C.src;
    const int n = a.ny * a.nx;
    const int stride = gridDim.x * blockDim.x;
    LedgerAcc L;
    L.init();
    for (int base = blockIdx.x * blockDim.x + threadIdx.x; base < n; base += stride * kLedgerU) {
        typename P::hst h[kLedgerU];
        typename P::real qx[kLedgerU], qy[kLedgerU], b[kLedgerU];
#pragma unroll
        for (int u = 0; u < kLedgerU; u++) {
            const int i = base + u * stride;
            if (i >= n) continue;
            b[u] = a.b[i];
            h[u] = a.h[fsrc][i];
            qx[u] = a.qx[fsrc][i];
            qy[u] = a.qy[fsrc][i];
        }
#pragma unroll
        for (int u = 0; u < kLedgerU; u++) {
            const int i = base + u * stride;
            if (i >= n) continue;
            if (fsrc == 1) {
                a.h[0][i] = h[u];
                a.qx[0][i] = qx[u];
                a.qy[0][i] = qy[u];
            }
            CellVals v;
            v.b = (double)b[u];
            v.depth = (double)P::depth(h[u]);
            v.quanta = P::kQuant ? (long long)h[u] : 0;
            v.qx = (double)qx[u];
            v.qy = (double)qy[u];
            L.add<P>(a, v, i);
        }
    }
    store_ledger<P>(a, L, false, smem);
}

// The begin pass over the wave's band of the input state (rows staged with
// their halo, buffer 0): ledger before, static diagnostic and the CFL maximum
// of the first substep, which goes to the spread set `cfl_set` (as
// v2_begin_kernel, FP32 strip layout).
template <int SH>
__device__ inline void begin32_work(const V2Args<F32P> &a, const StripRows<SH> &R, int r0, int c0,
                                    unsigned long long *cfl_set)
{
    const int ny = a.ny, nx = a.nx;
    const int lane = threadIdx.x & (kWave - 1);
    const int gc = c0 - 1 + lane;
    const bool owner = lane >= 1 && lane <= kStripW && gc < nx;
    const bool stat = a.cfg.static_exit != 0;
    double H[SH + 2];
#pragma unroll
    for (int j = 0; j < SH + 2; j++) H[j] = (double)R.b[j] + (double)((float)R.hq[j] * (float)kQuantum);
    const float area = uniform((float)(a.dx * a.dy)), f6 = uniform((float)(a.dx * a.dy / 6.0));
    const float dry = uniform((float)a.cfg.dry_tolerance);
    const float Kg = uniform((float)(a.cfg.earth_pressure * a.cfg.gravity));
    const StaticGrad sg(a.cfg.tan_start, a.dx, a.dy, gc > 0 && gc + 1 < nx);
    const ColWeights cw(a.cfg.weight_mode, ny, nx, area, f6, gc);
    Ledger32 L;
    L.init();
    double m[3] = {0.0, 0.0, 0.0};              // max |grad H|^2, max |q|, max wave speed
#pragma unroll
    for (int j = 1; j <= SH; j++) {
        const int r = r0 - 1 + j;
        const double HR = from_next_lane_d(H[j]), HL = from_prev_lane_d(H[j]);
        if (!owner || r >= ny) continue;
        const int hq = R.hq[j];
        const float qxv = R.qx[j], qyv = R.qy[j];
        m[2] = npmax(m[2], (double)L.add(r, ny, hq, qxv, qyv, R.b[j], cw.at(r, ny), dry, Kg));
        if (stat) {
            const double hl = gc > 0 ? HL : H[j], hr = gc + 1 < nx ? HR : H[j];
            const double hd = r > 0 ? H[j - 1] : H[j], hu = r + 1 < ny ? H[j + 1] : H[j];
            m[0] = npmax(m[0], sg.grad2(hl, hr, hd, hu, r > 0 && r + 1 < ny));
            m[1] = npmax(m[1], (double)fmaxf(fabsf(qxv), fabsf(qyv)));
        }
    }
    store_ledger32(a, L, gc, owner, true);
    // the wave's maxima, one spread slot per wave (no barrier)
    const double g2 = wave_max(m[0]), qm = wave_max(m[1]), wm = wave_max(m[2]);
    if (lane == 0) {
        const int slot = ((blockIdx.x * (kBlock / kWave) + wave_index()) % kSpread) * kSpreadStride;
        if (stat) {
            // the square root is monotonic: the root of the largest square is the largest root
            atomic_max_nonneg(&a.spread[kSetStaticGrad * kSpreadWords + slot], sqrt(g2));
            atomic_max_nonneg(&a.spread[kSetStaticQ * kSpreadWords + slot], qm);
        }
        atomic_max_nonneg(&cfl_set[slot], wm);
    }
}

// One pass over the input state (a step without speculation).
template <int SH>
__global__ void __launch_bounds__(kBlock) v2_begin32_kernel(V2Args<F32P> a)
{
    const int ny = a.ny, nx = a.nx;
    // the kernel arguments of the prologue in one batch of scalar loads
    __asm__ volatile("" ::"s"(ny), "s"(nx), "s"(a.b), "s"(a.h[0]), "s"(a.qx[0]), "s"(a.qy[0]),
                     "s"(a.cfg.static_exit), "s"(a.cfg.weight_mode), "s"(a.cfg.dry_tolerance),
                     "s"(a.cfg.earth_pressure), "s"(a.cfg.gravity), "s"(a.cfg.tan_start), "s"(a.dx), "s"(a.dy),
                     "s"((int)gridDim.x));
    const int tiles_x = (nx + kStripW - 1) / kStripW;
    const int t = xcd_block();
    const int r0 = (t / tiles_x) * (SH * (kBlock / kWave)) + wave_index() * SH, c0 = (t % tiles_x) * kStripW;
    StripRows<SH> R;
    strip_load<SH>(a, 0, r0, c0, R);
    begin32_work<SH>(a, R, r0, c0, a.spread);
    if (blockIdx.x == 0 && threadIdx.x == 0) a.sync->ctl[0].remaining = a.cfg.dt;
}

// Launch 0 of a speculative FP32 step: the begin pass over the input (ledger
// before, static diagnostic, the true CFL maximum into the spec_true set) and
// substep 0 with the CFL maximum the previous step's end pass predicted, in
// one launch.  The input in buffer 0 is not written; launch 1 keeps the
// substep if the true maximum equals the prediction and the input is not
// static, else restarts from buffer 0 or ends the step.
template <int SH, int DIAGF>
__global__ void __launch_bounds__(kBlock) v2_first32_kernel(V2Args<F32P> a)
{
    const int ny = a.ny, nx = a.nx;
    __asm__ volatile("" ::"s"(ny), "s"(nx), "s"(a.b), "s"(a.h[0]), "s"(a.h[1]), "s"(a.qx[0]), "s"(a.qx[1]),
                     "s"(a.qy[0]), "s"(a.qy[1]), "s"(a.spread), "s"(a.diag_part), "s"(a.diag_groups),
                     "s"(a.sync), "s"(a.ext_x), "s"((int)gridDim.x), "s"(a.pred), "s"(a.spec_true),
                     "s"(a.cfg.static_exit), "s"(a.cfg.weight_mode), "s"(a.cfg.dry_tolerance),
                     "s"(a.cfg.earth_pressure), "s"(a.cfg.gravity), "s"(a.cfg.tan_start), "s"(a.dx), "s"(a.dy));
    const int tiles_x = (nx + kStripW - 1) / kStripW;
    const int t = xcd_block();
    const int r0 = (t / tiles_x) * (SH * (kBlock / kWave)) + wave_index() * SH, c0 = (t % tiles_x) * kStripW;
    const int lane = threadIdx.x & (kWave - 1);
    const double pslot = __longlong_as_double((long long)a.pred[lane * kSpreadStride]);
    StripRows<SH> rows;
    strip_load<SH>(a, 0, r0, c0, rows);
    begin32_work<SH>(a, rows, r0, c0, a.spec_true);
    const double pred = wave_max(pslot);
    substep32_work<SH, DIAGF, false>(a, rows, fresh_control(a.cfg.dt), pred, 0, r0, c0, 0.0, pred);
}

// copy back when the final state sits in scratch; ledger after (FP32 strips).
// The body runs in the first substep launch that finds the loop ended (same
// tile grid as the substeps); each wave stores its own ledger partial.
template <int SH>
__device__ inline void end32_body(const V2Args<F32P> &a, int fsrc)
{
    const int ny = a.ny, nx = a.nx;
    // the configuration the ledger needs, in one batch of scalar loads
    __asm__ volatile("" ::"s"(a.cfg.weight_mode), "s"(a.cfg.dry_tolerance), "s"(a.cfg.density),
                     "s"(a.cfg.earth_pressure), "s"(a.cfg.gravity), "s"(a.dx), "s"(a.dy), "s"(a.groups),
                     "s"(a.ledger_f), "s"(a.ledger_q));
    const int tiles_x = (nx + kStripW - 1) / kStripW;
    const int t = xcd_block();
    const int r0 = (t / tiles_x) * (SH * (kBlock / kWave)) + wave_index() * SH, c0 = (t % tiles_x) * kStripW;
    const int lane = threadIdx.x & (kWave - 1);
    const int gc = c0 - 1 + lane;
    const bool owner = lane >= 1 && lane <= kStripW && gc < nx;
    const int32_t *gh = fsrc ? a.h[1] : a.h[0];
    const float *gqx = fsrc ? a.qx[1] : a.qx[0], *gqy = fsrc ? a.qy[1] : a.qy[0];
    int hq[SH];
    float qxv[SH], qyv[SH], bv[SH];
#pragma unroll
    for (int j = 0; j < SH; j++) {
        const int r = r0 + j;
        hq[j] = 0; qxv[j] = 0.0f; qyv[j] = 0.0f; bv[j] = 0.0f;
        if (owner && r < ny) {
            const int i = r * nx + gc;
            bv[j] = a.b[i];
            hq[j] = gh[i];
            qxv[j] = gqx[i];
            qyv[j] = gqy[i];
        }
    }
    const float area = uniform((float)(a.dx * a.dy)), f6 = uniform((float)(a.dx * a.dy / 6.0));
    const float dry = uniform((float)a.cfg.dry_tolerance);
    const float Kg = uniform((float)(a.cfg.earth_pressure * a.cfg.gravity));
    const ColWeights cw(a.cfg.weight_mode, ny, nx, area, f6, gc);
    Ledger32 L;
    L.init();
    double pred = 0.0;
#pragma unroll
    for (int j = 0; j < SH; j++) {
        const int r = r0 + j;
        if (!owner || r >= ny) continue;
        if (fsrc == 1) {
            const int i = r * nx + gc;
            a.h[0][i] = hq[j];
            a.qx[0][i] = qxv[j];
            a.qy[0][i] = qyv[j];
        }
        pred = npmax(pred, (double)L.add(r, ny, hq[j], qxv[j], qyv[j], bv[j], cw.at(r, ny), dry, Kg));
    }
    store_ledger32(a, L, gc, owner, false);
    // the CFL maximum of the result, as the next step's begin pass would
    // compute it: that step's launch 0 takes its first substep with it and
    // launch 1 checks it against the true maximum of that step's input
    const double pm = wave_max(pred);
    if (lane == 0)
        atomic_max_nonneg(&a.pred_next[((blockIdx.x * (kBlock / kWave) + wave_index()) % kSpread) * kSpreadStride], pm);
}

template <int SH>
__global__ void __launch_bounds__(kBlock) v2_end32_kernel(V2Args<F32P> a, int k)
{
    const Control C = resolve_control(a.sync, k);
    if (C.static_exit) return;
    end32_body<SH>(a, C.src);
}

// Fold the partial ledgers and diagnostics in fixed order.  One wave per
// field: lane l sums the partials l, l + 64, ... (their loads issued
// together), then a DPP tree combines the lanes -- no cross-wave or
// cross-workgroup step, so the workgroups of the launch run independently.
// The results go straight to pinned host memory (FoldOut); the last wave
// copies out the resolved loop state and zeroes the next step's loop state
// (the other parity's block and spread sets, unused during this step).
// FP32 launches carry more workgroups after the fold's, which copy the final
// state back into buffer 0 when the last substep left it in buffer 1.  The
// host sees the record before those copies end; later work on the stream is
// ordered after them.
constexpr int kFoldF = 2 * kSum + kDiag + 2;   // double fields: ledgers, diagnostics, static maxima
constexpr int kFoldQ = 2 * kClasses;           // int64 fields: quanta per weight class
// field 3 and kSum + 3 are maxima (speed), 7 and kSum + 7 minima (depth), the
// last two (static-gradient and momentum spread maxima) maxima
constexpr unsigned kFoldMax = (1u << 3) | (1u << (kSum + 3)) | (3u << (2 * kSum + kDiag));
constexpr unsigned kFoldMin = (1u << 7) | (1u << (kSum + 7));
constexpr int kFoldWaves = kFoldF + kFoldQ + 1;
static_assert(kFoldWaves <= kStampWords, "one completion stamp per fold wave");
constexpr int kFoldUnroll = 8;
constexpr int kFoldBlocks = (kFoldWaves + kBlock / kWave - 1) / (kBlock / kWave);
// FP32: workgroups after the fold's copy the final state back into buffer 0
// when the last substep stored the closing ledger and left the state in
// buffer 1, kCopyU cells per thread (kBlock x kCopyU consecutive cells per
// workgroup)
constexpr int kCopyU = 8;

struct FoldOut {
    double f[kFoldF];
    long long q[kFoldQ];
    Control ctl;
};
static_assert(sizeof(FoldOut) <= 1024, "the fold output fits the pinned result block");

// a wave's reduction of n values, lane l taking l, l + 64, ...; U loads per
// lane are in flight at once
template <int OP, class T, int U = kFoldUnroll>
__device__ inline T lane_fold(const T *src, int n, T ident)
{
    const int lane = threadIdx.x & (kWave - 1);
    T acc = ident;
    for (int p0 = lane; p0 < n; p0 += kWave * U) {
        T x[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int p = p0 + u * kWave;
            x[u] = p < n ? src[p] : ident;
        }
#pragma unroll
        for (int u = 0; u < U; u++) acc = red_op<OP>(acc, x[u]);
    }
    return wave_reduce<OP>(acc);
}

template <class P>
__global__ void __launch_bounds__(kBlock) v2_fold_kernel(V2Args<P> a, int k, int nparts, int ndiag, FoldOut *out,
                                                          unsigned *stamps, unsigned stamp)
{
    if constexpr (P::kQuant) {
        if (blockIdx.x >= kFoldBlocks) {
            const Control C = resolve_control(a.sync, k);
            if (!(C.done && C.end_done == 2 && C.src == 1)) return;
            const int n = a.ny * a.nx;
            const int base = (blockIdx.x - kFoldBlocks) * (kBlock * kCopyU) + threadIdx.x;
            int hv[kCopyU];
            float xv[kCopyU], yv[kCopyU];
#pragma unroll
            for (int u = 0; u < kCopyU; u++) {
                const int i = base + u * kBlock;
                if (i < n) {
                    hv[u] = a.h[1][i];
                    xv[u] = a.qx[1][i];
                    yv[u] = a.qy[1][i];
                }
            }
#pragma unroll
            for (int u = 0; u < kCopyU; u++) {
                const int i = base + u * kBlock;
                if (i < n) {
                    a.h[0][i] = hv[u];
                    a.qx[0][i] = xv[u];
                    a.qy[0][i] = yv[u];
                }
            }
            return;
        }
    }
    const int w = blockIdx.x * (kBlock / kWave) + wave_index();
    const int lane = threadIdx.x & (kWave - 1);
    if (w < 2 * kSum) {
        const double *src = a.ledger_f + (int64_t)w * nparts;
        double v;
        // FP32: one partial per wave of the ledger passes (1 728 at 701 x 701)
        if ((kFoldMax >> w) & 1u) v = lane_fold<1, double, 32>(src, nparts, 0.0);
        else if ((kFoldMin >> w) & 1u) v = lane_fold<2, double, 32>(src, nparts, 1e300);
        else v = lane_fold<0, double, 32>(src, nparts, 0.0);
        if (lane == 0) out->f[w] = v;
    } else if (w < 2 * kSum + kDiag) {
        const int d = w - 2 * kSum;
        // FP32: one partial per substep wave (1 728 at 701 x 701)
        const double v = lane_fold<0, double, 32>(a.diag_part + (int64_t)d * ndiag, ndiag, 0.0);
        if (lane == 0) out->f[w] = v;
    } else if (w < kFoldF) {
        const int set = w == 2 * kSum + kDiag ? kSetStaticGrad : kSetStaticQ;
        const double x = __longlong_as_double((long long)a.spread[set * kSpreadWords + lane * kSpreadStride]);
        const double v = wave_reduce<1>(x);
        if (lane == 0) out->f[w] = v;
    } else if (w < kFoldF + kFoldQ) {
        const int q = w - kFoldF;
        const long long v = lane_fold<0, long long, 32>(a.ledger_q + (int64_t)q * nparts, nparts, 0ll);
        if (lane == 0) out->q[q] = v;
    } else if (w == kFoldWaves - 1) {
        if (lane == 0) out->ctl = resolve_control(a.sync, k);
        unsigned *ps = reinterpret_cast<unsigned *>(a.sync_next);
        for (int i = lane; i < (int)(sizeof(SyncBlock) / 4); i += kWave) ps[i] = 0u;
        for (int i = lane; i < kStepSpreadSets * kSpread; i += kWave) a.spread_next[(int64_t)i * kSpreadStride] = 0ull;
        if constexpr (P::kQuant) {
            // the next step's true-maximum set, and this step's (consumed) prediction
            a.spec_true_next[lane * kSpreadStride] = 0ull;
            a.pred[lane * kSpreadStride] = 0ull;
        }
    }
    // each wave's lane 0 wrote one field of the host record: stamp word w
    if (w < kFoldWaves && lane == 0) host_stamp(&stamps[w], stamp);
}

// the step record from the folded values (host side, after the stream sync)
template <class P>
static void assemble_record(const FoldOut &fo, const V2Args<P> &a, mv2_step_record *rec, double static_out[2])
{
    const Control &C = fo.ctl;
    const bool after_valid = !C.static_exit;
    for (int q = 0; q < kSum; q++) {
        rec->summary_before[q] = fo.f[q];
        rec->summary_after[q] = after_valid ? fo.f[kSum + q] : fo.f[q];
    }
    for (int d = 0; d < kDiag; d++) rec->diag[d] = (after_valid && C.substeps > 0) ? fo.f[2 * kSum + d] : 0.0;
    rec->mass_quanta_before = rec->mass_quanta_after = 0.0;
    if (P::kQuant) {
        // exact mobile volume: integer quanta per weight class times its weight
        const double area = a.dx * a.dy, f6 = a.dx * a.dy / 6.0;
        const int rep_r[3] = {0, 1, a.ny - 1}, rep_c[3] = {0, 1, a.nx - 1};
        double mb = 0.0, ma = 0.0, qb = 0.0, qa = 0.0;
        for (int q = 0; q < kClasses; q++) {
            const double w = a.cfg.weight_mode ? vertex_weight(a.ny, a.nx, f6, rep_r[q / 3], rep_c[q % 3]) : area;
            mb += ((double)fo.q[q] * kQuantum) * w;
            ma += ((double)fo.q[kClasses + q] * kQuantum) * w;
            qb += (double)fo.q[q];
            qa += (double)fo.q[kClasses + q];
        }
        rec->summary_before[0] = mb;
        rec->summary_after[0] = after_valid ? ma : mb;
        rec->mass_quanta_before = qb;
        rec->mass_quanta_after = after_valid ? qa : qb;
    }
    static_out[0] = fo.f[2 * kSum + kDiag];
    static_out[1] = fo.f[2 * kSum + kDiag + 1];
    rec->substeps = C.substeps;
    rec->cfl_limited = C.cfl_limited;
    rec->static_exit = C.static_exit;
    rec->maximum_cfl = C.max_cfl;
    rec->remaining = C.remaining;
    rec->numeric_error = C.numeric;
    rec->status = C.numeric ? MV2_ENUMERIC : MV2_OK;
    rec->complete = C.done && (!P::kQuant || C.end_done);
}

// Launch geometry of one mv2_step call.
struct StepLaunch {
    int band;          // FP32 band height (rows per wave)
    int gs, gu, gfold; // begin / end workgroups, substep workgroups, fold workgroups
    FoldOut *fold;     // pinned host record the fold pass writes
};

template <class P>
static void enqueue_begin(const V2Args<P> &a, const StepLaunch &L, hipStream_t st)
{
    if constexpr (P::kQuant) {
        switch (L.band) {
        case 2: hipLaunchKernelGGL(v2_begin32_kernel<2>, dim3(L.gs), dim3(kBlock), 0, st, a); break;
        case 3: hipLaunchKernelGGL(v2_begin32_kernel<3>, dim3(L.gs), dim3(kBlock), 0, st, a); break;
        case 5: hipLaunchKernelGGL(v2_begin32_kernel<5>, dim3(L.gs), dim3(kBlock), 0, st, a); break;
        case 6: hipLaunchKernelGGL(v2_begin32_kernel<6>, dim3(L.gs), dim3(kBlock), 0, st, a); break;
        default: hipLaunchKernelGGL(v2_begin32_kernel<4>, dim3(L.gs), dim3(kBlock), 0, st, a); break;
        }
    } else {
        hipLaunchKernelGGL(v2_begin_kernel<P>, dim3(L.gs), dim3(kBlock), 0, st, a);
    }
}

// launch 0 of a speculative FP32 step (begin pass + substep 0)
template <class P>
static void enqueue_first(const V2Args<P> &a, const StepLaunch &L, hipStream_t st)
{
    if constexpr (P::kQuant) {
        const bool d = a.cfg.diagnostics == 1;
        switch (L.band) {
        case 2: if (d) hipLaunchKernelGGL((v2_first32_kernel<2, 1>), dim3(L.gu), dim3(kBlock), 0, st, a);
                else hipLaunchKernelGGL((v2_first32_kernel<2, 0>), dim3(L.gu), dim3(kBlock), 0, st, a); break;
        case 3: if (d) hipLaunchKernelGGL((v2_first32_kernel<3, 1>), dim3(L.gu), dim3(kBlock), 0, st, a);
                else hipLaunchKernelGGL((v2_first32_kernel<3, 0>), dim3(L.gu), dim3(kBlock), 0, st, a); break;
        case 5: if (d) hipLaunchKernelGGL((v2_first32_kernel<5, 1>), dim3(L.gu), dim3(kBlock), 0, st, a);
                else hipLaunchKernelGGL((v2_first32_kernel<5, 0>), dim3(L.gu), dim3(kBlock), 0, st, a); break;
        case 6: if (d) hipLaunchKernelGGL((v2_first32_kernel<6, 1>), dim3(L.gu), dim3(kBlock), 0, st, a);
                else hipLaunchKernelGGL((v2_first32_kernel<6, 0>), dim3(L.gu), dim3(kBlock), 0, st, a); break;
        default: if (d) hipLaunchKernelGGL((v2_first32_kernel<4, 1>), dim3(L.gu), dim3(kBlock), 0, st, a);
                 else hipLaunchKernelGGL((v2_first32_kernel<4, 0>), dim3(L.gu), dim3(kBlock), 0, st, a); break;
        }
    }
}

// one FP32 substep launch of band height SH, all ten diagnostic sums or the
// production record's five
template <class P, int SH, int DIAGF, int EAGER>
static void launch_substep32_c(const V2Args<P> &a, int grid, hipStream_t st, int k, bool close)
{
    // the first two launches of a step test the static exit / the speculation
    if (k <= 1) {
        if (close) hipLaunchKernelGGL((v2_substep_kernel<P, SH, DIAGF, EAGER, 1, true>), dim3(grid), dim3(kBlock), 0, st, a, k);
        else hipLaunchKernelGGL((v2_substep_kernel<P, SH, DIAGF, EAGER, 1, false>), dim3(grid), dim3(kBlock), 0, st, a, k);
    } else {
        if (close) hipLaunchKernelGGL((v2_substep_kernel<P, SH, DIAGF, EAGER, 0, true>), dim3(grid), dim3(kBlock), 0, st, a, k);
        else hipLaunchKernelGGL((v2_substep_kernel<P, SH, DIAGF, EAGER, 0, false>), dim3(grid), dim3(kBlock), 0, st, a, k);
    }
}

template <class P, int SH>
static void launch_substep32(const V2Args<P> &a, int grid, hipStream_t st, int k, int eager, bool close)
{
    const int d = a.cfg.diagnostics == 1 ? 1 : 0;
    if (d && eager) launch_substep32_c<P, SH, 1, 1>(a, grid, st, k, close);
    else if (d) launch_substep32_c<P, SH, 1, 0>(a, grid, st, k, close);
    else if (eager) launch_substep32_c<P, SH, 0, 1>(a, grid, st, k, close);
    else launch_substep32_c<P, SH, 0, 0>(a, grid, st, k, close);
}

// substep launches k0 .. k0 + count - 1.  Timing mode brackets the launches
// expected to work (0 .. eager_until - 1 of the first round) with one event
// pair: events between launches would add their own latency to every launch.
template <class P>
static int enqueue_substeps(mv2_context *ctx, const V2Args<P> &a, const StepLaunch &L, hipStream_t st, int k0,
                            int count, int eager_until, bool timing, int first_timed = 0)
{
    for (int k = k0; k < k0 + count; k++) {
        if (timing && k == first_timed) MV2_HIP(ctx, hipEventRecord(ctx->sub_events[0], st));
        const int eager = k < eager_until ? 1 : 0;
        if constexpr (P::kQuant) {
            // only the round's last launch carries the close path
            const bool close = k == k0 + count - 1;
            switch (L.band) {
            case 2: launch_substep32<P, 2>(a, L.gu, st, k, eager, close); break;
            case 3: launch_substep32<P, 3>(a, L.gu, st, k, eager, close); break;
            case 5: launch_substep32<P, 5>(a, L.gu, st, k, eager, close); break;
            case 6: launch_substep32<P, 6>(a, L.gu, st, k, eager, close); break;
            default: launch_substep32<P, 4>(a, L.gu, st, k, eager, close); break;
            }
        } else {
            hipLaunchKernelGGL(v2_substep_kernel<P>, dim3(L.gu), dim3(kBlock), 0, st, a, k);
        }
        if (timing && k0 == first_timed && k == eager_until - 1) MV2_HIP(ctx, hipEventRecord(ctx->sub_events[1], st));
    }
    return MV2_OK;
}

// after launch k - 1: FP32's first launch past the loop's end did the end pass;
// FP64 runs it as a launch of its own.  Then the fold into the host record,
// whose workgroups stamp `stamps` with `stamp` (0: no stamps).
template <class P>
static void enqueue_close(const V2Args<P> &a, const StepLaunch &L, hipStream_t st, int k, unsigned *stamps,
                          unsigned stamp)
{
    if constexpr (!P::kQuant) hipLaunchKernelGGL(v2_end_kernel<P>, dim3(L.gs), dim3(kBlock), 0, st, a, k);
    hipLaunchKernelGGL(v2_fold_kernel<P>, dim3(L.gfold), dim3(kBlock), 0, st, a, k, a.groups, a.diag_groups, L.fold, stamps,
                       stamp);
}

constexpr size_t kStepGraphsMax = 16;   // instantiated sequences kept per context
constexpr size_t kStepSeenMax = 16;     // keys remembered from a single sighting

// The first round of a step (begin, `chunk` substeps, end / fold) replayed
// from a cached graph when an earlier call had the same launch arguments.  A
// key is captured on its second sighting, so callers whose arguments change
// every call (a varying dt) stay on plain launches.  The graph is captured on
// a private stream (the bound stream may be the null stream, which cannot be
// captured) and launched on `st`; it records no memset and no host call.
template <class P>
static int replay_step(mv2_context *ctx, const V2Args<P> &a, const StepLaunch &L, int chunk, int eager_until,
                       hipStream_t st, bool *replayed)
{
    *replayed = false;
    struct Shape {
        StepLaunch launch;
        int chunk, eager_until, quantised, pad;
    } shape;
    std::memset(&shape, 0, sizeof(shape));
    shape.launch = L;
    shape.chunk = chunk;
    shape.eager_until = eager_until;
    shape.quantised = P::kQuant ? 1 : 0;
    std::vector<unsigned char> key(sizeof(Shape) + sizeof(V2Args<P>));
    std::memcpy(key.data(), &shape, sizeof(Shape));
    std::memcpy(key.data() + sizeof(Shape), &a, sizeof(V2Args<P>));
    StepGraph *hit = nullptr;
    for (StepGraph &g : ctx->step_graphs)
        if (g.key == key) {
            hit = &g;
            break;
        }
    if (!hit) {
        auto seen = std::find(ctx->step_seen.begin(), ctx->step_seen.end(), key);
        if (seen == ctx->step_seen.end()) {
            if (ctx->step_seen.size() >= kStepSeenMax) ctx->step_seen.erase(ctx->step_seen.begin());
            ctx->step_seen.push_back(std::move(key));
            return MV2_OK;
        }
        ctx->step_seen.erase(seen);
        if (!ctx->capture_stream)
            MV2_HIP(ctx, hipStreamCreateWithFlags(&ctx->capture_stream, hipStreamNonBlocking));
        hipStream_t cs = ctx->capture_stream;
        MV2_HIP(ctx, hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
        enqueue_begin<P>(a, L, cs);
        enqueue_substeps<P>(ctx, a, L, cs, 0, chunk, eager_until, false);
        enqueue_close<P>(a, L, cs, chunk, nullptr, 0u);
        hipGraph_t graph = nullptr;
        hipError_t e = hipStreamEndCapture(cs, &graph);
        if (e == hipSuccess) e = hipGetLastError();
        if (e != hipSuccess) {
            if (graph) hipGraphDestroy(graph);
            return check_hip(ctx, e, "capture of the step's launch sequence");
        }
        hipGraphExec_t exec = nullptr;
        e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        hipGraphDestroy(graph);
        if (e != hipSuccess) return check_hip(ctx, e, "hipGraphInstantiate");
        if (ctx->step_graphs.size() >= kStepGraphsMax) {
            auto oldest = std::min_element(ctx->step_graphs.begin(), ctx->step_graphs.end(),
                                           [](const StepGraph &x, const StepGraph &y) { return x.last_use < y.last_use; });
            hipGraphExecDestroy(oldest->exec);
            ctx->step_graphs.erase(oldest);
        }
        ctx->step_graphs.push_back(StepGraph{std::move(key), exec, 0});
        hit = &ctx->step_graphs.back();
    }
    hit->last_use = ++ctx->graph_clock;
    MV2_HIP(ctx, hipGraphLaunch(hit->exec, st));
    ctx->graph_replays++;
    *replayed = true;
    return MV2_OK;
}

template <class P>
static int launch_step(mv2_context *ctx, const mv2_step_config *cfg, mv2_step_record *out, int fixed_chunk = 0)
{
    const int64_t n = (int64_t)ctx->ny * ctx->nx;
    // FP32 band height of the substep (rows per wave): the workgroups of a
    // launch spread over the CUs, and the CUs that receive one more
    // workgroup than the rest set the launch time.  The band minimises that
    // busiest CU's work, ceil(workgroups / CUs) x (band + 1.5) -- a wave's
    // rows plus its halo rows and fixed costs (701^2: 5 rows, 432
    // workgroups, at most 2 per CU, against 3 per CU with 4 rows).
    // MV2_FRAME_BAND_ROWS forces a band of 2 to 6.
    int band = 4;
    if (P::kQuant) {
        const int strips = (ctx->nx + kStripW - 1) / kStripW;
        const int cus = ctx->cu_count > 0 ? ctx->cu_count : 256;
        double best = 1e300;
        for (int cand = 3; cand <= 6; cand++) {
            const int wgs = strips * ((ctx->ny + cand * (kBlock / kWave) - 1) / (cand * (kBlock / kWave)));
            const double cost = (double)((wgs + cus - 1) / cus) * (cand + 1.5);
            if (cost < best - 1e-9) {
                best = cost;
                band = cand;
            }
        }
        if (const char *e = getenv("MV2_FRAME_BAND_ROWS")) {
            const int forced = atoi(e);
            if (forced >= 2 && forced <= 6) band = forced;
        }
    }
    const int rows_sub = band * (kBlock / kWave);
    const int gu = P::kQuant ? ((ctx->nx + kStripW - 1) / kStripW) * ((ctx->ny + rows_sub - 1) / rows_sub)
                             : grid_for(n);                              // substep launches
    // the begin / end passes use the substep's tiles (same band)
    const int gl = P::kQuant ? ((ctx->nx + kStripW - 1) / kStripW) * ((ctx->ny + rows_sub - 1) / rows_sub) : 0;
    // begin / end passes: the substep tiles (FP32), a grid stride (FP64)
    const int gs = P::kQuant ? gl : grid_for(n, kBlock * kLedgerU, kLedgerGroupsMax);
    // ledger partials: one per wave of the FP32 begin / end passes, one per workgroup in FP64
    const int gpart = P::kQuant ? gs * (kBlock / kWave) : gs;
    if (int rc = ensure_scratch(ctx, ctx->partials,
                                (sizeof(double) * kLedgerF + sizeof(long long) * kLedgerQ) * (size_t)gpart))
        return rc;
    // diagnostic partials: one per substep wave (FP32) or workgroup (FP64)
    const int gdiag = P::kQuant ? gu * (kBlock / kWave) : gu;
    if (int rc = ensure_scratch(ctx, ctx->step_diag, sizeof(double) * kDiag * (size_t)gdiag)) return rc;
    if (int rc = ensure_scratch(ctx, ctx->spread, kSpreadScratchBytes)) return rc;
    V2Args<P> a;
    std::memset(&a, 0, sizeof(a));      // padding bytes are part of the graph key
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
    a.ext_x = (const typename P::real *)ctx->ext_x;
    a.ext_y = (const typename P::real *)ctx->ext_y;
    a.ext_max = ctx->ext_max;
    const int parity = ctx->v2_frame & 1;
    unsigned long long *spread_even = (unsigned long long *)ctx->spread.ptr;
    unsigned long long *spread_odd = spread_even + (size_t)kStepOddSpreadSet * kSpreadWords;
    a.sync = (SyncBlock *)((char *)ctx->sync.ptr + parity * kSyncParityBytes);
    a.sync_next = (SyncBlock *)((char *)ctx->sync.ptr + (parity ^ 1) * kSyncParityBytes);
    a.spread_next = parity ? spread_even : spread_odd;
    a.ledger_f = (double *)ctx->partials.ptr;
    a.ledger_q = (long long *)(a.ledger_f + (size_t)kLedgerF * gpart);
    a.groups = gpart;
    a.diag_part = (double *)ctx->step_diag.ptr;
    a.diag_groups = gdiag;
    a.spread = parity ? spread_odd : spread_even;
    a.spec_true = spread_even + (size_t)(kSpecTrueSet + parity) * kSpreadWords;
    a.spec_true_next = spread_even + (size_t)(kSpecTrueSet + (parity ^ 1)) * kSpreadWords;
    a.pred = spread_even + (size_t)(kSpecPredSet + parity) * kSpreadWords;
    a.pred_next = spread_even + (size_t)(kSpecPredSet + (parity ^ 1)) * kSpreadWords;
    // the fold pass writes its results straight into pinned host memory
    // (device-visible): no copy launch before the stream sync
    FoldOut *fold = (FoldOut *)ctx->host_result.ptr;
    a.clamp_quarter = cfg->cfl <= 0.25 ? 1 : 0;
    hipStream_t st = ctx->stream;
    const bool timing = ctx->timing != 0;
    // plain launches wait on the fold's completion stamps; timing mode (whose
    // events must have completed before they are read) and graph replays
    // (whose captured fold carries no stamp) wait on the stream
    const bool stamped = !timing && !(ctx->graphs && fixed_chunk == 0);
    if (timing) MV2_HIP(ctx, hipEventRecord(ctx->ev_start, st));
    // this parity's loop state is zero if the previous step completed; an
    // error anywhere below leaves v2_clean at -1 and the next step resets
    // Speculation (FP32): the previous step left a CFL prediction of this
    // state (its end pass, and no library call wrote mobile or momentum
    // since) and this parity's loop state is clean; not under graph replay
    // or a fixed chunk (the summary pass).
    const bool spec = P::kQuant && ctx->v2_spec && ctx->v2_pred_ready && ctx->v2_clean == parity &&
                      !ctx->graphs && fixed_chunk == 0;
    ctx->v2_pred_ready = 0;
    if (ctx->v2_clean != parity)
        hipLaunchKernelGGL(v2_reset_kernel, dim3(1), dim3(kBlock), 0, st, a.sync, a.spread, a.spec_true, a.pred_next);
    ctx->v2_clean = -1;
    int k = 0;
    // Substep launches of the first round, for a step that takes as many
    // substeps as the previous one: FP32's last substep stores the closing
    // ledger itself (a speculative launch 0 cannot: a step of one substep
    // closes in launch 1, after the check); FP64 closes in a launch of its own
    // after one more that finds the loop ended.
    const int pred_n = ctx->v2_predicted;
    int chunk = fixed_chunk > 0 ? fixed_chunk
                : pred_n <= 0 ? 4
                : !P::kQuant ? pred_n + 1
                : spec ? (pred_n > 2 ? pred_n : 2) : pred_n;
    // launches expected to work stage their tile before the loop state arrives
    int eager_until = fixed_chunk > 0 ? 0 : (pred_n > 0 ? (P::kQuant ? chunk : pred_n) : 4);
    // timing mode: launches first .. bracket - 1 between the event pair (a
    // speculative step's launch 0 also carries the begin pass: left out)
    const int first_timed = spec ? 1 : 0;
    const int bracket = eager_until;
    const int gcopy = P::kQuant ? (int)((n + kBlock * kCopyU - 1) / (kBlock * kCopyU)) : 0;
    const StepLaunch L{band, gs, gu, kFoldBlocks + gcopy, fold};
    // the first round (begin, predicted substeps, end / fold) as one graph launch
    bool replayed = false;
    if (ctx->graphs && !timing && fixed_chunk == 0)
        if (int rc = replay_step<P>(ctx, a, L, chunk, eager_until, st, &replayed)) return rc;
    if (replayed) {
        k = chunk;
    } else if (spec) {
        // launch 0: begin pass and the speculative substep 0
        enqueue_first<P>(a, L, st);
        k = 1;
        chunk -= 1;
    } else {
        enqueue_begin<P>(a, L, st);
    }
    MV2_HIP(ctx, hipGetLastError());
    unsigned stamp = 0u;
    for (;;) {
        if (!replayed) {
            if (timing && (int)ctx->sub_events.size() < 2) {
                while ((int)ctx->sub_events.size() < 2) {
                    hipEvent_t e;
                    MV2_HIP(ctx, hipEventCreate(&e));
                    ctx->sub_events.push_back(e);
                }
            }
            if (int rc = enqueue_substeps<P>(ctx, a, L, st, k, chunk, eager_until, timing, first_timed)) return rc;
            k += chunk;
            stamp = stamped ? next_stamp(ctx, kFoldWaves) : 0u;
            enqueue_close<P>(a, L, st, k, stamp_words(ctx), stamp);
            MV2_HIP(ctx, hipGetLastError());
        } else {
            stamp = 0u;
        }
        replayed = false;
        if (timing) MV2_HIP(ctx, hipEventRecord(ctx->ev_stop, st));
        if (int rc = stamp_wait(ctx, st, stamp, kFoldWaves)) return rc;
        if (fold->ctl.done && (!P::kQuant || fold->ctl.end_done)) break;
        // more substeps than enqueued: the state so far is valid, enqueue more
        chunk = chunk * 2 > 4 ? chunk * 2 : 4;
        eager_until = k + chunk;
    }
    *out = mv2_step_record{};
    assemble_record<P>(*fold, a, out, ctx->v2_static);
    ctx->v2_clean = parity ^ 1;          // zeroed by this step's fold pass
    ctx->v2_frame++;
    ctx->v2_last_sync = parity;
    if (fixed_chunk == 0) ctx->v2_predicted = out->substeps;
    // the end pass ran (no static exit): it left the CFL prediction of this
    // result for the next step
    if (P::kQuant && !out->static_exit) ctx->v2_pred_ready = 1;
    if (spec) {
        ctx->spec_steps++;
        if (fold->ctl.restarted) ctx->spec_restarts++;
    }
    out->launches = k;
    out->kernel_ms = 0.0;
    out->substep_kernel_ms = 0.0;
    if (timing) {
        float ms = 0.0f;
        MV2_HIP(ctx, hipEventElapsedTime(&ms, ctx->ev_start, ctx->ev_stop));
        out->kernel_ms = ms;
    }
    out->timed_launches = 0;
    if (timing && bracket > first_timed && out->substeps == bracket) {
        // substeps are the first launches: the bracket held exactly the working ones
        float e = 0.0f;
        MV2_HIP(ctx, hipEventElapsedTime(&e, ctx->sub_events[0], ctx->sub_events[1]));
        out->substep_kernel_ms = e;
        out->timed_launches = bracket - first_timed;
    }
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
    if (int rc = launch_step<P>(ctx, &c, &rec, 1)) return rc;
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
    if (cfg->diagnostics != 0 && cfg->diagnostics != 1)
        return set_error(ctx, MV2_EINVAL, "[MobileV2] diagnostics must be 0 (all sums) or 1 (production record)");
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
    int rc = ctx->precision == MV2_FP64 ? launch_step<F64P>(ctx, &c, &rec, 1)
                                        : launch_step<F32P>(ctx, &c, &rec, 1);
    if (rc) return rc;
    out[0] = ctx->v2_static[0];
    out[1] = ctx->v2_static[1];
    return MV2_OK;
}

extern "C" int mv2_set_external(mv2_context *ctx, const void *ax, const void *ay, double max_abs)
{
    if (!ctx) return set_error(nullptr, MV2_EINVAL, "[MobileV2] null context");
    if ((ax == nullptr) != (ay == nullptr))
        return set_error(ctx, MV2_EINVAL, "MobileV2 external acceleration needs both components");
    if (ax && !(max_abs >= 0.0 && std::isfinite(max_abs)))
        return set_error(ctx, MV2_EINVAL, "MobileV2 external acceleration must be finite");
    ctx->ext_x = ax;
    ctx->ext_y = ay;
    ctx->ext_max = ax ? max_abs : 0.0;
    return MV2_OK;
}

extern "C" int mv2_set_timing(mv2_context *ctx, int enable)
{
    if (!ctx) return set_error(nullptr, MV2_EINVAL, "[MobileV2] null context");
    ctx->timing = enable ? 1 : 0;
    return MV2_OK;
}

extern "C" int mv2_set_graphs(mv2_context *ctx, int enable)
{
    if (!ctx) return set_error(nullptr, MV2_EINVAL, "[MobileV2] null context");
    ctx->graphs = enable ? 1 : 0;
    return MV2_OK;
}

extern "C" long long mv2_graph_replays(const mv2_context *ctx) { return ctx ? ctx->graph_replays : -1; }

extern "C" int mv2_set_speculation(mv2_context *ctx, int enable)
{
    if (!ctx) return set_error(nullptr, MV2_EINVAL, "[MobileV2] null context");
    ctx->v2_spec = enable ? 1 : 0;
    return MV2_OK;
}

extern "C" int mv2_speculation_stats(const mv2_context *ctx, long long out[2])
{
    if (!ctx || !out) return set_error(nullptr, MV2_EINVAL, "[MobileV2] null argument");
    out[0] = ctx->spec_steps;
    out[1] = ctx->spec_restarts;
    return MV2_OK;
}
