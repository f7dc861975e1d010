// Batched Mobile V2 step: hundreds of independent terrains of uneven size
// advance together (FP32 fields, int32 mobile quanta).
//
// The instances' fields are packed back to back in caller-owned arrays; every
// instance keeps its own CFL loop (own dt sequence, own substep count, own
// static early exit), exactly as if it were stepped alone.  One launch per
// substep covers the tiles of *all* instances, so a launch moves the whole
// batch (hundreds of MB) and the per-launch costs that dominate a single 701^2
// terrain are amortised.
//
// Per-instance loop state follows the single-terrain design: launch k reads
// ctl[k & 1][i] and the instance's first tile writes ctl[(k+1) & 1][i]; the
// instance's CFL maximum is a spread maximum over 16 slots of the instance
// (tile g uses slot g mod 16) in one of four rotating sets (launch k reads set
// k % 4, fills k+1, clears k+2).  Launch indices enter only as k mod 8, so the
// host can capture eight substep launches once into a HIP graph and replay it.
//
// Tiles are 62 columns by four wave bands of SH rows (the strip layout of
// mv2_strip32.hpp); SH is a template parameter chosen by the host-side
// autotuner.  Every face is computed from the same two cells with the same
// instructions whatever the band height, so fields are bit-identical across
// shapes and to the single-terrain step of each instance.
#include "mv2_internal.hpp"
#include "mv2_strip32.hpp"
#include "mv2_relax_tile.hpp"
#include "mv2_batch.hpp"

#include <cmath>
#include <cstring>
#include <vector>

namespace mv2 {

constexpr int kInstSlots = 16;          // spread slots per instance
constexpr int kInstStride = 16;         // u64 words between slots (128 B)
constexpr int kInstWords = kInstSlots * kInstStride;
constexpr int kWaveSetsB = 4;           // rotating CFL sets
constexpr int kBLedger = 3;             // px, py, kinetic (doubles) per tile and pass

struct BControl {
    double remaining;
    double max_cfl;
    double last_remaining;
    int substeps, src;
    int last_substeps, last_src;
    int done, numeric, cfl_limited, static_exit;
    int rolled, started;
};

struct BatchArgs {
    int count, tiles;
    double dx, dy;
    const float *b;
    int32_t *h[2];
    float *qx[2];
    float *qy[2];
    const int32_t *inst_ny, *inst_nx;
    const int64_t *inst_off;
    const int32_t *inst_first_tile, *inst_tiles;
    const int32_t *tile_inst, *tile_r0, *tile_c0;
    // loop state of this step's parity (steps alternate between two copies;
    // the finalize pass zeroes the other copy for the next step)
    BControl *ctl;                      // [2][count]
    int *flags;                         // [2][count]
    unsigned long long *wave;           // [4][count][kInstWords]
    unsigned long long *stat;           // [2][count][kInstWords]: max |grad H|, max |q|
    BControl *ctl_next;
    int *flags_next;
    unsigned long long *wave_next, *stat_next;
    double *ledger;                     // [2 passes][kBLedger][tiles]
    long long *quanta;                  // [2 passes][tiles]
    mv2_batch_record *rec;              // [count]
    mv2_step_config cfg;
    int clamp_quarter;
};

// the arguments of one instance in the shape the tile functions expect
struct InstView {
    int ny, nx;
    double dx, dy;
    const float *b;
    int32_t *h[2];
    float *qx[2];
    float *qy[2];
    mv2_step_config cfg;                // a copy: no address of the kernel arguments is taken
    int clamp_quarter;
    const float *ext_x = nullptr, *ext_y = nullptr;   // batches take the uniform cfg acceleration
    __device__ InstView(const BatchArgs &a, int i)
        : ny(a.inst_ny[i]), nx(a.inst_nx[i]), dx(a.dx), dy(a.dy), cfg(a.cfg), clamp_quarter(a.clamp_quarter)
    {
        const int64_t off = a.inst_off[i];
        b = a.b + off;
        for (int s = 0; s < 2; s++) {
            h[s] = a.h[s] + off;
            qx[s] = a.qx[s] + off;
            qy[s] = a.qy[s] + off;
        }
    }
};

__device__ inline unsigned long long *inst_set(unsigned long long *base, int set, int count, int i)
{
    return base + ((int64_t)set * count + i) * kInstWords;
}

__device__ inline void inst_max(unsigned long long *slots, double v)
{
    atomic_max_nonneg(&slots[(blockIdx.x % kInstSlots) * kInstStride], v);
}

// fold an instance's 16 slots: every thread of the block receives the value
__device__ inline double inst_read_max(const unsigned long long *slots, double *smem1)
{
    if (threadIdx.x < kWave) {
        double v = threadIdx.x < kInstSlots ? __longlong_as_double((long long)slots[threadIdx.x * kInstStride]) : 0.0;
        v = wave_max(v);
        if (threadIdx.x == 0) *smem1 = v;
    }
    __syncthreads();
    const double r = *smem1;
    __syncthreads();
    return r;
}

// the same fold in one wave (every wave of the block folds the slots
// itself: no barrier)
__device__ inline double inst_read_max_wave(const unsigned long long *slots)
{
    const int lane = threadIdx.x & (kWave - 1);
    const double v = __longlong_as_double((long long)slots[(lane % kInstSlots) * kInstStride]);
    return wave_max(lane < kInstSlots ? v : 0.0);
}

__device__ inline BControl resolve_bcontrol(const BatchArgs &a, int k, int i)
{
    BControl C = a.ctl[(k & 1) * a.count + i];
    const int f = a.flags[(k & 1) * a.count + i];
    if (f && !C.rolled) {
        C.numeric = (f & 2) ? 3 : 1;
        C.src = C.last_src;
        C.remaining = C.last_remaining;
        C.substeps = C.last_substeps;
        C.done = 1;
        C.rolled = 1;
    }
    return C;
}

// zero the loop state of one instance (words of the four arrays)
__device__ inline void zero_instance_state(BControl *ctl, int *flags, unsigned long long *wave,
                                           unsigned long long *stat, int count, int i, int t, int nt)
{
    constexpr int kCtlWords = (int)(sizeof(BControl) / 4);
    for (int w = t; w < 2 * kCtlWords; w += nt)
        reinterpret_cast<unsigned *>(ctl + (int64_t)(w / kCtlWords) * count + i)[w % kCtlWords] = 0u;
    if (t < 2) flags[(int64_t)t * count + i] = 0;
    for (int w = t; w < kWaveSetsB * kInstSlots; w += nt)
        wave[((int64_t)(w / kInstSlots) * count + i) * kInstWords + (w % kInstSlots) * kInstStride] = 0ull;
    for (int w = t; w < 2 * kInstSlots; w += nt)
        stat[((int64_t)(w / kInstSlots) * count + i) * kInstWords + (w % kInstSlots) * kInstStride] = 0ull;
}

// zero this parity's loop state (first step, or after an error); a zero
// record is an instance that has not started: its first substep launch
// takes remaining = dt
__global__ void batch_reset_kernel(BatchArgs a)
{
    for (int i = blockIdx.x; i < a.count; i += gridDim.x)
        zero_instance_state(a.ctl, a.flags, a.wave, a.stat, a.count, i, threadIdx.x, blockDim.x);
}

// One tile: ledger before, static diagnostic, CFL maximum of the input.  The
// tile's waves use the strip layout of the substep: lanes 1..62 own the
// columns, a lane holds its column's rows, and the static gradient takes its
// x-neighbours by DPP and its y-neighbours from the rows in registers.
template <int SH>
__global__ void __launch_bounds__(kBlock) batch_begin_kernel(BatchArgs a)
{
    __shared__ double lds[6 * kBlock];
    const int g = xcd_block();
    const int i = a.tile_inst[g];
    const InstView v(a, i);
    const int lane = threadIdx.x & (kWave - 1);
    const int r0 = a.tile_r0[g] + wave_index() * SH, c0 = a.tile_c0[g];
    const int ny = v.ny, nx = v.nx;
    const int gc = c0 - 1 + lane;
    const bool owner = lane >= 1 && lane <= kStripW && gc < nx;
    const double K = a.cfg.earth_pressure, gr = a.cfg.gravity, dry = a.cfg.dry_tolerance;
    const float Qf = (float)kQuantum;
    StripRows<SH> R;
    strip_load<SH>(v, 0, r0, c0, R);
    double H[SH + 2];
#pragma unroll
    for (int j = 0; j < SH + 2; j++) H[j] = (double)R.b[j] + (double)((float)R.hq[j] * Qf);
    double s[kBLedger] = {0.0, 0.0, 0.0};
    long long q[1] = {0};
    double m[3] = {0.0, 0.0, 0.0};                // max |grad H|^2, max |q|, max wave speed
    const StaticGrad sg(a.cfg.tan_start, a.dx, a.dy, gc > 0 && gc + 1 < nx);
    const float Kg = uniform((float)(K * gr)), dryf = uniform((float)dry);
#pragma unroll
    for (int j = 1; j <= SH; j++) {
        const int r = r0 - 1 + j;
        const double HR = from_next_lane_d(H[j]), HL = from_prev_lane_d(H[j]);
        if (!owner || r >= ny) continue;
        const int32_t hq = R.hq[j];
        const float d = (float)hq * Qf;
        const float qxv = R.qx[j], qyv = R.qy[j];
        q[0] += hq;
        s[0] += (double)qxv;
        s[1] += (double)qyv;
        float vx = 0.0f, vy = 0.0f;
        if (d > dryf) {
            const float inv = rcpf(d);
            vx = qxv * inv;
            vy = qyv * inv;
            s[2] += (double)(0.5f * (qxv * vx + qyv * vy));
        }
        m[2] = npmax(m[2], (double)(sqrtf_hw(vx * vx + vy * vy) + sqrtf_hw(Kg * d)));
        if (a.cfg.static_exit) {
            // central differences inside, one-sided at the terrain edge
            const double hl = gc > 0 ? HL : H[j], hr = gc + 1 < nx ? HR : H[j];
            const double hd = r > 0 ? H[j - 1] : H[j], hu = r + 1 < ny ? H[j + 1] : H[j];
            const double g2 = sg.grad2(hl, hr, hd, hu, r > 0 && r + 1 < ny);
            m[0] = npmax(m[0], g2);
            m[1] = npmax(m[1], fabs((double)qxv));
            m[1] = npmax(m[1], fabs((double)qyv));
        }
    }
    double f6[6] = {s[0], s[1], s[2], m[0], m[1], m[2]};
    block_fold<double, 6, 0x38u>(f6, lds);
    block_fold<long long, 1>(q, reinterpret_cast<long long *>(lds));
    if (threadIdx.x == 0) {
        for (int f = 0; f < kBLedger; f++) a.ledger[(int64_t)f * a.tiles + g] = f6[f];
        a.quanta[g] = q[0];
        inst_max(inst_set(a.wave, 0, a.count, i), f6[5]);
        if (a.cfg.static_exit) {
            // the square root is monotonic: the root of the largest square is the largest root
            inst_max(inst_set(a.stat, 0, a.count, i), sqrt(f6[3]));
            inst_max(inst_set(a.stat, 1, a.count, i), f6[4]);
        }
    }
}

// One substep of every instance whose loop has not ended (k is used mod 8).
template <int SH>
__global__ void __launch_bounds__(kBlock) batch_substep_kernel(BatchArgs a, int k)
{
    __shared__ double smem[64 * 4];
    const int g = xcd_block();
    const int i = a.tile_inst[g];
    const bool lead = g == a.inst_first_tile[i] && threadIdx.x == 0;
    const mv2_step_config &cfg = a.cfg;
    BControl C = resolve_bcontrol(a, k, i);
    if (!C.done && !C.started) {
        C.started = 1;
        C.remaining = cfg.dt;
        if (cfg.static_exit) {
            const double ext_max = npmax(0.0, npmax(fabs(cfg.external_ax), fabs(cfg.external_ay)));
            const double qmax = inst_read_max(inst_set(a.stat, 1, a.count, i), smem);
            const double gmax = inst_read_max(inst_set(a.stat, 0, a.count, i), smem);
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
        if (lead) a.ctl[((k + 1) & 1) * a.count + i] = C;
        return;
    }
    const InstView v(a, i);
    const int src = C.src;
    const int r0 = a.tile_r0[g] + wave_index() * SH, c0 = a.tile_c0[g];
    StripRows<SH> rows;
    const double maxw = inst_read_max_wave(inst_set(a.wave, k % kWaveSetsB, a.count, i));
    strip_load<SH>(v, src, r0, c0, rows);
    const double mind = a.dx <= a.dy ? a.dx : a.dy;
    const double stable = maxw <= 1.0e-14 ? C.remaining : cfg.cfl * mind / maxw;
    const double sub_dt = npmin(C.remaining, stable);
    if (g == a.inst_first_tile[i] && threadIdx.x < kInstSlots)
        inst_set(a.wave, (k + 2) % kWaveSetsB, a.count, i)[threadIdx.x * kInstStride] = 0ull;
    if (lead) {
        BControl nc = C;
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
        a.ctl[((k + 1) & 1) * a.count + i] = nc;
    }
    float dg[kDiag];
    for (int d = 0; d < kDiag; d++) dg[d] = 0.0f;
    float wv = 0.0f;
    long long hn_min = 0;
    int ovf = 0;
    if (strip_interior<SH>(v.ny, v.nx, r0, c0))
        strip_update<SH, InstView, 0, true>(v, rows, src, r0, c0, sub_dt, dg, wv, hn_min, ovf);
    else
        strip_update<SH, InstView, 0, false>(v, rows, src, r0, c0, sub_dt, dg, wv, hn_min, ovf);
    // per wave, no barrier: the wave-speed maximum into one of the instance's
    // slots, a failure (negative thickness before clamping, int32 range) into
    // its flags
    const float wmx = wave_max(wv);
    const bool neg = __ballot(hn_min < 0) != 0ull;
    const bool rng = __ballot(ovf != 0) != 0ull;
    if ((threadIdx.x & (kWave - 1)) == 0) {
        unsigned long long *slots = inst_set(a.wave, (k + 1) % kWaveSetsB, a.count, i);
        const int slot = (g * (kBlock / kWave) + wave_index()) % kInstSlots;
        atomic_max_nonneg(&slots[slot * kInstStride], (double)wmx);
        const int f = (neg ? 1 : 0) | (rng ? 2 : 0);
        if (f) atomicOr(&a.flags[((k + 1) & 1) * a.count + i], f);
    }
}

// copy back when an instance's final state sits in scratch; ledger after
template <int SH>
__global__ void __launch_bounds__(kBlock) batch_end_kernel(BatchArgs a, int k)
{
    __shared__ double lds[kBLedger * kBlock];
    const int g = xcd_block();
    const int i = a.tile_inst[g];
    const BControl C = resolve_bcontrol(a, k, i);
    const InstView v(a, i);
    const int lane = threadIdx.x & (kWave - 1);
    const int r0 = a.tile_r0[g] + wave_index() * SH, c0 = a.tile_c0[g];
    const int ny = v.ny, nx = v.nx;
    const int gc = c0 - 1 + lane;
    const bool owner = lane >= 1 && lane <= kStripW && gc < nx;
    const double dry = a.cfg.dry_tolerance;
    const int fsrc = C.src;
    const int32_t *gh = fsrc ? v.h[1] : v.h[0];
    const float *gqx = fsrc ? v.qx[1] : v.qx[0], *gqy = fsrc ? v.qy[1] : v.qy[0];
    int32_t hq[SH];
    float qxv[SH], qyv[SH];
#pragma unroll
    for (int j = 0; j < SH; j++) {
        const int r = r0 + j;
        hq[j] = 0; qxv[j] = 0.0f; qyv[j] = 0.0f;
        if (owner && r < ny) {
            const int idx = r * nx + gc;
            hq[j] = gh[idx];
            qxv[j] = gqx[idx];
            qyv[j] = gqy[idx];
        }
    }
    double s[kBLedger] = {0.0, 0.0, 0.0};
    long long q[1] = {0};
#pragma unroll
    for (int j = 0; j < SH; j++) {
        const int r = r0 + j;
        if (!owner || r >= ny) continue;
        if (fsrc == 1) {
            const int idx = r * nx + gc;
            v.h[0][idx] = hq[j];
            v.qx[0][idx] = qxv[j];
            v.qy[0][idx] = qyv[j];
        }
        q[0] += hq[j];
        const double depth = (double)((float)hq[j] * (float)kQuantum);
        s[0] += (double)qxv[j];
        s[1] += (double)qyv[j];
        if (depth > dry) {
            const float inv = rcpf((float)depth);
            s[2] += (double)(0.5f * (qxv[j] * (qxv[j] * inv) + qyv[j] * (qyv[j] * inv)));
        }
    }
    block_fold<double, kBLedger>(s, lds);
    block_fold<long long, 1>(q, reinterpret_cast<long long *>(lds));
    if (threadIdx.x == 0) {
        for (int f = 0; f < kBLedger; f++) a.ledger[(int64_t)(kBLedger + f) * a.tiles + g] = s[f];
        a.quanta[a.tiles + g] = q[0];
    }
}

// one workgroup per instance: fold its tiles' partials in fixed order
__global__ void __launch_bounds__(kBlock) batch_finalize_kernel(BatchArgs a, int k)
{
    __shared__ double smem[64 * 4];
    const int i = blockIdx.x;
    const BControl C = resolve_bcontrol(a, k, i);
    const int t0 = a.inst_first_tile[i], nt = a.inst_tiles[i];
    double s[2 * kBLedger];
    for (int f = 0; f < 2 * kBLedger; f++) s[f] = 0.0;
    long long q[2] = {0, 0};
    for (int t = threadIdx.x; t < nt; t += blockDim.x) {
        for (int f = 0; f < 2 * kBLedger; f++) s[f] += a.ledger[(int64_t)f * a.tiles + t0 + t];
        q[0] += a.quanta[t0 + t];
        q[1] += a.quanta[a.tiles + t0 + t];
    }
    // the next step's loop state of this instance (other parity): unused now
    zero_instance_state(a.ctl_next, a.flags_next, a.wave_next, a.stat_next, a.count, i, threadIdx.x, blockDim.x);
    block_reduce<double, 2 * kBLedger, 0>(s, smem);
    block_reduce<long long, 2, 0>(q, reinterpret_cast<long long *>(smem));
    if (threadIdx.x == 0) {
        const double area = a.dx * a.dy;
        const bool after = !C.static_exit;
        mv2_batch_record r;
        r.substeps = C.substeps;
        r.numeric_error = C.numeric;
        r.static_exit = C.static_exit;
        r.complete = C.done;
        r.maximum_cfl = C.max_cfl;
        r.remaining = C.remaining;
        r.mass_quanta_before = (double)q[0];
        r.mass_quanta_after = after ? (double)q[1] : (double)q[0];
        r.momentum_before[0] = s[0] * area;
        r.momentum_before[1] = s[1] * area;
        r.momentum_after[0] = (after ? s[kBLedger + 0] : s[0]) * area;
        r.momentum_after[1] = (after ? s[kBLedger + 1] : s[1]) * area;
        r.kinetic_before = s[2] * area;
        r.kinetic_after = (after ? s[kBLedger + 2] : s[2]) * area;
        a.rec[i] = r;
    }
}

// ---------------------------------------------------------------------------
// Batched relaxation: the relaxation of every instance of the batch, each with
// its own iteration count and convergence test (as if relaxed alone), one
// launch per iteration over the 44 x 48 relaxation tiles of all instances
// (the single-field tile iteration of mv2_relax_tile.hpp with one iteration).
// ---------------------------------------------------------------------------
using BRelaxGeom = RelaxGeom<44, 48, 640, 1, false>;
constexpr int kBRelaxTH = BRelaxGeom::TH, kBRelaxTW = BRelaxGeom::TW;
struct BRelaxControl {
    int iterations, src, done, converged;
    double last;            // maximum excess of the last applied iteration
};

struct BatchRelaxArgs {
    int count, tiles;
    float *h[2];                        // packed field and its ping-pong copy
    const int32_t *inst_ny, *inst_nx;
    const int64_t *inst_off;
    const int32_t *inst_first, *inst_tiles;
    const int32_t *tile_inst, *tile_r0, *tile_c0;
    float crit[4];
    int max_iterations;
    double tolerance;
    BRelaxControl *ctl;                 // [2][count]
    unsigned long long *spread;         // [3][count][kInstWords]: maximum excess
    mv2_relax_result *out;              // [count]
    const int32_t *tile_list;           // tiles to run (the instances still relaxing), null: every tile
    int32_t *tchanged;                  // [2][tiles]: the tile's interior changed in launch k (parity k & 1)
    double *texcess;                    // [tiles]: the tile's maximum excess in its last computed iteration
};

// tile of this workgroup: XCD-aware order over every tile or over the list
__device__ inline int brelax_tile(const BatchRelaxArgs &a)
{
    const int g = xcd_block();
    return a.tile_list ? a.tile_list[g] : g;
}

// loop state of instance i for launch k (every thread of the block calls it)
__device__ inline BRelaxControl brelax_resolve(const BatchRelaxArgs &a, int k, int i, double *smem1)
{
    BRelaxControl C = a.ctl[(k & 1) * a.count + i];
    if (!C.done && k > 0) {
        C.last = inst_read_max(inst_set(a.spread, (k - 1) % 3, a.count, i), smem1);
        if (C.last <= a.tolerance) {
            C.converged = 1;
            C.done = 1;
        } else if (C.iterations >= a.max_iterations) {
            C.done = 1;
        }
    }
    return C;
}

__global__ void batch_relax_reset_kernel(BatchRelaxArgs a)
{
    for (int i = blockIdx.x; i < a.count; i += gridDim.x) {
        for (int w = threadIdx.x; w < 2 * (int)(sizeof(BRelaxControl) / 4); w += blockDim.x) {
            constexpr int kW = (int)(sizeof(BRelaxControl) / 4);
            reinterpret_cast<unsigned *>(a.ctl + (int64_t)(w / kW) * a.count + i)[w % kW] = 0u;
        }
        for (int w = threadIdx.x; w < 3 * kInstSlots; w += blockDim.x)
            inst_set(a.spread, w / kInstSlots, a.count, i)[(w % kInstSlots) * kInstStride] = 0ull;
    }
}

__global__ void __launch_bounds__(BRelaxGeom::BLOCK) batch_relax_kernel(BatchRelaxArgs a, int k)
{
    __shared__ float tile[BRelaxGeom::kCells];
    __shared__ double smem[BRelaxGeom::BLOCK / kWave];
    const int g = brelax_tile(a);
    const int i = a.tile_inst[g];
    const bool first = g == a.inst_first[i];
    const BRelaxControl C = brelax_resolve(a, k, i, smem);
    if (C.done) {
        if (first && threadIdx.x == 0) a.ctl[((k + 1) & 1) * a.count + i] = C;
        return;
    }
    const int src = C.src;
    if (first && threadIdx.x < kInstSlots)
        inst_set(a.spread, (k + 1) % 3, a.count, i)[threadIdx.x * kInstStride] = 0ull;
    if (first && threadIdx.x == 0) {
        BRelaxControl nc = C;
        nc.iterations = C.iterations + 1;
        nc.src = src ^ 1;
        a.ctl[((k + 1) & 1) * a.count + i] = nc;
    }
    const int64_t off = a.inst_off[i];
    // Quiet tiles: an iteration of a tile reads its interior and an 8-cell
    // halo (smaller than a tile), i.e. its 3 x 3 tile neighbourhood.  When none of those tiles
    // changed in launch k - 1, the tile's input equals the input of its last
    // computed iteration, which changed nothing, so this iteration would
    // reproduce the same field (already in both ping-pong buffers) and the same
    // excess: the stored excess stands in for it.  Launch 0 runs every tile.
    if (k > 0) {
        const int nx = a.inst_nx[i], ny = a.inst_ny[i];
        const int tx = (nx + kBRelaxTW - 1) / kBRelaxTW, ty = (ny + kBRelaxTH - 1) / kBRelaxTH;
        const int trow = a.tile_r0[g] / kBRelaxTH, tcol = a.tile_c0[g] / kBRelaxTW;
        int flag = 0;
        if (threadIdx.x < 9) {
            const int rr = trow + (int)threadIdx.x / 3 - 1, cc = tcol + (int)threadIdx.x % 3 - 1;
            if (rr >= 0 && rr < ty && cc >= 0 && cc < tx)
                flag = a.tchanged[((k - 1) & 1) * a.tiles + a.inst_first[i] + rr * tx + cc];
        }
        if (!__syncthreads_or(flag)) {
            if (threadIdx.x == 0) {
                a.tchanged[(k & 1) * a.tiles + g] = 0;
                inst_max(inst_set(a.spread, k % 3, a.count, i), a.texcess[g]);
            }
            return;
        }
    }
    bool changed = false;
    float tm[1] = {0.0f};
    relax_tile_iterations<float, BRelaxGeom>(a.h[src] + off, a.h[src ^ 1] + off, a.inst_ny[i], a.inst_nx[i], a.crit,
                                             a.tile_r0[g], a.tile_c0[g], tile, 1, tm, &changed);
    double m[1] = {(double)tm[0]};
    block_reduce<double, 1, 1, BRelaxGeom::BLOCK / kWave>(m, smem);
    const int any = __syncthreads_or(changed ? 1 : 0);
    if (threadIdx.x == 0) {
        inst_max(inst_set(a.spread, k % 3, a.count, i), m[0]);
        a.texcess[g] = m[0];
        a.tchanged[(k & 1) * a.tiles + g] = any;
    }
}

// after launch k - 1: copy each tile's interior back when its instance's field
// sits in scratch; the instance's first tile writes the instance's result
__global__ void __launch_bounds__(kBlock) batch_relax_end_kernel(BatchRelaxArgs a, int k)
{
    __shared__ double smem[1];
    const int g = brelax_tile(a);
    const int i = a.tile_inst[g];
    const BRelaxControl C = brelax_resolve(a, k, i, smem);
    if (C.src == 1) {
        const int64_t off = a.inst_off[i];
        const int ny = a.inst_ny[i], nx = a.inst_nx[i], r0 = a.tile_r0[g], c0 = a.tile_c0[g];
        for (int e = threadIdx.x; e < kBRelaxTH * kBRelaxTW; e += blockDim.x) {
            const int r = r0 + e / kBRelaxTW, c = c0 + e % kBRelaxTW;
            if (r < ny && c < nx) a.h[0][off + (int64_t)r * nx + c] = a.h[1][off + (int64_t)r * nx + c];
        }
    }
    if (g == a.inst_first[i] && threadIdx.x == 0) {
        mv2_relax_result r;
        r.iterations = C.iterations;
        r.converged = C.converged;
        r.max_excess_last = C.last;
        r.complete = C.done;
        r.launches = k;
        a.out[i] = r;
    }
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
// tile shapes: kStripW columns by 4 bands of SH rows (one band per wave)
struct TileShape {
    int th, tw;
};
constexpr TileShape kShapes[] = {{8, kStripW}, {16, kStripW}, {32, kStripW}, {64, kStripW}};
constexpr int kShapeCount = 4;

template <int SH>
struct ShapeKernels {
    static_assert(SH * (kBlock / kWave) > 0, "bands per tile");
    static void begin(dim3 g, hipStream_t s, const BatchArgs &a)
    {
        batch_begin_kernel<SH><<<g, dim3(kBlock), 0, s>>>(a);
    }
    static void substep(dim3 g, hipStream_t s, const BatchArgs &a, int k)
    {
        batch_substep_kernel<SH><<<g, dim3(kBlock), 0, s>>>(a, k);
    }
    static void end(dim3 g, hipStream_t s, const BatchArgs &a, int k)
    {
        batch_end_kernel<SH><<<g, dim3(kBlock), 0, s>>>(a, k);
    }
};

static void launch_begin(int shape, dim3 g, hipStream_t s, const BatchArgs &a)
{
    switch (shape) {
    case 0: ShapeKernels<2>::begin(g, s, a); break;
    case 1: ShapeKernels<4>::begin(g, s, a); break;
    case 2: ShapeKernels<8>::begin(g, s, a); break;
    default: ShapeKernels<16>::begin(g, s, a); break;
    }
}
static void launch_substep(int shape, dim3 g, hipStream_t s, const BatchArgs &a, int k)
{
    switch (shape) {
    case 0: ShapeKernels<2>::substep(g, s, a, k); break;
    case 1: ShapeKernels<4>::substep(g, s, a, k); break;
    case 2: ShapeKernels<8>::substep(g, s, a, k); break;
    default: ShapeKernels<16>::substep(g, s, a, k); break;
    }
}
static void launch_end(int shape, dim3 g, hipStream_t s, const BatchArgs &a, int k)
{
    switch (shape) {
    case 0: ShapeKernels<2>::end(g, s, a, k); break;
    case 1: ShapeKernels<4>::end(g, s, a, k); break;
    case 2: ShapeKernels<8>::end(g, s, a, k); break;
    default: ShapeKernels<16>::end(g, s, a, k); break;
    }
}

}  // namespace mv2

using namespace mv2;

static mv2_batch_state *batch_of(mv2_context *ctx)
{
    if (!ctx->batch) ctx->batch = new mv2_batch_state();
    return ctx->batch;
}

void mv2_batch_release(mv2_context *ctx)
{
    mv2_batch_state *bs = ctx->batch;
    if (!bs) return;
    for (int p = 0; p < 2; p++)
        if (bs->graph[p]) hipGraphExecDestroy(bs->graph[p]);
    if (bs->capture) hipStreamDestroy(bs->capture);
    if (bs->ev_sub0) hipEventDestroy(bs->ev_sub0);
    if (bs->ev_sub1) hipEventDestroy(bs->ev_sub1);
    for (Scratch *s : {&bs->inst, &bs->tiles_tab, &bs->alt_h, &bs->alt_qx, &bs->alt_qy, &bs->ctl, &bs->wave,
                       &bs->stat, &bs->ledger, &bs->rec, &bs->rtab, &bs->relax_alt, &bs->relax_ctl,
                       &bs->relax_spread, &bs->relax_out, &bs->ctab, &bs->label_counts, &bs->label_stats,
                       &bs->moved, &bs->relax_list, &bs->relax_quiet})
        if (s->ptr) hipFree(s->ptr);
    if (bs->relax_list_host.ptr) hipHostFree(bs->relax_list_host.ptr);
    delete bs;
    ctx->batch = nullptr;
}

static int build_tiles(mv2_context *ctx, mv2_batch_state *bs, int shape)
{
    const int th = kShapes[shape].th, tw = kShapes[shape].tw;
    std::vector<int32_t> first(bs->count), ntiles(bs->count), ti, tr, tc;
    for (int i = 0; i < bs->count; i++) {
        first[i] = (int32_t)ti.size();
        const int tx = (bs->nx[i] + tw - 1) / tw, ty = (bs->ny[i] + th - 1) / th;
        // XCD-aware order inside an instance is unnecessary here: consecutive
        // tiles of one instance land on different XCDs, and the batch is large
        for (int r = 0; r < ty; r++)
            for (int c = 0; c < tx; c++) {
                ti.push_back(i);
                tr.push_back(r * th);
                tc.push_back(c * tw);
            }
        ntiles[i] = ty * tx;
    }
    const int T = (int)ti.size();
    if (int rc = ensure_scratch(ctx, bs->tiles_tab, sizeof(int32_t) * 3 * (size_t)T)) return rc;
    int32_t *tab = (int32_t *)bs->tiles_tab.ptr;
    MV2_HIP(ctx, hipMemcpy(tab, ti.data(), sizeof(int32_t) * T, hipMemcpyHostToDevice));
    MV2_HIP(ctx, hipMemcpy(tab + T, tr.data(), sizeof(int32_t) * T, hipMemcpyHostToDevice));
    MV2_HIP(ctx, hipMemcpy(tab + 2 * (size_t)T, tc.data(), sizeof(int32_t) * T, hipMemcpyHostToDevice));
    int32_t *inst = (int32_t *)bs->inst.ptr;
    MV2_HIP(ctx, hipMemcpy(inst + 2 * (size_t)bs->count + 2 * (size_t)bs->count, first.data(),
                           sizeof(int32_t) * bs->count, hipMemcpyHostToDevice));
    MV2_HIP(ctx, hipMemcpy(inst + 5 * (size_t)bs->count, ntiles.data(), sizeof(int32_t) * bs->count,
                           hipMemcpyHostToDevice));
    bs->tiles = T;
    bs->shape = shape;
    if (int rc = ensure_scratch(ctx, bs->ledger, (sizeof(double) * 2 * kBLedger + sizeof(long long) * 2) * (size_t)T))
        return rc;
    for (int p = 0; p < 2; p++)
        if (bs->graph[p]) {
            hipGraphExecDestroy(bs->graph[p]);
            bs->graph[p] = nullptr;
        }
    return MV2_OK;
}

// relaxation tiles of every instance (44 x 48, row-major per instance)
static int build_relax_tiles(mv2_context *ctx, mv2_batch_state *bs)
{
    std::vector<int32_t> first(bs->count), ntiles(bs->count), ti, tr, tc;
    for (int i = 0; i < bs->count; i++) {
        first[i] = (int32_t)ti.size();
        const int tx = (bs->nx[i] + kBRelaxTW - 1) / kBRelaxTW, ty = (bs->ny[i] + kBRelaxTH - 1) / kBRelaxTH;
        for (int r = 0; r < ty; r++)
            for (int c = 0; c < tx; c++) {
                ti.push_back(i);
                tr.push_back(r * kBRelaxTH);
                tc.push_back(c * kBRelaxTW);
            }
        ntiles[i] = tx * ty;
    }
    const int T = (int)ti.size();
    const size_t n = (size_t)bs->count;
    if (int rc = ensure_scratch(ctx, bs->rtab, sizeof(int32_t) * (2 * n + 3 * (size_t)T))) return rc;
    int32_t *tab = (int32_t *)bs->rtab.ptr;
    MV2_HIP(ctx, hipMemcpy(tab, first.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice));
    MV2_HIP(ctx, hipMemcpy(tab + n, ntiles.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice));
    MV2_HIP(ctx, hipMemcpy(tab + 2 * n, ti.data(), sizeof(int32_t) * T, hipMemcpyHostToDevice));
    MV2_HIP(ctx, hipMemcpy(tab + 2 * n + T, tr.data(), sizeof(int32_t) * T, hipMemcpyHostToDevice));
    MV2_HIP(ctx, hipMemcpy(tab + 2 * n + 2 * (size_t)T, tc.data(), sizeof(int32_t) * T, hipMemcpyHostToDevice));
    bs->rtiles = T;
    bs->relax_predicted = 0;
    bs->rfirst = first;
    bs->rcount = ntiles;
    return MV2_OK;
}

static BatchArgs batch_args(mv2_batch_state *bs, const mv2_step_config *cfg)
{
    BatchArgs a;
    a.count = bs->count;
    a.tiles = bs->tiles;
    a.dx = bs->dx;
    a.dy = bs->dy;
    a.b = bs->b;
    a.h[0] = bs->h;
    a.h[1] = (int32_t *)bs->alt_h.ptr;
    a.qx[0] = bs->qx;
    a.qx[1] = (float *)bs->alt_qx.ptr;
    a.qy[0] = bs->qy;
    a.qy[1] = (float *)bs->alt_qy.ptr;
    int32_t *inst = (int32_t *)bs->inst.ptr;
    const size_t n = (size_t)bs->count;
    a.inst_ny = inst;
    a.inst_nx = inst + n;
    a.inst_off = (const int64_t *)(inst + 2 * n);      // 2n int32 words
    a.inst_first_tile = inst + 4 * n;
    a.inst_tiles = inst + 5 * n;
    const int32_t *tab = (const int32_t *)bs->tiles_tab.ptr;
    a.tile_inst = tab;
    a.tile_r0 = tab + bs->tiles;
    a.tile_c0 = tab + 2 * (size_t)bs->tiles;
    // two copies of the loop state, by step parity
    const int p = bs->frame & 1;
    BControl *ctl0 = (BControl *)bs->ctl.ptr;
    int *flags0 = (int *)((char *)bs->ctl.ptr + sizeof(BControl) * 4 * n);
    unsigned long long *wave0 = (unsigned long long *)bs->wave.ptr;
    unsigned long long *stat0 = (unsigned long long *)bs->stat.ptr;
    const size_t wave_copy = (size_t)kWaveSetsB * kInstWords * n, stat_copy = (size_t)2 * kInstWords * n;
    a.ctl = ctl0 + (size_t)p * 2 * n;
    a.flags = flags0 + (size_t)p * 2 * n;
    a.wave = wave0 + p * wave_copy;
    a.stat = stat0 + p * stat_copy;
    a.ctl_next = ctl0 + (size_t)(p ^ 1) * 2 * n;
    a.flags_next = flags0 + (size_t)(p ^ 1) * 2 * n;
    a.wave_next = wave0 + (p ^ 1) * wave_copy;
    a.stat_next = stat0 + (p ^ 1) * stat_copy;
    a.ledger = (double *)bs->ledger.ptr;
    a.quanta = (long long *)(a.ledger + (size_t)2 * kBLedger * bs->tiles);
    a.rec = (mv2_batch_record *)bs->rec.ptr;
    a.cfg = *cfg;
    a.clamp_quarter = cfg->cfl <= 0.25 ? 1 : 0;
    return a;
}

extern "C" int mv2_batch_bind(mv2_context *ctx, int count, const int32_t *ny, const int32_t *nx,
                              const int64_t *offset, double dx, double dy, float *b, int32_t *h_quanta, float *qx,
                              float *qy)
{
    if (!ctx) return set_error(nullptr, MV2_EINVAL, "[BatchedMobileV2] null context");
    if (count < 1 || !ny || !nx || !offset || !b || !h_quanta || !qx || !qy)
        return set_error(ctx, MV2_EINVAL, "[BatchedMobileV2] invalid batch description");
    if (!(dx > 0.0) || !(dy > 0.0)) return set_error(ctx, MV2_EINVAL, "[BatchedMobileV2] dx, dy must be positive");
    hipSetDevice(ctx->device);
    int64_t cells = 0;
    for (int i = 0; i < count; i++) {
        if (ny[i] < 2 || nx[i] < 2 || (int64_t)ny[i] * nx[i] >= ((int64_t)1 << 31))
            return set_error(ctx, MV2_EINVAL, "[BatchedMobileV2] instance %d has an invalid size %d x %d", i, ny[i], nx[i]);
        if (offset[i] < cells)
            return set_error(ctx, MV2_EINVAL, "[BatchedMobileV2] instance %d overlaps the previous one", i);
        cells = offset[i] + (int64_t)ny[i] * nx[i];
    }
    mv2_batch_state *bs = batch_of(ctx);
    bs->count = count;
    bs->cells = cells;
    bs->dx = dx;
    bs->dy = dy;
    bs->b = b;
    bs->h = h_quanta;
    bs->qx = qx;
    bs->qy = qy;
    bs->ny.assign(ny, ny + count);
    bs->nx.assign(nx, nx + count);
    bs->off.assign(offset, offset + count);
    bs->predicted = 0;
    if (int rc = ensure_scratch(ctx, bs->inst, sizeof(int32_t) * 6 * (size_t)count)) return rc;
    int32_t *inst = (int32_t *)bs->inst.ptr;
    MV2_HIP(ctx, hipMemcpy(inst, ny, sizeof(int32_t) * count, hipMemcpyHostToDevice));
    MV2_HIP(ctx, hipMemcpy(inst + count, nx, sizeof(int32_t) * count, hipMemcpyHostToDevice));
    MV2_HIP(ctx, hipMemcpy(inst + 2 * (size_t)count, offset, sizeof(int64_t) * count, hipMemcpyHostToDevice));
    if (int rc = ensure_scratch(ctx, bs->alt_h, sizeof(int32_t) * (size_t)cells)) return rc;
    if (int rc = ensure_scratch(ctx, bs->alt_qx, sizeof(float) * (size_t)cells)) return rc;
    if (int rc = ensure_scratch(ctx, bs->alt_qy, sizeof(float) * (size_t)cells)) return rc;
    if (int rc = ensure_scratch(ctx, bs->ctl, (sizeof(BControl) + sizeof(int)) * 4 * (size_t)count)) return rc;
    if (int rc = ensure_scratch(ctx, bs->wave, sizeof(unsigned long long) * 2 * kWaveSetsB * kInstWords * (size_t)count))
        return rc;
    if (int rc = ensure_scratch(ctx, bs->stat, sizeof(unsigned long long) * 2 * 2 * kInstWords * (size_t)count))
        return rc;
    bs->clean = -1;                      // new scratch: reset before the next step
    bs->ctiles = 0;                      // labelling tiles of the new layout: built on first use
    if (int rc = ensure_scratch(ctx, bs->rec, sizeof(mv2_batch_record) * (size_t)count)) return rc;
    if (int rc = build_relax_tiles(ctx, bs)) return rc;
    return build_tiles(ctx, bs, bs->shape);
}

extern "C" int mv2_batch_set_tile_shape(mv2_context *ctx, int shape)
{
    if (!ctx || !ctx->batch || !ctx->batch->count)
        return set_error(ctx, MV2_ESTATE, "[BatchedMobileV2] bind a batch first");
    if (shape < 0 || shape >= kShapeCount) return set_error(ctx, MV2_EINVAL, "[BatchedMobileV2] unknown tile shape %d", shape);
    hipSetDevice(ctx->device);
    if (shape == ctx->batch->shape && ctx->batch->tiles) return MV2_OK;
    return build_tiles(ctx, ctx->batch, shape);
}

extern "C" int mv2_batch_tile_shape(int shape, int *th, int *tw)
{
    if (shape < 0 || shape >= kShapeCount || !th || !tw) return MV2_EINVAL;
    *th = kShapes[shape].th;
    *tw = kShapes[shape].tw;
    return MV2_OK;
}

// capture eight substep launches (k = 0..7) into a graph for this config
static int ensure_graph(mv2_context *ctx, mv2_batch_state *bs, const BatchArgs &a)
{
    const int p = bs->frame & 1;
    std::vector<unsigned char> key(sizeof(BatchArgs) + sizeof(int));
    memcpy(key.data(), &a, sizeof(BatchArgs));
    memcpy(key.data() + sizeof(BatchArgs), &bs->shape, sizeof(int));
    if (bs->graph[p] && key == bs->graph_key[p]) return MV2_OK;
    if (bs->graph[p]) {
        hipGraphExecDestroy(bs->graph[p]);
        bs->graph[p] = nullptr;
    }
    if (!bs->capture) MV2_HIP(ctx, hipStreamCreateWithFlags(&bs->capture, hipStreamNonBlocking));
    MV2_HIP(ctx, hipStreamBeginCapture(bs->capture, hipStreamCaptureModeThreadLocal));
    for (int k = 0; k < 8; k++) launch_substep(bs->shape, dim3(bs->tiles), bs->capture, a, k);
    hipGraph_t graph = nullptr;
    MV2_HIP(ctx, hipStreamEndCapture(bs->capture, &graph));
    hipError_t e = hipGraphInstantiate(&bs->graph[p], graph, nullptr, nullptr, 0);
    hipGraphDestroy(graph);
    if (e != hipSuccess) {
        bs->graph[p] = nullptr;
        return check_hip(ctx, e, "hipGraphInstantiate");
    }
    bs->graph_key[p] = key;
    return MV2_OK;
}

extern "C" int mv2_batch_step(mv2_context *ctx, const mv2_step_config *cfg, int use_graph, mv2_batch_record *out,
                              double *timing_ms)
{
    if (!ctx) return set_error(nullptr, MV2_EINVAL, "[BatchedMobileV2] null context");
    mv2_batch_state *bs = ctx->batch;
    if (!bs || !bs->count) return set_error(ctx, MV2_ESTATE, "[BatchedMobileV2] bind a batch first");
    if (!cfg || !out) return set_error(ctx, MV2_EINVAL, "[BatchedMobileV2] null config/record");
    if (!(cfg->dt > 0.0) || !std::isfinite(cfg->dt)) return set_error(ctx, MV2_EINVAL, "dt must be finite/positive");
    if (!(cfg->cfl > 0.0) || !(cfg->earth_pressure > 0.0) || !(cfg->gravity > 0.0) || !(cfg->dry_tolerance > 0.0) ||
        !(cfg->basal_friction > 0.0))
        return set_error(ctx, MV2_EINVAL, "MobileV2 configuration must be finite/positive");
    hipSetDevice(ctx->device);
    hipStream_t st = ctx->stream;
    const BatchArgs a = batch_args(bs, cfg);
    if (use_graph)
        if (int rc = ensure_graph(ctx, bs, a)) return rc;
    const dim3 grid(bs->tiles);
    MV2_HIP(ctx, hipEventRecord(ctx->ev_start, st));
    // this parity's loop state is zero if the previous step completed; an
    // error anywhere below leaves clean at -1 and the next step resets
    const int parity = bs->frame & 1;
    if (bs->clean != parity)
        hipLaunchKernelGGL(batch_reset_kernel, dim3(bs->count < 1024 ? bs->count : 1024), dim3(kWave), 0, st, a);
    bs->clean = -1;
    launch_begin(bs->shape, grid, st, a);
    MV2_HIP(ctx, hipGetLastError());
    int k = 0;
    int want = bs->predicted > 0 ? bs->predicted + 1 : 8;
    const size_t rec_bytes = sizeof(mv2_batch_record) * (size_t)bs->count;
    if (int rc = ensure_scratch(ctx, ctx->host_result, rec_bytes + 4096, true)) return rc;
    const mv2_batch_record *hrec = (const mv2_batch_record *)ctx->host_result.ptr;
    if (!bs->ev_sub0) {
        MV2_HIP(ctx, hipEventCreate(&bs->ev_sub0));
        MV2_HIP(ctx, hipEventCreate(&bs->ev_sub1));
    }
    double sub_ms = 0.0;
    for (;;) {
        MV2_HIP(ctx, hipEventRecord(bs->ev_sub0, st));
        // whole groups of eight from the graph when the phase allows, the rest
        // as single launches
        for (int left = want; left > 0;) {
            if (use_graph && k % 8 == 0 && left >= 8) {
                MV2_HIP(ctx, hipGraphLaunch(bs->graph[parity], st));
                k += 8;
                left -= 8;
            } else {
                launch_substep(bs->shape, grid, st, a, k % 8);
                k++;
                left--;
            }
        }
        MV2_HIP(ctx, hipEventRecord(bs->ev_sub1, st));
        launch_end(bs->shape, grid, st, a, k % 8);
        hipLaunchKernelGGL(batch_finalize_kernel, dim3(bs->count), dim3(kBlock), 0, st, a, k % 8);
        MV2_HIP(ctx, hipGetLastError());
        MV2_HIP(ctx, hipEventRecord(ctx->ev_stop, st));
        MV2_HIP(ctx, hipMemcpyAsync(ctx->host_result.ptr, a.rec, rec_bytes, hipMemcpyDeviceToHost, st));
        if (int rc_w = stream_wait(ctx, st)) return rc_w;
        float sms = 0.0f;
        MV2_HIP(ctx, hipEventElapsedTime(&sms, bs->ev_sub0, bs->ev_sub1));
        sub_ms += sms;
        bool all = true;
        int most = 0;
        for (int i = 0; i < bs->count; i++) {
            all = all && hrec[i].complete;
            most = hrec[i].substeps > most ? hrec[i].substeps : most;
        }
        if (all) {
            bs->predicted = most;
            break;
        }
        // more substeps than enqueued: the states so far are valid, continue
        want = want * 2 > 8 ? want * 2 : 8;
        (void)most;
    }
    memcpy(out, hrec, rec_bytes);
    bs->clean = parity ^ 1;              // zeroed by this step's finalize pass
    bs->frame++;
    if (timing_ms) {
        float ms = 0.0f;
        MV2_HIP(ctx, hipEventElapsedTime(&ms, ctx->ev_start, ctx->ev_stop));
        timing_ms[0] = ms;          // the whole call on the device
        timing_ms[1] = sub_ms;      // the substep launches (including those after every loop ended)
        timing_ms[2] = (double)k;   // substep launches
    }
    for (int i = 0; i < bs->count; i++)
        if (out[i].numeric_error)
            return set_error(ctx, MV2_ENUMERIC, "[BatchedMobileV2] instance %d: %s", i,
                             out[i].numeric_error == 1 ? "MobileV2 positivity failure"
                             : out[i].numeric_error == 2 ? "MobileV2 maximum_substeps exceeded"
                                                         : "mobile thickness exceeds the FP32 storage range");
    return MV2_OK;
}

extern "C" int mv2_batch_relax(mv2_context *ctx, float *field, const double crit_diff[4], int max_iterations,
                               double tolerance, mv2_relax_result *out)
{
    if (!ctx) return set_error(nullptr, MV2_EINVAL, "[BatchedMiniSlope] null context");
    mv2_batch_state *bs = ctx->batch;
    if (!bs || !bs->count) return set_error(ctx, MV2_ESTATE, "[BatchedMiniSlope] bind a batch first");
    if (!field || !crit_diff || !out) return set_error(ctx, MV2_EINVAL, "[BatchedMiniSlope] null argument");
    if (max_iterations < 1) return set_error(ctx, MV2_EINVAL, "[BatchedMiniSlope] max_iterations must be >= 1");
    hipSetDevice(ctx->device);
    const size_t n = (size_t)bs->count;
    if (int rc = ensure_scratch(ctx, bs->relax_alt, sizeof(float) * (size_t)bs->cells)) return rc;
    if (int rc = ensure_scratch(ctx, bs->relax_ctl, sizeof(BRelaxControl) * 2 * n)) return rc;
    if (int rc = ensure_scratch(ctx, bs->relax_spread, sizeof(unsigned long long) * 3 * kInstWords * n)) return rc;
    if (int rc = ensure_scratch(ctx, bs->relax_out, sizeof(mv2_relax_result) * n)) return rc;
    BatchRelaxArgs a;
    a.count = bs->count;
    a.tiles = bs->rtiles;
    a.h[0] = field;
    a.h[1] = (float *)bs->relax_alt.ptr;
    const int32_t *inst = (const int32_t *)bs->inst.ptr;
    a.inst_ny = inst;
    a.inst_nx = inst + n;
    a.inst_off = (const int64_t *)(inst + 2 * n);
    const int32_t *tab = (const int32_t *)bs->rtab.ptr;
    const size_t T = (size_t)bs->rtiles;
    a.inst_first = tab;
    a.inst_tiles = tab + n;
    a.tile_inst = tab + 2 * n;
    a.tile_r0 = tab + 2 * n + T;
    a.tile_c0 = tab + 2 * n + 2 * T;
    for (int d = 0; d < 4; d++) a.crit[d] = (float)crit_diff[d];
    a.max_iterations = max_iterations;
    a.tolerance = tolerance;
    a.ctl = (BRelaxControl *)bs->relax_ctl.ptr;
    a.spread = (unsigned long long *)bs->relax_spread.ptr;
    a.out = (mv2_relax_result *)bs->relax_out.ptr;
    a.tile_list = nullptr;
    if (int rc = ensure_scratch(ctx, bs->relax_quiet, (2 * sizeof(int32_t) + sizeof(double)) * T)) return rc;
    a.texcess = (double *)bs->relax_quiet.ptr;
    a.tchanged = (int32_t *)(a.texcess + T);
    if (int rc = ensure_scratch(ctx, bs->relax_list, sizeof(int32_t) * T)) return rc;
    if (int rc = ensure_scratch(ctx, bs->relax_list_host, sizeof(int32_t) * T, true)) return rc;
    int grid = bs->rtiles;
    hipStream_t st = ctx->stream;
    hipLaunchKernelGGL(batch_relax_reset_kernel, dim3(bs->count < 1024 ? bs->count : 1024), dim3(kWave), 0, st, a);
    MV2_HIP(ctx, hipGetLastError());
    const size_t out_bytes = sizeof(mv2_relax_result) * n;
    if (int rc = ensure_scratch(ctx, ctx->host_result, out_bytes + 4096, true)) return rc;
    const mv2_relax_result *hres = (const mv2_relax_result *)ctx->host_result.ptr;
    // as the single-field relaxation: the previous call's count + 1 (2..32),
    // then geometric growth while some instance has not finished
    int chunk = bs->relax_predicted + 1;
    if (chunk < 2) chunk = 2;
    if (chunk > 32) chunk = 32;
    int k = 0;
    for (;;) {
        if (chunk > max_iterations - k) chunk = max_iterations - k;
        for (int j = 0; j < chunk; j++, k++)
            hipLaunchKernelGGL(batch_relax_kernel, dim3(grid), dim3(BRelaxGeom::BLOCK), 0, st, a, k);
        hipLaunchKernelGGL(batch_relax_end_kernel, dim3(grid), dim3(kBlock), 0, st, a, k);
        MV2_HIP(ctx, hipGetLastError());
        MV2_HIP(ctx, hipMemcpyAsync(ctx->host_result.ptr, a.out, out_bytes, hipMemcpyDeviceToHost, st));
        if (int rc_w = stream_wait(ctx, st)) return rc_w;
        bool all = true;
        int most = 0;
        for (int i = 0; i < bs->count; i++) {
            all = all && hres[i].complete;
            most = hres[i].iterations > most ? hres[i].iterations : most;
        }
        if (all || k >= max_iterations) {
            bs->relax_predicted = most;
            break;
        }
        chunk = k / 2 > 8 ? k / 2 : 8;
        // later launches run only the tiles of the instances still relaxing: a
        // finished instance's field, result and loop state were settled by the
        // end pass above and nothing reads them again in this call.  The
        // staging buffer is rewritten only after the next synchronisation.
        int32_t *list = (int32_t *)bs->relax_list_host.ptr;
        int live = 0;
        for (int i = 0; i < bs->count; i++)
            if (!hres[i].complete)
                for (int t = 0; t < bs->rcount[i]; t++) list[live++] = bs->rfirst[i] + t;
        if (live < grid) {
            MV2_HIP(ctx, hipMemcpyAsync(bs->relax_list.ptr, list, sizeof(int32_t) * (size_t)live,
                                        hipMemcpyHostToDevice, st));
            a.tile_list = (const int32_t *)bs->relax_list.ptr;
            grid = live;
        }
    }
    memcpy(out, hres, out_bytes);
    return MV2_OK;
}
