// Mobile -> Resting deposition: one WarpDepositionOperator.step_resident call
// (warp_deposition.py:23-115 deposit, :117-129 weighted_sum, :180-235).
//
// A cell with mobile material, no tool forcing, speed |q|/h at or below the
// threshold and a free surface b_eff + mobile inside the cohesive stop yield
// moves min(h, rate dt) of mobile into b_eff (all of h when the caller
// settles a sub-cell tail), removes the same fraction of its momentum and
// reports it.  Every cell is judged on the field as it was before the call:
// the first kernel reads the pre-call fields and writes only the per-cell
// outputs (deposited, removed momentum) and per-workgroup sums; the second
// applies the amounts in place.  The reference's single Warp kernel reads
// neighbours that other threads may already have changed; the free surface
// each of them sees differs from the pre-call one by rounding only, so the
// two-pass form is the same operator made reproducible.
//
// FP32 storage: the decision is evaluated in double on the widened values;
// amounts are whole quanta of the mobile storage, min(h, floor(rate dt / q)).
#include "mv2_internal.hpp"

#include <cmath>
#include <cstring>

namespace mv2 {

constexpr int kDepMaxGrid = 1 << 16;

struct DepArgs {
    void *b_eff, *mobile, *qx, *qy;
    const int32_t *forcing;       // may be null (no tool forcing)
    void *deposited, *removed_x, *removed_y;
    double *cumulative;           // avalanche_m2r_cumulative
    double *partials;             // [grid][3] sums of the first pass
    double *result;               // [3] folded sums
    int64_t n;
    int ny, nx, settle, weight_mode;
    int cap_q;                    // FP32: floor(rate dt / quantum), clamped to int32
    double dx, dy, rate_dt, speed_threshold, density, rho_g, cohesion, tan_stop, area, f6;
};

__device__ inline double dep_min(double a, double b) { return a <= b ? a : b; }

// warp_deposition.py:69-102 on pre-call fields: 1 when the slow cell is
// inside the cohesive stop yield of its free surface.  Differences are taken
// in the reference's order, ((b_r + m_r) - b_l) - m_l.
template <bool F32>
__device__ inline bool dep_settles(const DepArgs &a, int r, int c, int64_t i, double h, double qx, double qy)
{
    const double vx = qx / h, vy = qy / h;
    if (sqrt(vx * vx + vy * vy) > a.speed_threshold) return false;
    const int64_t left = c > 0 ? i - 1 : i, right = c + 1 < a.nx ? i + 1 : i;
    const int64_t down = r > 0 ? i - a.nx : i, up = r + 1 < a.ny ? i + a.nx : i;
    const double hx = (c == 0 || c + 1 == a.nx) ? a.dx : 2.0 * a.dx;
    const double hy = (r == 0 || r + 1 == a.ny) ? a.dy : 2.0 * a.dy;
    double bl, br, bd, bu, ml, mr, md, mu;
    if (!F32) {
        const double *b = (const double *)a.b_eff, *m = (const double *)a.mobile;
        bl = b[left]; br = b[right]; bd = b[down]; bu = b[up];
        ml = m[left]; mr = m[right]; md = m[down]; mu = m[up];
    } else {
        const float *b = (const float *)a.b_eff;
        const int32_t *m = (const int32_t *)a.mobile;
        bl = b[left]; br = b[right]; bd = b[down]; bu = b[up];
        ml = (double)m[left] * kQuantum; mr = (double)m[right] * kQuantum;
        md = (double)m[down] * kQuantum; mu = (double)m[up] * kQuantum;
    }
    const double sx = (br + mr - bl - ml) / hx;
    const double sy = (bu + mu - bd - md) / hy;
    const double slope_tan = sqrt(sx * sx + sy * sy);
    const double normalization = sqrt(1.0 + slope_tan * slope_tan);
    const double drive = a.rho_g * h * slope_tan / normalization;
    const double resistance = a.cohesion + a.rho_g * h * a.tan_stop / normalization;
    return !(drive > resistance);
}

// pass 1: decisions and per-cell outputs from the pre-call fields, sums per
// workgroup (deposited w, density rx w, density ry w)
template <bool F32>
__global__ void __launch_bounds__(kBlock) deposition_decide_kernel(DepArgs a)
{
    __shared__ double smem[3 * (kBlock / kWave)];
    double v[3] = {0.0, 0.0, 0.0};
    const int64_t stride = (int64_t)gridDim.x * kBlock;
    for (int64_t i = (int64_t)xcd_block() * kBlock + threadIdx.x; i < a.n; i += stride) {
        const int r = (int)((unsigned)i / (unsigned)a.nx), c = (int)i - r * a.nx;
        const bool forced = a.forcing && a.forcing[i] != 0;
        if (!F32) {
            const double h = ((const double *)a.mobile)[i];
            double dep = 0.0, rx = 0.0, ry = 0.0;
            if (h > 0.0 && !forced) {
                const double qx = ((const double *)a.qx)[i], qy = ((const double *)a.qy)[i];
                if (a.settle) {
                    dep = h; rx = qx; ry = qy;
                } else if (dep_settles<false>(a, r, c, i, h, qx, qy)) {
                    dep = dep_min(h, a.rate_dt);
                    const double fraction_removed = dep / h;
                    rx = qx * fraction_removed;
                    ry = qy * fraction_removed;
                }
            }
            ((double *)a.deposited)[i] = dep;
            ((double *)a.removed_x)[i] = rx;
            ((double *)a.removed_y)[i] = ry;
            const double w = cell_weight(a.weight_mode, a.ny, a.nx, a.area, a.f6, r, c);
            v[0] += dep * w;
            v[1] += a.density * rx * w;
            v[2] += a.density * ry * w;
        } else {
            const int32_t hq = ((const int32_t *)a.mobile)[i];
            int32_t aq = 0;
            float rx = 0.0f, ry = 0.0f;
            if (hq > 0 && !forced) {
                const float qx = ((const float *)a.qx)[i], qy = ((const float *)a.qy)[i];
                const double h = (double)hq * kQuantum;
                if (a.settle) {
                    aq = hq; rx = qx; ry = qy;
                } else if (dep_settles<true>(a, r, c, i, h, (double)qx, (double)qy)) {
                    aq = hq < a.cap_q ? hq : a.cap_q;
                    if (aq > 0) {
                        const double fraction_removed = (double)aq / (double)hq;
                        rx = (float)((double)qx * fraction_removed);
                        ry = (float)((double)qy * fraction_removed);
                    }
                }
            }
            ((int32_t *)a.deposited)[i] = aq;
            ((float *)a.removed_x)[i] = rx;
            ((float *)a.removed_y)[i] = ry;
            const double w = cell_weight(a.weight_mode, a.ny, a.nx, a.area, a.f6, r, c);
            v[0] += (double)aq * kQuantum * w;
            v[1] += a.density * (double)rx * w;
            v[2] += a.density * (double)ry * w;
        }
    }
    block_reduce<double, 3, 0>(v, smem);
    if (threadIdx.x == 0) {
        double *p = a.partials + (size_t)blockIdx.x * 3;
        p[0] = v[0]; p[1] = v[1]; p[2] = v[2];
    }
}

// pass 2: apply the decided amounts in place; workgroup 0 also folds the
// first pass's sums in workgroup order (complete: the first pass ran before
// this launch on the same stream)
template <bool F32>
__global__ void __launch_bounds__(kBlock) deposition_apply_kernel(DepArgs a, int nparts)
{
    const int64_t stride = (int64_t)gridDim.x * kBlock;
    for (int64_t i = (int64_t)xcd_block() * kBlock + threadIdx.x; i < a.n; i += stride) {
        const int r = (int)((unsigned)i / (unsigned)a.nx), c = (int)i - r * a.nx;
        if (!F32) {
            const double dep = ((const double *)a.deposited)[i];
            if (!(dep > 0.0)) continue;
            double *b = (double *)a.b_eff, *m = (double *)a.mobile;
            double *qx = (double *)a.qx, *qy = (double *)a.qy;
            const double h = m[i];
            const double w = cell_weight(a.weight_mode, a.ny, a.nx, a.area, a.f6, r, c);
            if (a.settle) {
                b[i] = b[i] + h;
                m[i] = 0.0;
                qx[i] = 0.0;
                qy[i] = 0.0;
                a.cumulative[i] = a.cumulative[i] + h * w;
            } else {
                b[i] = b[i] + dep;
                m[i] = h - dep;
                const double fraction = (h - dep) / h;
                qx[i] = qx[i] * fraction;
                qy[i] = qy[i] * fraction;
                a.cumulative[i] = a.cumulative[i] + dep * w;
            }
        } else {
            const int32_t aq = ((const int32_t *)a.deposited)[i];
            if (aq <= 0) continue;
            float *b = (float *)a.b_eff, *qx = (float *)a.qx, *qy = (float *)a.qy;
            int32_t *m = (int32_t *)a.mobile;
            const int32_t hq = m[i];
            const double amount = (double)aq * kQuantum;
            const double w = cell_weight(a.weight_mode, a.ny, a.nx, a.area, a.f6, r, c);
            b[i] = (float)((double)b[i] + amount);
            m[i] = hq - aq;
            if (a.settle) {
                qx[i] = 0.0f;
                qy[i] = 0.0f;
            } else {
                const double fraction = (double)(hq - aq) / (double)hq;
                qx[i] = (float)((double)qx[i] * fraction);
                qy[i] = (float)((double)qy[i] * fraction);
            }
            a.cumulative[i] = a.cumulative[i] + amount * w;
        }
    }
    if (blockIdx.x == 0) {
        __shared__ double smem[3 * (kBlock / kWave)];
        double v[3] = {0.0, 0.0, 0.0};
        for (int p = threadIdx.x; p < nparts; p += kBlock)
            for (int k = 0; k < 3; k++) v[k] += a.partials[(size_t)p * 3 + k];
        block_reduce<double, 3, 0>(v, smem);
        if (threadIdx.x == 0)
            for (int k = 0; k < 3; k++) a.result[k] = v[k];
    }
}

static bool finite_nonneg(double v) { return std::isfinite(v) && v >= 0.0; }

}  // namespace mv2

using namespace mv2;

extern "C" int mv2_deposition_step(mv2_context *ctx, const mv2_deposition_config *cfg,
                                   const int32_t *forcing, void *deposited, void *removed_x,
                                   void *removed_y, double *m2r_cumulative, double out[3])
{
    if (!ctx) return set_error(nullptr, MV2_EINVAL, "[WarpDeposition] null context");
    if (!ctx->bound) return set_error(ctx, MV2_ESTATE, "[WarpDeposition] bind DeviceBulkState first");
    if (!cfg || !out) return set_error(ctx, MV2_EINVAL, "[WarpDeposition] null config/result");
    state_written(ctx);
    if (!deposited || !removed_x || !removed_y || !m2r_cumulative)
        return set_error(ctx, MV2_EINVAL,
                         "[WarpDeposition] missing shared fields: deposition_work, "
                         "deposition_removed_momentum_x/y, avalanche_m2r_cumulative");
    if (!std::isfinite(cfg->dt) || !(cfg->dt > 0.0))
        return set_error(ctx, MV2_EINVAL, "[WarpDeposition] dt_s must be finite/positive");
    if (!std::isfinite(cfg->settling_rate) || !(cfg->settling_rate > 0.0) ||
        !finite_nonneg(cfg->speed_threshold))
        return set_error(ctx, MV2_EINVAL, "[WarpDeposition] invalid configuration");
    if (!std::isfinite(cfg->density) || !(cfg->density > 0.0) || !finite_nonneg(cfg->gravity) ||
        !finite_nonneg(cfg->cohesion) || !finite_nonneg(cfg->tan_stop))
        return set_error(ctx, MV2_EINVAL, "[WarpDeposition] invalid material parameters");
    hipSetDevice(ctx->device);
    out[0] = out[1] = out[2] = 0.0;
    DepArgs a{};
    a.b_eff = ctx->b_eff;
    a.mobile = ctx->mobile;
    a.qx = ctx->qx;
    a.qy = ctx->qy;
    a.forcing = forcing;
    a.deposited = deposited;
    a.removed_x = removed_x;
    a.removed_y = removed_y;
    a.cumulative = m2r_cumulative;
    a.n = (int64_t)ctx->ny * ctx->nx;
    a.ny = ctx->ny;
    a.nx = ctx->nx;
    a.settle = cfg->settle_subcell_tail ? 1 : 0;
    a.weight_mode = cfg->weight_mode;
    a.dx = ctx->dx;
    a.dy = ctx->dy;
    a.rate_dt = cfg->settling_rate * cfg->dt;
    const double cap = floor(a.rate_dt * kInvQuantum);
    a.cap_q = cap >= 2147483647.0 ? 2147483647 : (int)cap;
    a.speed_threshold = cfg->speed_threshold;
    a.density = cfg->density;
    a.rho_g = cfg->density * cfg->gravity;
    a.cohesion = cfg->cohesion;
    a.tan_stop = cfg->tan_stop;
    a.area = ctx->dx * ctx->dy;
    a.f6 = a.area / 6.0;
    const int grid = grid_for(a.n, kBlock, kDepMaxGrid);
    if (int rc = ensure_scratch(ctx, ctx->partials, sizeof(double) * 3 * (size_t)grid)) return rc;
    a.partials = (double *)ctx->partials.ptr;
    a.result = (double *)ctx->result.ptr;
    if (ctx->precision == MV2_FP64) {
        deposition_decide_kernel<false><<<grid, kBlock, 0, ctx->stream>>>(a);
        MV2_HIP(ctx, hipGetLastError());
        deposition_apply_kernel<false><<<grid, kBlock, 0, ctx->stream>>>(a, grid);
    } else {
        deposition_decide_kernel<true><<<grid, kBlock, 0, ctx->stream>>>(a);
        MV2_HIP(ctx, hipGetLastError());
        deposition_apply_kernel<true><<<grid, kBlock, 0, ctx->stream>>>(a, grid);
    }
    MV2_HIP(ctx, hipGetLastError());
    MV2_HIP(ctx, hipMemcpyAsync(ctx->host_result.ptr, a.result, sizeof(double) * 3, hipMemcpyDeviceToHost,
                                ctx->stream));
    if (int rc_w = stream_wait(ctx, ctx->stream)) return rc_w;
    memcpy(out, ctx->host_result.ptr, sizeof(double) * 3);
    return MV2_OK;
}
