// Resting <-> Mobile exchange contracts and reservoir ledger.
//
// entrain/deposit restate bulk_state_authority.py:104-150 (donor-limited,
// H_free = b_eff + mobile invariant per cell, entrainment never writes
// momentum, deposition reports the momentum it removes); reservoirs restates
// sum_reservoirs (:60-75).  Sums are formed per workgroup and combined in a
// fixed order by a second single-workgroup kernel, so every ledger value is
// reproducible run to run (the reference uses FP64 atomics).
//
// FP32 storage: b_eff/z_base/momentum are float, mobile is int32 quanta.  A
// requested depth is rounded down to the largest amount that is whole quanta
// and leaves b_eff exactly on the float32 grid (entrain_quanta_f32 /
// deposit_quanta_f32), so the donor limit holds, H_free is bit-invariant per
// cell and the ledger reports exactly what both reservoirs changed by.
#include "mv2_internal.hpp"

#include <cmath>
#include <cstring>

namespace mv2 {

constexpr int kExGrid = 256;

}  // namespace mv2

// Workgroup partials of the allocation-free ledger (mv2_reservoirs_noalloc):
// kExGrid partials of 4 values, every grid size folds into at most kExGrid
// workgroups.  Each GPU holds its own copy for the life of the process; host
// code never uses this name as an address -- mv2_create resolves the device
// address once per context (noalloc_scratch_address) and the entry point
// passes that pointer to the kernels.
static __device__ __align__(256) double scratch[4 * mv2::kExGrid];

namespace mv2 {

int noalloc_scratch_address(mv2_context *ctx, double **out)
{
    void *p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
        *out = nullptr;
        return check_hip(ctx, e, "hipGetSymbolAddress(scratch)");
    }
    *out = static_cast<double *>(p);
    return MV2_OK;
}

__device__ inline double dmaxd(double a, double b) { return a >= b ? a : b; }
__device__ inline double dmind(double a, double b) { return a <= b ? a : b; }

template <int NV>
__global__ void __launch_bounds__(kBlock) reduce_partials(const double *partials, int nparts, double *out)
{
    __shared__ double smem[NV * 4];
    double v[NV];
    for (int k = 0; k < NV; k++) v[k] = 0.0;
    for (int p = threadIdx.x; p < nparts; p += blockDim.x)
        for (int k = 0; k < NV; k++) v[k] += partials[(size_t)p * NV + k];
    block_reduce<double, NV, 0>(v, smem);
    if (threadIdx.x == 0)
        for (int k = 0; k < NV; k++) out[k] = v[k];
}

template <bool F32>
__global__ void __launch_bounds__(kBlock) entrain_kernel(const void *z_base, void *b_eff, void *mobile,
                                                         const int32_t *idx, const double *req, int n,
                                                         int ny, int nx, double area, double f6,
                                                         int weight_mode, double *partials)
{
    __shared__ double smem[4];
    double v[1] = {0.0};
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
        const int i = idx[k];
        const int r = i / nx, c = i - r * nx;
        const double w = cell_weight(weight_mode, ny, nx, area, f6, r, c);
        if (!F32) {
            const double *z = (const double *)z_base;
            double *b = (double *)b_eff;
            double *m = (double *)mobile;
            double available = dmaxd(b[i] - z[i], 0.0);
            double depth = dmind(dmaxd(req[k], 0.0), available);
            b[i] = b[i] - depth;
            m[i] = m[i] + depth;
            v[0] += depth * w;
        } else {
            const float *z = (const float *)z_base;
            float *b = (float *)b_eff;
            int32_t *m = (int32_t *)mobile;
            double available = dmaxd((double)b[i] - (double)z[i], 0.0);
            double depth = dmind(dmaxd(req[k], 0.0), available);
            const long long dq = entrain_quanta_f32(b[i], depth, 2147483647LL - (long long)m[i]);
            double moved = (double)dq * kQuantum;
            b[i] = (float)((double)b[i] - moved);            // exact
            m[i] = (int32_t)((long long)m[i] + dq);
            v[0] += moved * w;
        }
    }
    block_reduce<double, 1, 0>(v, smem);
    if (threadIdx.x == 0) partials[blockIdx.x] = v[0];
}

template <bool F32>
__global__ void __launch_bounds__(kBlock) deposit_kernel(void *b_eff, void *mobile, void *qx_, void *qy_,
                                                         const int32_t *idx, const double *req, int n,
                                                         int ny, int nx, double area, double f6,
                                                         int weight_mode, double density, double *partials)
{
    __shared__ double smem[12];
    double v[3] = {0.0, 0.0, 0.0};
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
        const int i = idx[k];
        const int r = i / nx, c = i - r * nx;
        const double w = cell_weight(weight_mode, ny, nx, area, f6, r, c);
        if (!F32) {
            double *b = (double *)b_eff;
            double *m = (double *)mobile;
            double *qx = (double *)qx_;
            double *qy = (double *)qy_;
            double before = dmaxd(m[i], 0.0);
            double depth = dmind(dmaxd(req[k], 0.0), before);
            double fraction = before > 0.0 ? depth / before : 0.0;
            double rx = qx[i] * fraction, ry = qy[i] * fraction;
            b[i] = b[i] + depth;
            m[i] = before - depth;
            qx[i] = qx[i] - rx;
            qy[i] = qy[i] - ry;
            v[0] += depth * w;
            v[1] += density * rx * w;
            v[2] += density * ry * w;
        } else {
            float *b = (float *)b_eff;
            int32_t *m = (int32_t *)mobile;
            float *qx = (float *)qx_;
            float *qy = (float *)qy_;
            long long before = m[i] > 0 ? m[i] : 0;
            long long dq = (long long)floor(dmaxd(req[k], 0.0) * kInvQuantum);
            if (dq > before) dq = before;
            float b_new;
            dq = deposit_quanta_f32(b[i], dq, &b_new);
            float fraction = before > 0 ? (float)((double)dq / (double)before) : 0.0f;
            float rx = qx[i] * fraction, ry = qy[i] * fraction;
            double moved = (double)dq * kQuantum;
            b[i] = b_new;
            m[i] = (int32_t)(before - dq);
            qx[i] = qx[i] - rx;
            qy[i] = qy[i] - ry;
            v[0] += moved * w;
            v[1] += density * (double)rx * w;
            v[2] += density * (double)ry * w;
        }
    }
    block_reduce<double, 3, 0>(v, smem);
    if (threadIdx.x == 0)
        for (int k = 0; k < 3; k++) partials[(size_t)blockIdx.x * 3 + k] = v[k];
}

template <bool F32>
__global__ void __launch_bounds__(kBlock) reservoir_kernel(const void *z_base, const void *b_eff,
                                                           const void *mobile, const void *qx_,
                                                           const void *qy_, int ny, int nx, double area,
                                                           double f6, int weight_mode, double density,
                                                           double *partials)
{
    __shared__ double smem[16];
    double v[4] = {0.0, 0.0, 0.0, 0.0};
    const int64_t n = (int64_t)ny * nx;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int r = (int)i / nx, c = (int)i - r * nx;
        const double w = cell_weight(weight_mode, ny, nx, area, f6, r, c);
        double z, b, m, qx, qy;
        if (!F32) {
            z = ((const double *)z_base)[i];
            b = ((const double *)b_eff)[i];
            m = ((const double *)mobile)[i];
            qx = ((const double *)qx_)[i];
            qy = ((const double *)qy_)[i];
        } else {
            z = ((const float *)z_base)[i];
            b = ((const float *)b_eff)[i];
            m = (double)((const int32_t *)mobile)[i] * kQuantum;
            qx = ((const float *)qx_)[i];
            qy = ((const float *)qy_)[i];
        }
        v[0] += dmaxd(b - z, 0.0) * w;
        v[1] += m * w;
        v[2] += density * qx * w;
        v[3] += density * qy * w;
    }
    block_reduce<double, 4, 0>(v, smem);
    if (threadIdx.x == 0)
        for (int k = 0; k < 4; k++) partials[(size_t)blockIdx.x * 4 + k] = v[k];
}

static int finish(mv2_context *ctx, int nv, int nparts, double *out)
{
    double *partials = (double *)ctx->partials.ptr;
    double *dres = (double *)ctx->result.ptr;
    switch (nv) {
    case 1: hipLaunchKernelGGL(reduce_partials<1>, dim3(1), dim3(kBlock), 0, ctx->stream, partials, nparts, dres); break;
    case 3: hipLaunchKernelGGL(reduce_partials<3>, dim3(1), dim3(kBlock), 0, ctx->stream, partials, nparts, dres); break;
    default: hipLaunchKernelGGL(reduce_partials<4>, dim3(1), dim3(kBlock), 0, ctx->stream, partials, nparts, dres); break;
    }
    MV2_HIP(ctx, hipGetLastError());
    MV2_HIP(ctx, hipMemcpyAsync(ctx->host_result.ptr, dres, sizeof(double) * nv, hipMemcpyDeviceToHost, ctx->stream));
    if (int rc_w = stream_wait(ctx, ctx->stream)) return rc_w;
    memcpy(out, ctx->host_result.ptr, sizeof(double) * nv);
    return MV2_OK;
}

static int check_bound(mv2_context *ctx, const char *who)
{
    if (!ctx) return set_error(nullptr, MV2_EINVAL, "%s null context", who);
    if (!ctx->bound) return set_error(ctx, MV2_ESTATE, "%s bind state first", who);
    hipSetDevice(ctx->device);
    return MV2_OK;
}

}  // namespace mv2

using namespace mv2;

extern "C" int mv2_entrain(mv2_context *ctx, const int32_t *idx, const double *depth, int n,
                           int weight_mode, double *volume_out)
{
    if (int rc = check_bound(ctx, "[DeviceBulkState]")) return rc;
    if (n < 0 || (n > 0 && (!idx || !depth)) || !volume_out)
        return set_error(ctx, MV2_EINVAL, "[DeviceBulkState] invalid entrainment request");
    state_written(ctx);
    *volume_out = 0.0;
    if (n == 0) return MV2_OK;
    if (int rc = ensure_scratch(ctx, ctx->partials, sizeof(double) * 4 * kExGrid)) return rc;
    const int grid = grid_for(n, kBlock, kExGrid);
    const double area = ctx->dx * ctx->dy, f6 = area / 6.0;
    if (ctx->precision == MV2_FP64)
        hipLaunchKernelGGL(entrain_kernel<false>, dim3(grid), dim3(kBlock), 0, ctx->stream, ctx->z_base, ctx->b_eff,
                           ctx->mobile, idx, depth, n, ctx->ny, ctx->nx, area, f6, weight_mode,
                           (double *)ctx->partials.ptr);
    else
        hipLaunchKernelGGL(entrain_kernel<true>, dim3(grid), dim3(kBlock), 0, ctx->stream, ctx->z_base, ctx->b_eff,
                           ctx->mobile, idx, depth, n, ctx->ny, ctx->nx, area, f6, weight_mode,
                           (double *)ctx->partials.ptr);
    MV2_HIP(ctx, hipGetLastError());
    return finish(ctx, 1, grid, volume_out);
}

extern "C" int mv2_deposit(mv2_context *ctx, const int32_t *idx, const double *depth, int n,
                           int weight_mode, double density, double out[3])
{
    if (int rc = check_bound(ctx, "[DeviceBulkState]")) return rc;
    if (n < 0 || (n > 0 && (!idx || !depth)) || !out || !(density > 0.0))
        return set_error(ctx, MV2_EINVAL, "[DeviceBulkState] invalid deposition indices/density");
    state_written(ctx);
    out[0] = out[1] = out[2] = 0.0;
    if (n == 0) return MV2_OK;
    if (int rc = ensure_scratch(ctx, ctx->partials, sizeof(double) * 4 * kExGrid)) return rc;
    const int grid = grid_for(n, kBlock, kExGrid);
    const double area = ctx->dx * ctx->dy, f6 = area / 6.0;
    if (ctx->precision == MV2_FP64)
        hipLaunchKernelGGL(deposit_kernel<false>, dim3(grid), dim3(kBlock), 0, ctx->stream, ctx->b_eff, ctx->mobile,
                           ctx->qx, ctx->qy, idx, depth, n, ctx->ny, ctx->nx, area, f6, weight_mode,
                           density, (double *)ctx->partials.ptr);
    else
        hipLaunchKernelGGL(deposit_kernel<true>, dim3(grid), dim3(kBlock), 0, ctx->stream, ctx->b_eff, ctx->mobile,
                           ctx->qx, ctx->qy, idx, depth, n, ctx->ny, ctx->nx, area, f6, weight_mode,
                           density, (double *)ctx->partials.ptr);
    MV2_HIP(ctx, hipGetLastError());
    return finish(ctx, 3, grid, out);
}

extern "C" int mv2_reservoirs(mv2_context *ctx, int weight_mode, double density, double out[4])
{
    if (int rc = check_bound(ctx, "[DeviceBulkState]")) return rc;
    if (!out || !(density > 0.0))
        return set_error(ctx, MV2_EINVAL, "[DeviceBulkState] density must be finite/positive");
    if (int rc = ensure_scratch(ctx, ctx->partials, sizeof(double) * 4 * kExGrid)) return rc;
    const int64_t n = (int64_t)ctx->ny * ctx->nx;
    const int grid = grid_for(n, kBlock, kExGrid);
    const double area = ctx->dx * ctx->dy, f6 = area / 6.0;
    if (ctx->precision == MV2_FP64)
        hipLaunchKernelGGL(reservoir_kernel<false>, dim3(grid), dim3(kBlock), 0, ctx->stream, ctx->z_base, ctx->b_eff,
                           ctx->mobile, ctx->qx, ctx->qy, ctx->ny, ctx->nx, area, f6, weight_mode, density,
                           (double *)ctx->partials.ptr);
    else
        hipLaunchKernelGGL(reservoir_kernel<true>, dim3(grid), dim3(kBlock), 0, ctx->stream, ctx->z_base, ctx->b_eff,
                           ctx->mobile, ctx->qx, ctx->qy, ctx->ny, ctx->nx, area, f6, weight_mode, density,
                           (double *)ctx->partials.ptr);
    MV2_HIP(ctx, hipGetLastError());
    return finish(ctx, 4, grid, out);
}

// Allocation-free variant of mv2_reservoirs: the same kernels and grid (so the
// same sums, bit for bit), the workgroup partials in the file-scope array and
// the folded result written straight to the caller's device buffer.  No
// allocation, no host copy, no stream sync; every partial that the fold reads
// is written first by its workgroup, so nothing needs clearing.
extern "C" int mv2_reservoirs_noalloc(mv2_context *ctx, int weight_mode, double density, double *out_device)
{
    if (int rc = check_bound(ctx, "[DeviceBulkState]")) return rc;
    if (!out_device || !(density > 0.0))
        return set_error(ctx, MV2_EINVAL, "[DeviceBulkState] density must be finite/positive");
    if (!ctx->noalloc_scratch)
        return set_error(ctx, MV2_EDEVICE, "[DeviceBulkState] allocation-free scratch unavailable on device %d",
                         ctx->device);
    const int64_t n = (int64_t)ctx->ny * ctx->nx;
    const int grid = grid_for(n, kBlock, kExGrid);
    if (grid > kExGrid)
        return set_error(ctx, MV2_EINVAL, "[DeviceBulkState] %d ledger partials exceed the allocation-free "
                         "scratch (%d)", grid, kExGrid);
    double *partials = ctx->noalloc_scratch;
    const double area = ctx->dx * ctx->dy, f6 = area / 6.0;
    if (ctx->precision == MV2_FP64)
        hipLaunchKernelGGL(reservoir_kernel<false>, dim3(grid), dim3(kBlock), 0, ctx->stream, ctx->z_base, ctx->b_eff,
                           ctx->mobile, ctx->qx, ctx->qy, ctx->ny, ctx->nx, area, f6, weight_mode, density, partials);
    else
        hipLaunchKernelGGL(reservoir_kernel<true>, dim3(grid), dim3(kBlock), 0, ctx->stream, ctx->z_base, ctx->b_eff,
                           ctx->mobile, ctx->qx, ctx->qy, ctx->ny, ctx->nx, area, f6, weight_mode, density, partials);
    hipLaunchKernelGGL(reduce_partials<4>, dim3(1), dim3(kBlock), 0, ctx->stream, partials, grid, out_device);
    MV2_HIP(ctx, hipGetLastError());
    return MV2_OK;
}
