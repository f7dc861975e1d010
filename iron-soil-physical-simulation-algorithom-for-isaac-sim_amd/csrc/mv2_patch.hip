// Compact host <-> device state bridges and the FailureZone activation.
//
// DeviceBulkState's compact mutation and read paths
// (pkg/runtime/bulk_state_authority.py:804-918, :1036-1137): a rectangular
// patch scattered into one float field (apply_host_patch), integer values
// scattered at flat indices (apply_host_indices), float source terms added at
// flat indices (add_host_indices) and a bounded region of every state field
// gathered in one launch (download_patch / download_region, which the
// reference assembles from one gather launch + sync per field and per tile).
//
// FailureZone activation (pkg/runtime/gpu_failure_bridge.py:114-136): the
// entrainment contract over the activation patch, then the tool forcing
// mask -- material_mask = 1 on the contact cells (affected and carrying
// mobile material after the activation) and their 4-neighbourhood inside the
// patch, where the cell carries mobile material -- with the forced flat
// indices compacted in row-major order.  The CAD geometry that produces the
// activated heights and the affected mask stays with the caller.
#include "mv2_internal.hpp"

#include <cmath>
#include <cstring>

namespace mv2 {

constexpr int kActParts = 256;       // workgroups of the activation's entrainment pass
constexpr int kActBlock = 1024;      // threads of the forcing / compaction workgroup

enum PatchField { kFieldB = 0, kFieldMobile = 1, kFieldQx = 2, kFieldQy = 3 };

// the bound state fields, passed to the kernels by value
struct mv2_context_view {
    int f32, ny, nx;
    double dx, dy;
    void *z, *b, *mobile, *qx, *qy;
};

static mv2_context_view context_view(const mv2_context *ctx)
{
    return {ctx->precision == MV2_FP32 ? 1 : 0, ctx->ny, ctx->nx, ctx->dx, ctx->dy,
            ctx->z_base, ctx->b_eff, ctx->mobile, ctx->qx, ctx->qy};
}

struct ActivationAux {
    double volume;
    int forcing_count;
    int pad;
};

__device__ inline void *field_ptr(const mv2_context_view &v, int field)
{
    return field == kFieldB ? v.b : (field == kFieldMobile ? v.mobile : (field == kFieldQx ? v.qx : v.qy));
}

// apply_host_patch (:804-845): values of a row-major h x w patch at (row0, col0)
__global__ void __launch_bounds__(kBlock) patch_scatter_kernel(mv2_context_view v, int field, int row0, int col0,
                                                               int h, int w, const double *values)
{
    const int64_t n = (int64_t)h * w;
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x) {
        const int r = (int)(k / w), c = (int)(k - (int64_t)r * w);
        const int64_t i = (int64_t)(row0 + r) * v.nx + col0 + c;
        const double x = values[k];
        if (!v.f32) {
            ((double *)field_ptr(v, field))[i] = x;
        } else if (field == kFieldMobile) {
            ((int32_t *)v.mobile)[i] = (int32_t)rint(x * kInvQuantum);   // range checked by the host
        } else {
            ((float *)field_ptr(v, field))[i] = (float)x;
        }
    }
}

// apply_host_indices (:847-885): target[idx[k]] = values[k] (or the scalar)
__global__ void __launch_bounds__(kBlock) index_scatter_int_kernel(int32_t *target, int64_t cells, const int32_t *idx,
                                                                   const int32_t *values, int32_t scalar, int n)
{
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
        const int32_t i = idx[k];
        if (i >= 0 && i < cells) target[i] = values ? values[k] : scalar;
    }
}

// add_host_indices (:887-918): field[idx[k]] += values[k], atomically (indices
// may repeat); FP32 mobile adds whole quanta
__global__ void __launch_bounds__(kBlock) index_add_kernel(mv2_context_view v, int field, const int32_t *idx,
                                                           const double *values, int n)
{
    const int64_t cells = (int64_t)v.ny * v.nx;
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
        const int32_t i = idx[k];
        if (i < 0 || i >= cells) continue;
        const double x = values[k];
        if (!v.f32) {
            atomicAdd(&((double *)field_ptr(v, field))[i], x);
        } else if (field == kFieldMobile) {
            atomicAdd(&((int32_t *)v.mobile)[i], (int32_t)rint(x * kInvQuantum));
        } else {
            atomicAdd(&((float *)field_ptr(v, field))[i], (float)x);
        }
    }
}

// download_patch / download_region (:1036-1137): out[f][k], fields b_eff,
// z_base, mobile (metres), momentum_x, momentum_y of the h x w region
__global__ void __launch_bounds__(kBlock) region_gather_kernel(mv2_context_view v, int row0, int col0, int h, int w,
                                                               double *out)
{
    const int64_t n = (int64_t)h * w;
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x) {
        const int r = (int)(k / w), c = (int)(k - (int64_t)r * w);
        const int64_t i = (int64_t)(row0 + r) * v.nx + col0 + c;
        double b, z, m, qx, qy;
        if (!v.f32) {
            b = ((const double *)v.b)[i];
            z = ((const double *)v.z)[i];
            m = ((const double *)v.mobile)[i];
            qx = ((const double *)v.qx)[i];
            qy = ((const double *)v.qy)[i];
        } else {
            b = ((const float *)v.b)[i];
            z = ((const float *)v.z)[i];
            m = (double)((const int32_t *)v.mobile)[i] * kQuantum;
            qx = ((const float *)v.qx)[i];
            qy = ((const float *)v.qy)[i];
        }
        out[k] = b;
        out[n + k] = z;
        out[2 * n + k] = m;
        out[3 * n + k] = qx;
        out[4 * n + k] = qy;
    }
}

// Activation pass 1: per patch cell the flags of the forcing rule, from the
// mobile depth before the activation plus the activated height
// (gpu_failure_bridge.py:123-124: mobile_after = patch mobile + activated;
// bit 0 mobile_after > 0, bit 1 contact = affected && mobile_after > 0), then
// the entrainment contract of that cell (bulk_state_authority.py:104-127 /
// :920-960 with the Triangle-A-C weights) and the workgroup's moved volume.
__global__ void __launch_bounds__(kBlock) activation_entrain_kernel(mv2_context_view v, int row0, int col0, int h,
                                                                    int w, const double *activated,
                                                                    const int32_t *affected, uint8_t *flags,
                                                                    double *partials)
{
    __shared__ double smem[4];
    double vol[1] = {0.0};
    const int64_t n = (int64_t)h * w;
    const double area = v.dx * v.dy, f6 = area / 6.0;
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x) {
        const int r = (int)(k / w), c = (int)(k - (int64_t)r * w);
        const int gr = row0 + r, gc = col0 + c;
        const int64_t i = (int64_t)gr * v.nx + gc;
        const double act = activated[k];
        const double wgt = vertex_weight(v.ny, v.nx, f6, gr, gc);
        double before;
        if (!v.f32) {
            const double *z = (const double *)v.z;
            double *b = (double *)v.b;
            double *m = (double *)v.mobile;
            before = m[i];
            const double available = fmax(b[i] - z[i], 0.0);
            const double depth = fmin(fmax(act, 0.0), available);
            b[i] = b[i] - depth;
            m[i] = m[i] + depth;
            vol[0] += depth * wgt;
        } else {
            const float *z = (const float *)v.z;
            float *b = (float *)v.b;
            int32_t *m = (int32_t *)v.mobile;
            before = (double)m[i] * kQuantum;
            const double available = fmax((double)b[i] - (double)z[i], 0.0);
            const double depth = fmin(fmax(act, 0.0), available);
            const long long dq = entrain_quanta_f32(b[i], depth, 2147483647LL - (long long)m[i]);
            const double moved = (double)dq * kQuantum;
            b[i] = (float)((double)b[i] - moved);
            m[i] = (int32_t)((long long)m[i] + dq);
            vol[0] += moved * wgt;
        }
        const bool wet = before + act > 0.0;
        flags[k] = (uint8_t)((wet ? 1 : 0) | ((wet && affected[k] != 0) ? 2 : 0));
    }
    block_reduce<double, 1, 0>(vol, smem);
    if (threadIdx.x == 0) partials[blockIdx.x] = vol[0];
}

// Activation pass 2 (one workgroup): forced = wet && (contact here or in a
// 4-neighbour inside the patch); material_mask = 1 there, and the forced flat
// indices in row-major patch order (each thread owns a contiguous run of
// cells, a block scan gives its output offset).  Also folds the moved volume.
__global__ void __launch_bounds__(kActBlock) activation_forcing_kernel(int nx, int row0, int col0, int h, int w,
                                                                       const uint8_t *flags, int32_t *material_mask,
                                                                       int32_t *forcing_out, const double *partials,
                                                                       int nparts, ActivationAux *aux)
{
    __shared__ int scan[kActBlock];
    __shared__ double vsum[kActBlock / kWave];
    const int64_t n = (int64_t)h * w;
    const int64_t per = (n + kActBlock - 1) / kActBlock;
    const int64_t k0 = (int64_t)threadIdx.x * per;
    const int64_t k1 = k0 + per < n ? k0 + per : n;
    auto forced = [&](int64_t k) {
        if (!(flags[k] & 1)) return false;
        const int r = (int)(k / w), c = (int)(k - (int64_t)r * w);
        return ((flags[k] >> 1) & 1) || (r > 0 && ((flags[k - w] >> 1) & 1)) ||
               (r + 1 < h && ((flags[k + w] >> 1) & 1)) || (c > 0 && ((flags[k - 1] >> 1) & 1)) ||
               (c + 1 < w && ((flags[k + 1] >> 1) & 1));
    };
    int count = 0;
    for (int64_t k = k0; k < k1; k++) count += forced(k) ? 1 : 0;
    scan[threadIdx.x] = count;
    __syncthreads();
    for (int off = 1; off < kActBlock; off <<= 1) {          // inclusive Hillis-Steele scan
        const int add = threadIdx.x >= off ? scan[threadIdx.x - off] : 0;
        __syncthreads();
        scan[threadIdx.x] += add;
        __syncthreads();
    }
    int pos = scan[threadIdx.x] - count;
    for (int64_t k = k0; k < k1; k++) {
        if (!forced(k)) continue;
        const int r = (int)(k / w), c = (int)(k - (int64_t)r * w);
        const int32_t i = (int32_t)((int64_t)(row0 + r) * nx + col0 + c);
        material_mask[i] = 1;
        forcing_out[pos++] = i;
    }
    // moved volume: partials folded in workgroup order
    double x = 0.0;
    for (int p = threadIdx.x; p < nparts; p += kActBlock) x += partials[p];
    x = wave_reduce<0>(x);
    if ((threadIdx.x & (kWave - 1)) == 0) vsum[threadIdx.x / kWave] = x;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int q = 0; q < kActBlock / kWave; q++) t += vsum[q];
        aux->volume = t;
        aux->forcing_count = scan[kActBlock - 1];
        aux->pad = 0;
    }
}

static int check_bbox(mv2_context *ctx, int row0, int row1, int col0, int col1, bool allow_empty, const char *what)
{
    const bool ok = allow_empty ? (0 <= row0 && row0 <= row1 && row1 <= ctx->ny && 0 <= col0 && col0 <= col1 &&
                                   col1 <= ctx->nx)
                                : (0 <= row0 && row0 < row1 && row1 <= ctx->ny && 0 <= col0 && col0 < col1 &&
                                   col1 <= ctx->nx);
    if (!ok) return set_error(ctx, MV2_EINVAL, "[DeviceBulkState] %s bbox outside state", what);
    return MV2_OK;
}

static int bound(mv2_context *ctx)
{
    if (!ctx) return set_error(nullptr, MV2_EINVAL, "[DeviceBulkState] null context");
    if (!ctx->bound) return set_error(ctx, MV2_ESTATE, "[DeviceBulkState] bind state first");
    hipSetDevice(ctx->device);
    return MV2_OK;
}

}  // namespace mv2

using namespace mv2;

extern "C" int mv2_patch_scatter(mv2_context *ctx, int field, int row0, int row1, int col0, int col1,
                                 const double *values)
{
    if (int rc = bound(ctx)) return rc;
    if (field < kFieldB || field > kFieldQy) return set_error(ctx, MV2_EINVAL, "[DeviceBulkState] unsupported mutable field");
    if (int rc = check_bbox(ctx, row0, row1, col0, col1, true, "patch")) return rc;
    state_written(ctx);
    const int64_t n = (int64_t)(row1 - row0) * (col1 - col0);
    if (n == 0) return MV2_OK;
    if (!values) return set_error(ctx, MV2_EINVAL, "[DeviceBulkState] patch values missing");
    hipLaunchKernelGGL(patch_scatter_kernel, dim3(grid_for(n, kBlock, 4096)), dim3(kBlock), 0, ctx->stream,
                       context_view(ctx), field, row0, col0, row1 - row0, col1 - col0, values);
    MV2_HIP(ctx, hipGetLastError());
    return MV2_OK;
}

extern "C" int mv2_index_scatter_int(mv2_context *ctx, int32_t *target, const int32_t *idx, const int32_t *values,
                                     int32_t scalar, int n)
{
    if (int rc = bound(ctx)) return rc;
    if (n < 0 || (n > 0 && (!target || !idx)))
        return set_error(ctx, MV2_EINVAL, "[DeviceBulkState] invalid compact scatter");
    if (n == 0) return MV2_OK;
    hipLaunchKernelGGL(index_scatter_int_kernel, dim3(grid_for(n, kBlock, 4096)), dim3(kBlock), 0, ctx->stream,
                       target, (int64_t)ctx->ny * ctx->nx, idx, values, scalar, n);
    MV2_HIP(ctx, hipGetLastError());
    return MV2_OK;
}

extern "C" int mv2_index_add(mv2_context *ctx, int field, const int32_t *idx, const double *values, int n)
{
    if (int rc = bound(ctx)) return rc;
    if (field < kFieldB || field > kFieldQy) return set_error(ctx, MV2_EINVAL, "[DeviceBulkState] unsupported compact float field");
    if (n < 0 || (n > 0 && (!idx || !values)))
        return set_error(ctx, MV2_EINVAL, "[DeviceBulkState] invalid compact add");
    state_written(ctx);
    if (n == 0) return MV2_OK;
    hipLaunchKernelGGL(index_add_kernel, dim3(grid_for(n, kBlock, 4096)), dim3(kBlock), 0, ctx->stream,
                       context_view(ctx), field, idx, values, n);
    MV2_HIP(ctx, hipGetLastError());
    return MV2_OK;
}

extern "C" int mv2_region_gather(mv2_context *ctx, int row0, int row1, int col0, int col1, double *out)
{
    if (int rc = bound(ctx)) return rc;
    if (int rc = check_bbox(ctx, row0, row1, col0, col1, false, "region")) return rc;
    if (!out) return set_error(ctx, MV2_EINVAL, "[DeviceBulkState] region output missing");
    const int64_t n = (int64_t)(row1 - row0) * (col1 - col0);
    hipLaunchKernelGGL(region_gather_kernel, dim3(grid_for(n, kBlock, 4096)), dim3(kBlock), 0, ctx->stream,
                       context_view(ctx), row0, col0, row1 - row0, col1 - col0, out);
    MV2_HIP(ctx, hipGetLastError());
    return MV2_OK;
}

extern "C" int mv2_failure_activate(mv2_context *ctx, int row0, int row1, int col0, int col1,
                                    const double *activated, const int32_t *affected, int32_t *material_mask,
                                    int32_t *forcing_out, mv2_activation_result *out)
{
    if (int rc = bound(ctx)) return rc;
    if (int rc = check_bbox(ctx, row0, row1, col0, col1, false, "activation patch")) return rc;
    state_written(ctx);
    if (!activated || !affected || !material_mask || !forcing_out || !out)
        return set_error(ctx, MV2_EINVAL, "[GpuFailureBridge] null argument");
    const int h = row1 - row0, w = col1 - col0;
    const int64_t n = (int64_t)h * w;
    const size_t flag_bytes = ((size_t)n + 255) & ~(size_t)255;
    if (int rc = ensure_scratch(ctx, ctx->patch, flag_bytes + sizeof(double) * kActParts + sizeof(ActivationAux)))
        return rc;
    uint8_t *flags = (uint8_t *)ctx->patch.ptr;
    double *partials = (double *)((char *)ctx->patch.ptr + flag_bytes);
    ActivationAux *aux = (ActivationAux *)(partials + kActParts);
    const int grid = grid_for(n, kBlock, kActParts);
    hipLaunchKernelGGL(activation_entrain_kernel, dim3(grid), dim3(kBlock), 0, ctx->stream, context_view(ctx), row0,
                       col0, h, w, activated, affected, flags, partials);
    hipLaunchKernelGGL(activation_forcing_kernel, dim3(1), dim3(kActBlock), 0, ctx->stream, ctx->nx, row0, col0, h, w,
                       flags, material_mask, forcing_out, partials, grid, aux);
    MV2_HIP(ctx, hipGetLastError());
    MV2_HIP(ctx, hipMemcpyAsync(ctx->host_result.ptr, aux, sizeof(ActivationAux), hipMemcpyDeviceToHost, ctx->stream));
    if (int rc_w = stream_wait(ctx, ctx->stream)) return rc_w;
    ActivationAux host;
    memcpy(&host, ctx->host_result.ptr, sizeof(host));
    out->activated_volume = host.volume;
    out->forcing_count = host.forcing_count;
    out->reserved = 0;
    return MV2_OK;
}
