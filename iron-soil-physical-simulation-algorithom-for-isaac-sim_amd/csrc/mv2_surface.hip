// Surface publication for the visual/contact consumers of the terrain
// (pkg/runtime/bulk_state_authority.py:156-232, :733-781, :1138-1182,
// :1280-1328): the free surface b_eff + mobile is snapshotted before a
// mutating operator and compared after it to flag the tiles that changed;
// changed tiles are gathered (with their one-cell overlap) for publication,
// and airborne parcels sample the surface bilinearly.
//
// The reference flags tiles with atomic_max and then downloads the whole flag
// array; it gathers one tile per launch with a host round trip each.  Here the
// flags are compacted to ascending tile ids on the device (one workgroup), and
// every selected tile is gathered by one launch into a packed buffer with a
// fixed (tile_size + 1)^2 stride.
#include "mv2_internal.hpp"

#include <cmath>
#include <cstring>

namespace mv2 {

__device__ inline double free_surface(const void *b, const void *m, int f32, int64_t i)
{
    if (f32) return (double)((const float *)b)[i] + (double)((const int32_t *)m)[i] * kQuantum;
    return ((const double *)b)[i] + ((const double *)m)[i];
}

__global__ void __launch_bounds__(kBlock) surface_capture_kernel(const void *b, const void *m, int f32, int64_t n,
                                                                 double *before)
{
    for (int64_t i = (int64_t)xcd_block() * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock)
        before[i] = free_surface(b, m, f32, i);
}

// flag_surface_change (:170-187) without the transient per-cell mask (the
// reference zeroes it again before returning)
__global__ void __launch_bounds__(kBlock) surface_flag_kernel(const void *b, const void *m, int f32, int64_t n,
                                                              int nx, const double *before, double tolerance,
                                                              int tile_size, int tiles_x, int32_t *tile_flags)
{
    for (int64_t i = (int64_t)xcd_block() * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
        if (fabs(free_surface(b, m, f32, i) - before[i]) > tolerance) {
            const int row = (int)(i / nx), col = (int)(i - (int64_t)row * nx);
            const int tile = (row / tile_size) * tiles_x + col / tile_size;
            if (tile_flags[tile] == 0) tile_flags[tile] = 1;   // every writer stores 1
        }
    }
}

// ascending ids of the flagged tiles: out[0] = count, out[1..count] = ids;
// one workgroup, block-wide exclusive scan per 256 flags
__global__ void __launch_bounds__(kBlock) surface_compact_kernel(const int32_t *tile_flags, int tile_count,
                                                                 int32_t *out)
{
    __shared__ int32_t scan[kBlock];
    __shared__ int32_t base;
    if (threadIdx.x == 0) base = 0;
    __syncthreads();
    for (int t0 = 0; t0 < tile_count; t0 += kBlock) {
        const int t = t0 + threadIdx.x;
        const int on = t < tile_count && tile_flags[t] != 0;
        scan[threadIdx.x] = on;
        __syncthreads();
        for (int off = 1; off < kBlock; off <<= 1) {
            const int add = threadIdx.x >= off ? scan[threadIdx.x - off] : 0;
            __syncthreads();
            scan[threadIdx.x] += add;
            __syncthreads();
        }
        if (on) out[1 + base + scan[threadIdx.x] - 1] = t;
        __syncthreads();
        if (threadIdx.x == kBlock - 1) base += scan[kBlock - 1];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = base;
}

// bilinear_surface (:196-228): clamp to the grid, the four corners' free
// surface, the reference's weighting order
__global__ void __launch_bounds__(kBlock) surface_sample_kernel(const void *b, const void *m, int f32, int ny,
                                                                int nx, const double *rows_cols, int n,
                                                                double *out)
{
    const int k = blockIdx.x * kBlock + threadIdx.x;
    if (k >= n) return;
    const double row = fmin(fmax(rows_cols[2 * k], 0.0), (double)(ny - 1));
    const double col = fmin(fmax(rows_cols[2 * k + 1], 0.0), (double)(nx - 1));
    const int r0 = (int)floor(row), c0 = (int)floor(col);
    const int r1 = min(r0 + 1, ny - 1), c1 = min(c0 + 1, nx - 1);
    const double fr = row - (double)r0, fc = col - (double)c0;
    const double v00 = free_surface(b, m, f32, (int64_t)r0 * nx + c0);
    const double v01 = free_surface(b, m, f32, (int64_t)r0 * nx + c1);
    const double v10 = free_surface(b, m, f32, (int64_t)r1 * nx + c0);
    const double v11 = free_surface(b, m, f32, (int64_t)r1 * nx + c1);
    out[k] = (1.0 - fr) * (1.0 - fc) * v00 + (1.0 - fr) * fc * v01 + fr * (1.0 - fc) * v10 + fr * fc * v11;
}

// gather of tiles with their one-cell overlap (:1280-1326): tile j of the
// list fills out[j * stride ...] row-major over its (r1 - r0) x (c1 - c0) cells
__global__ void __launch_bounds__(kBlock) surface_gather_kernel(const void *b, const void *m, int f32, int ny,
                                                                int nx, const int32_t *tile_ids, int tile_size,
                                                                int tiles_x, int tile_count, int64_t stride,
                                                                double *out)
{
    const int j = blockIdx.y;
    const int tile = tile_ids[j];
    if (tile < 0 || tile >= tile_count) return;     // ids are checked by the caller; never read outside
    const int tr = tile / tiles_x, tc = tile - tr * tiles_x;
    const int r0 = tr * tile_size, c0 = tc * tile_size;
    const int h = min(r0 + tile_size + 1, ny) - r0, w = min(c0 + tile_size + 1, nx) - c0;
    for (int e = blockIdx.x * kBlock + threadIdx.x; e < h * w; e += gridDim.x * kBlock) {
        const int r = e / w, c = e - r * w;
        out[(int64_t)j * stride + e] = free_surface(b, m, f32, (int64_t)(r0 + r) * nx + (c0 + c));
    }
}

static int surface_ctx(mv2_context *ctx, const char *who)
{
    if (!ctx) return set_error(nullptr, MV2_EINVAL, "%s null context", who);
    if (!ctx->bound) return set_error(ctx, MV2_ESTATE, "%s bind state first", who);
    hipSetDevice(ctx->device);
    return MV2_OK;
}

static bool tile_grid_ok(const mv2_context *ctx, int tile_size, int tiles_x, int tile_count)
{
    return tile_size >= 1 && tiles_x == (ctx->nx + tile_size - 1) / tile_size &&
           tile_count == tiles_x * ((ctx->ny + tile_size - 1) / tile_size);
}

}  // namespace mv2

using namespace mv2;

extern "C" int mv2_surface_capture(mv2_context *ctx, double *surface_before)
{
    if (int rc = surface_ctx(ctx, "[DeviceBulkState]")) return rc;
    if (!surface_before) return set_error(ctx, MV2_EINVAL, "[DeviceBulkState] null surface buffer");
    const int64_t n = (int64_t)ctx->ny * ctx->nx;
    surface_capture_kernel<<<grid_for(n), kBlock, 0, ctx->stream>>>(ctx->b_eff, ctx->mobile,
                                                                    ctx->precision == MV2_FP32, n, surface_before);
    MV2_HIP(ctx, hipGetLastError());
    return MV2_OK;
}

extern "C" int mv2_surface_dirty_tiles(mv2_context *ctx, const double *surface_before, double tolerance,
                                       int tile_size, int tiles_x, int tile_count, int32_t *tile_flags,
                                       int32_t *tile_ids_out, int *count_out)
{
    if (int rc = surface_ctx(ctx, "[DeviceBulkState]")) return rc;
    if (!surface_before || !tile_flags || !tile_ids_out || !count_out)
        return set_error(ctx, MV2_EINVAL, "[DeviceBulkState] null dirty-tile argument");
    if (!std::isfinite(tolerance) || tolerance < 0.0)
        return set_error(ctx, MV2_EINVAL, "[DeviceBulkState] dirty tolerance must be finite/non-negative");
    if (!tile_grid_ok(ctx, tile_size, tiles_x, tile_count))
        return set_error(ctx, MV2_EINVAL, "[DeviceBulkState] tile grid does not match the state");
    const int64_t n = (int64_t)ctx->ny * ctx->nx;
    const size_t bytes = sizeof(int32_t) * ((size_t)tile_count + 1);
    if (int rc = ensure_scratch(ctx, ctx->label_aux, bytes)) return rc;
    if (int rc = ensure_scratch(ctx, ctx->host_result, bytes < 4096 ? 4096 : bytes, true)) return rc;
    int32_t *dout = (int32_t *)ctx->label_aux.ptr;
    MV2_HIP(ctx, hipMemsetAsync(tile_flags, 0, sizeof(int32_t) * (size_t)tile_count, ctx->stream));
    surface_flag_kernel<<<grid_for(n), kBlock, 0, ctx->stream>>>(ctx->b_eff, ctx->mobile, ctx->precision == MV2_FP32,
                                                                 n, ctx->nx, surface_before, tolerance, tile_size,
                                                                 tiles_x, tile_flags);
    surface_compact_kernel<<<1, kBlock, 0, ctx->stream>>>(tile_flags, tile_count, dout);
    MV2_HIP(ctx, hipGetLastError());
    MV2_HIP(ctx, hipMemcpyAsync(ctx->host_result.ptr, dout, bytes, hipMemcpyDeviceToHost, ctx->stream));
    if (int rc_w = stream_wait(ctx, ctx->stream)) return rc_w;
    const int32_t *h = (const int32_t *)ctx->host_result.ptr;
    *count_out = h[0];
    memcpy(tile_ids_out, h + 1, sizeof(int32_t) * (size_t)h[0]);
    return MV2_OK;
}

extern "C" int mv2_surface_sample(mv2_context *ctx, const double *rows_cols, int n, double *out)
{
    if (int rc = surface_ctx(ctx, "[DeviceBulkState]")) return rc;
    if (n < 0 || (n > 0 && (!rows_cols || !out)))
        return set_error(ctx, MV2_EINVAL, "[DeviceBulkState] terrain queries must be finite [N,2]");
    if (n == 0) return MV2_OK;
    surface_sample_kernel<<<(n + kBlock - 1) / kBlock, kBlock, 0, ctx->stream>>>(
        ctx->b_eff, ctx->mobile, ctx->precision == MV2_FP32, ctx->ny, ctx->nx, rows_cols, n, out);
    MV2_HIP(ctx, hipGetLastError());
    return MV2_OK;
}

extern "C" int mv2_surface_gather_tiles(mv2_context *ctx, const int32_t *tile_ids, int n_tiles, int tile_size,
                                        int tiles_x, int tile_count, double *out)
{
    if (int rc = surface_ctx(ctx, "[DeviceBulkState]")) return rc;
    if (n_tiles < 0 || (n_tiles > 0 && (!tile_ids || !out)))
        return set_error(ctx, MV2_EINVAL, "[DeviceBulkState] invalid tile gather");
    if (!tile_grid_ok(ctx, tile_size, tiles_x, tile_count))
        return set_error(ctx, MV2_EINVAL, "[DeviceBulkState] tile grid does not match the state");
    if (n_tiles == 0) return MV2_OK;
    if (n_tiles > 65535) return set_error(ctx, MV2_EINVAL, "[DeviceBulkState] at most 65535 tiles per gather");
    const int64_t stride = (int64_t)(tile_size + 1) * (tile_size + 1);
    const int gx = (int)((stride + kBlock - 1) / kBlock < 64 ? (stride + kBlock - 1) / kBlock : 64);
    surface_gather_kernel<<<dim3(gx, n_tiles), kBlock, 0, ctx->stream>>>(
        ctx->b_eff, ctx->mobile, ctx->precision == MV2_FP32, ctx->ny, ctx->nx, tile_ids, tile_size, tiles_x,
        tile_count, stride, out);
    MV2_HIP(ctx, hipGetLastError());
    return MV2_OK;
}
