// Angle-of-repose relaxation (slope_model.relax_critical_slope,
// slope_model.py:273-343) and the compact colour-batch sweep
// (pkg/solvers/warp_compact_frontier.py:57-130).
//
// relax: one launch per iteration.  Each workgroup stages a tile plus an
// 8-cell halo in LDS and performs all 8 red-black sub-phases of an iteration
// there.  A sub-phase moves information one cell along its direction, so after
// the 8 sub-phases (two per direction, directions x, y and both diagonals) at
// most 6 halo layers are stale; the interior and every edge it owns are exact.
// Then the interior is written (ping-pong buffers) and the iteration's maximum
// excess is reduced into a spread maximum (64 slots).  The next launch tests
// it against the tolerance on the device and returns at once once converged,
// so the host enqueues batches of launches and reads one result per batch.
// An iteration costs one read + one write of the field instead of eight
// full-array passes; a tile whose 3 x 3 tile neighbourhood did not change in
// the previous launch is skipped (its field and excess are already known).
// Directions and parities: mv2_relax_tile.hpp.
#include "mv2_internal.hpp"
#include "mv2_frontier_edge.hpp"
#include "mv2_relax_tile.hpp"

#include <cmath>
#include <cstring>

namespace mv2 {

// Iteration state: launch k reads ctl[k & 1], its workgroup 0 writes
// ctl[(k + 1) & 1]; launch k reduces its maximum excess into spread set k % 3,
// tests set (k - 1) % 3 (the previous iteration) against the tolerance and
// clears set (k + 1) % 3 for the next launch.
struct RelaxControl {
    int iterations;      // iterations applied to the field
    int src;             // buffer holding the current field
    int done, converged;
    double last;         // maximum excess of the last applied iteration
};
struct RelaxSync {
    RelaxControl ctl[2];
};

template <class R>
struct RelaxArgs {
    R *h[2];
    int ny, nx;
    R crit[4];
    int max_iterations;
    double tolerance;
    RelaxSync *sync;
    unsigned long long *spread;   // three spread-maximum sets
    mv2_relax_result *out;
    int tiles_x, tiles;
    int32_t *tchanged;            // [2][tiles]: the tile's interior changed in launch k (parity k & 1)
    double *texcess;              // [tiles]: the tile's maximum excess in its last computed iteration
};

// loop state for launch k: did iteration k (applied by launch k - 1) converge?
// Every thread of the block calls it.
template <class R>
__device__ inline RelaxControl relax_resolve(const RelaxArgs<R> &a, int k, double *smem1)
{
    RelaxControl C = a.sync->ctl[k & 1];      // written by an earlier launch: plain load
    if (!C.done && k > 0) {
        C.last = spread_read_max(a.spread + ((k - 1) % 3) * kSpreadWords, smem1);
        if (C.last <= a.tolerance) {
            C.converged = 1;
            C.done = 1;
        } else if (C.iterations >= a.max_iterations) {
            C.done = 1;
        }
    }
    return C;
}

// One relaxation iteration per launch (the tile work is in mv2_relax_tile.hpp).
template <class R>
__global__ void __launch_bounds__(kRelaxBlock) relax_kernel(RelaxArgs<R> a, int k)
{
    __shared__ R tile[kRelaxLH * kRelaxLW];
    __shared__ double smem[8];
    RelaxSync *s = a.sync;
    const bool lead = blockIdx.x == 0 && threadIdx.x == 0;
    const RelaxControl C = relax_resolve(a, k, smem);
    if (C.done) {
        if (lead) s->ctl[(k + 1) & 1] = C;
        return;
    }
    const int src = C.src;
    if (blockIdx.x == 0) spread_clear(a.spread + ((k + 1) % 3) * kSpreadWords);
    if (lead) {
        RelaxControl nc = C;
        nc.iterations = C.iterations + 1;
        nc.src = src ^ 1;
        s->ctl[(k + 1) & 1] = nc;
    }
    const int t = blockIdx.x;
    const int trow = t / a.tiles_x, tcol = t % a.tiles_x;
    const int r0 = trow * kRelaxTH, c0 = tcol * kRelaxTW;
    // Quiet tiles: an iteration of a tile reads its interior and an 8-cell
    // halo, i.e. its 3 x 3 tile neighbourhood.  When none of those tiles
    // changed in launch k - 1, the tile's input equals the input of its last
    // computed iteration, which changed nothing, so this iteration would
    // reproduce the same field (already in both ping-pong buffers) and the same
    // excess: the stored excess stands in for it.  Launch 0 runs every tile.
    if (k > 0) {
        const int ty = a.tiles / a.tiles_x;
        int flag = 0;
        if (threadIdx.x < 9) {
            const int rr = trow + (int)threadIdx.x / 3 - 1, cc = tcol + (int)threadIdx.x % 3 - 1;
            if (rr >= 0 && rr < ty && cc >= 0 && cc < a.tiles_x)
                flag = a.tchanged[((k - 1) & 1) * a.tiles + rr * a.tiles_x + cc];
        }
        if (!__syncthreads_or(flag)) {
            if (threadIdx.x == 0) {
                a.tchanged[(k & 1) * a.tiles + t] = 0;
                spread_max(a.spread + (k % 3) * kSpreadWords, a.texcess[t]);
            }
            return;
        }
    }
    bool changed = false;
    const R tmax = relax_tile_iteration<R>(a.h[src], a.h[src ^ 1], a.ny, a.nx, a.crit, r0, c0, tile, &changed);
    double m[1] = {(double)tmax};
    block_reduce<double, 1, 1, kRelaxBlock / kWave>(m, smem);
    const int any = __syncthreads_or(changed ? 1 : 0);
    if (threadIdx.x == 0) {
        spread_max(a.spread + (k % 3) * kSpreadWords, m[0]);
        a.texcess[t] = m[0];
        a.tchanged[(k & 1) * a.tiles + t] = any;
    }
}

// after launch k - 1: copy back when the field sits in scratch, write the result
template <class R>
__global__ void __launch_bounds__(kBlock) relax_end_kernel(RelaxArgs<R> a, int k)
{
    __shared__ double smem[1];
    const RelaxControl C = relax_resolve(a, k, smem);
    if (C.src == 1) {
        const int64_t n = (int64_t)a.ny * a.nx;
        for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
             i += (int64_t)gridDim.x * blockDim.x)
            a.h[0][i] = a.h[1][i];
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        a.out->iterations = C.iterations;
        a.out->converged = C.converged;
        a.out->max_excess_last = C.last;
        a.out->complete = C.done;
    }
}

__global__ void zero_relax_sync(RelaxSync *s, unsigned long long *spread)
{
    unsigned *p = reinterpret_cast<unsigned *>(s);
    for (int i = threadIdx.x; i < (int)(sizeof(RelaxSync) / 4); i += blockDim.x) p[i] = 0u;
    for (int i = threadIdx.x; i < 3 * kSpread; i += blockDim.x) spread[(int64_t)i * kSpreadStride] = 0ull;
}

template <class R>
static int launch_relax(mv2_context *ctx, void *height, int ny, int nx, const double crit_diff[4],
                        int max_iterations, double tolerance, mv2_relax_result *out)
{
    const size_t n = (size_t)ny * nx;
    if (int rc = ensure_scratch(ctx, ctx->relax_alt, n * sizeof(R))) return rc;
    RelaxArgs<R> a;
    a.h[0] = (R *)height;
    a.h[1] = (R *)ctx->relax_alt.ptr;
    a.ny = ny;
    a.nx = nx;
    for (int d = 0; d < 4; d++) a.crit[d] = (R)crit_diff[d];
    a.max_iterations = max_iterations;
    a.tolerance = tolerance;
    // the relaxation keeps its own barrier block after the V2 sync block
    a.sync = (RelaxSync *)((char *)ctx->sync.ptr + 1024);
    if (int rc = ensure_scratch(ctx, ctx->spread, kSpreadScratchBytes)) return rc;
    a.spread = (unsigned long long *)ctx->spread.ptr + (size_t)kRelaxSpreadSet * kSpreadWords;
    a.out = (mv2_relax_result *)((char *)ctx->result.ptr + 2048);
    a.tiles_x = (nx + kRelaxTW - 1) / kRelaxTW;
    a.tiles = a.tiles_x * ((ny + kRelaxTH - 1) / kRelaxTH);
    if (int rc = ensure_scratch(ctx, ctx->relax_quiet, (2 * sizeof(int32_t) + sizeof(double)) * (size_t)a.tiles))
        return rc;
    a.texcess = (double *)ctx->relax_quiet.ptr;
    a.tchanged = (int32_t *)(a.texcess + a.tiles);
    hipStream_t st = ctx->stream;
    const int gend = grid_for((int64_t)n, kBlock, 4 * ctx->cu_count);
    hipLaunchKernelGGL(zero_relax_sync, dim3(1), dim3(kBlock), 0, st, a.sync, a.spread);
    MV2_HIP(ctx, hipGetLastError());
    // One launch per iteration; launches after convergence return at once.
    // Enqueue about as many as the previous call needed (at most 32), then
    // grow the batch geometrically while the field has not converged.
    int chunk = ctx->relax_predicted + 1;
    if (chunk < 2) chunk = 2;
    if (chunk > 32) chunk = 32;
    int k = 0;
    const mv2_relax_result *hres = (const mv2_relax_result *)ctx->host_result.ptr;
    for (;;) {
        if (chunk > max_iterations - k) chunk = max_iterations - k;
        for (int j = 0; j < chunk; j++, k++)
            hipLaunchKernelGGL(relax_kernel<R>, dim3(a.tiles), dim3(kRelaxBlock), 0, st, a, k);
        hipLaunchKernelGGL(relax_end_kernel<R>, dim3(gend), dim3(kBlock), 0, st, a, k);
        MV2_HIP(ctx, hipGetLastError());
        MV2_HIP(ctx, hipMemcpyAsync(ctx->host_result.ptr, a.out, sizeof(mv2_relax_result), hipMemcpyDeviceToHost, st));
        if (int rc_w = stream_wait(ctx, st)) return rc_w;
        if (hres->complete || k >= max_iterations) break;
        chunk = k / 2 > 8 ? k / 2 : 8;
    }
    memcpy(out, hres, sizeof(mv2_relax_result));
    out->launches = k;
    ctx->relax_predicted = out->iterations;
    return MV2_OK;
}

// ---- compact colour batch (warp_compact_frontier.py:57-130) -------------------
__device__ inline void mark_tile_halo(int row, int col, int rows, int cols, int tile_size, int tiles_x,
                                      int32_t *flags)
{
    for (int orow = -1; orow < 2; orow++)
        for (int ocol = -1; ocol < 2; ocol++) {
            int nr = row + orow, nc = col + ocol;
            if (nr >= 0 && nr < rows && nc >= 0 && nc < cols)
                atomicMax(&flags[(nr / tile_size) * tiles_x + nc / tile_size], 1);
        }
}

__global__ void __launch_bounds__(kBlock) frontier_relax_kernel(
    double *height, const double *weights, int32_t *reached, const int32_t *a_i, const int32_t *a_j,
    const int32_t *b_i, const int32_t *b_j, int n_edges, int rows, int cols, int tile_size, int tiles_x,
    double critical_difference, int32_t *moved_flags, int32_t *moved_count)
{
    int moved = 0;
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < n_edges; e += gridDim.x * blockDim.x) {
        const int64_t a = (int64_t)a_i[e] * cols + a_j[e];
        const int64_t b = (int64_t)b_i[e] * cols + b_j[e];
        if (frontier_edge_transfer(height, weights, reached, a, b, critical_difference)) {
            moved++;
            mark_tile_halo(a_i[e], a_j[e], rows, cols, tile_size, tiles_x, moved_flags);
            mark_tile_halo(b_i[e], b_j[e], rows, cols, tile_size, tiles_x, moved_flags);
        }
    }
    int total = moved;
    for (int off = kWave / 2; off > 0; off >>= 1) total += __shfl_down(total, off, kWave);
    if ((threadIdx.x & (kWave - 1)) == 0 && total) atomicAdd(moved_count, total);
}

__global__ void __launch_bounds__(kBlock) frontier_scan_kernel(
    const double *height, const int32_t *reached, const int32_t *tile_ids, const int32_t *a_i,
    const int32_t *a_j, const int32_t *b_i, const int32_t *b_j, int n_edges, int cols,
    double critical_difference, double tolerance, int32_t *flags)
{
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < n_edges; e += gridDim.x * blockDim.x) {
        const int64_t a = (int64_t)a_i[e] * cols + a_j[e];
        const int64_t b = (int64_t)b_i[e] * cols + b_j[e];
        if (frontier_edge_unstable(height, reached, a, b, critical_difference, tolerance))
            atomicMax(&flags[tile_ids[e]], 1);
    }
}

__global__ void fill_i32(int32_t *p, int n, int32_t v)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = v;
}

}  // namespace mv2

using namespace mv2;

extern "C" int mv2_relax(mv2_context *ctx, int precision, void *height, int ny, int nx,
                         const double crit_diff[4], int max_iterations, double tolerance,
                         mv2_relax_result *out)
{
    if (!ctx) return set_error(nullptr, MV2_EINVAL, "[MiniSlope] null context");
    if (!height || !crit_diff || !out) return set_error(ctx, MV2_EINVAL, "[MiniSlope] null argument");
    if (ny < 2 || nx < 2)
        return set_error(ctx, MV2_EINVAL, "height must be a 2-D array with both dimensions >= 2");
    if (max_iterations < 1) return set_error(ctx, MV2_EINVAL, "[MiniSlope] max_iterations must be >= 1");
    if (precision != MV2_FP64 && precision != MV2_FP32)
        return set_error(ctx, MV2_EINVAL, "[MiniSlope] unknown precision %d", precision);
    hipSetDevice(ctx->device);
    return precision == MV2_FP64
               ? launch_relax<double>(ctx, height, ny, nx, crit_diff, max_iterations, tolerance, out)
               : launch_relax<float>(ctx, height, ny, nx, crit_diff, max_iterations, tolerance, out);
}

extern "C" int mv2_frontier_relax(mv2_context *ctx, double *height, const double *weights, int32_t *reached,
                                  const int32_t *a_i, const int32_t *a_j, const int32_t *b_i,
                                  const int32_t *b_j, int n_edges, int rows, int cols, int tile_size,
                                  int tiles_x, int tile_count, double critical_difference,
                                  int32_t *moved_tile_flags, int32_t *moved_count_out)
{
    if (!ctx) return set_error(nullptr, MV2_EINVAL, "[WarpMiniSlope] null context");
    if (n_edges < 0 || !moved_tile_flags || !moved_count_out || tile_size < 1 || tile_count < 1)
        return set_error(ctx, MV2_EINVAL, "[WarpMiniSlope] invalid phase arguments");
    hipSetDevice(ctx->device);
    int32_t *dcount = (int32_t *)((char *)ctx->result.ptr + 3072);
    hipLaunchKernelGGL(fill_i32, dim3(grid_for(tile_count)), dim3(kBlock), 0, ctx->stream, moved_tile_flags,
                       tile_count, 0);
    hipLaunchKernelGGL(fill_i32, dim3(1), dim3(64), 0, ctx->stream, dcount, 1, 0);
    if (n_edges > 0)
        hipLaunchKernelGGL(frontier_relax_kernel, dim3(grid_for(n_edges)), dim3(kBlock), 0, ctx->stream, height,
                           weights, reached, a_i, a_j, b_i, b_j, n_edges, rows, cols, tile_size, tiles_x,
                           critical_difference, moved_tile_flags, dcount);
    MV2_HIP(ctx, hipGetLastError());
    MV2_HIP(ctx, hipMemcpyAsync(ctx->host_result.ptr, dcount, sizeof(int32_t), hipMemcpyDeviceToHost, ctx->stream));
    if (int rc_w = stream_wait(ctx, ctx->stream)) return rc_w;
    memcpy(moved_count_out, ctx->host_result.ptr, sizeof(int32_t));
    return MV2_OK;
}

extern "C" int mv2_frontier_scan(mv2_context *ctx, const double *height, const int32_t *reached,
                                 const int32_t *tile_ids, const int32_t *a_i, const int32_t *a_j,
                                 const int32_t *b_i, const int32_t *b_j, int n_edges, int cols, int tile_count,
                                 double critical_difference, double tolerance, int32_t *unstable_tile_flags)
{
    if (!ctx) return set_error(nullptr, MV2_EINVAL, "[WarpMiniSlope] null context");
    if (n_edges < 0 || !unstable_tile_flags || tile_count < 1)
        return set_error(ctx, MV2_EINVAL, "[WarpMiniSlope] invalid phase arguments");
    hipSetDevice(ctx->device);
    hipLaunchKernelGGL(fill_i32, dim3(grid_for(tile_count)), dim3(kBlock), 0, ctx->stream, unstable_tile_flags,
                       tile_count, 0);
    if (n_edges > 0)
        hipLaunchKernelGGL(frontier_scan_kernel, dim3(grid_for(n_edges)), dim3(kBlock), 0, ctx->stream, height,
                           reached, tile_ids, a_i, a_j, b_i, b_j, n_edges, cols, critical_difference, tolerance,
                           unstable_tile_flags);
    MV2_HIP(ctx, hipGetLastError());
    return MV2_OK;
}
