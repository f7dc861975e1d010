// Angle-of-repose relaxation (slope_model.relax_critical_slope,
// slope_model.py:273-343) and the compact colour-batch sweep
// (pkg/solvers/warp_compact_frontier.py:57-130).
//
// relax: one cooperative launch runs the whole data-dependent iteration loop.
// Each workgroup stages a tile plus an 8-cell halo in LDS and performs all 8
// red-black sub-phases of an iteration there.  A sub-phase moves information
// one cell along its direction, so after the 8 sub-phases (two per direction,
// directions x, y and both diagonals) at most 6 halo layers are stale; the
// interior and every edge it owns are exact.  Then the interior is written
// (ping-pong buffers), the per-iteration maximum excess is reduced with one
// atomicMax per workgroup, and one grid barrier decides convergence.  So an
// iteration costs one read + one write of the field instead of eight
// full-array passes, and the host is not involved per iteration.
//
// Directions and parities follow the reference's legacy H[x, y] convention
// behind MinimumSlopeAdapter (pkg/solvers/minimum_slope_adapter.py:170-179):
// on the canonical H[y, x] field with view index (r, c):
//   d0 legacy (1,0):  a = (r, c),   b = (r, c+1),   parity on c, dist dx
//   d1 legacy (0,1):  a = (r, c),   b = (r+1, c),   parity on r, dist dy
//   d2 legacy (1,1):  a = (r, c),   b = (r+1, c+1), parity on c, dist hypot
//   d3 legacy (1,-1): a = (r+1, c), b = (r, c+1),   parity on c, dist hypot
#include "mv2_internal.hpp"

#include <cmath>
#include <cstring>

namespace mv2 {

constexpr int kRelaxHalo = 8;
constexpr int kRelaxTH = 32;   // interior rows per tile
constexpr int kRelaxTW = 48;   // interior columns per tile
constexpr int kRelaxLH = kRelaxTH + 2 * kRelaxHalo;
constexpr int kRelaxLW = kRelaxTW + 2 * kRelaxHalo;

struct RelaxSync {
    unsigned counter;
    unsigned pad;
    unsigned long long slot[3];
};

template <class R>
struct RelaxArgs {
    R *h[2];
    int ny, nx;
    R crit[4];
    int max_iterations;
    double tolerance;
    RelaxSync *sync;
    mv2_relax_result *out;
    int tiles_x, tiles;
};

__device__ inline void edge_cells(int d, int r, int c, int &ra, int &ca, int &rb, int &cb)
{
    switch (d) {
    case 0: ra = r; ca = c; rb = r; cb = c + 1; break;
    case 1: ra = r; ca = c; rb = r + 1; cb = c; break;
    case 2: ra = r; ca = c; rb = r + 1; cb = c + 1; break;
    default: ra = r + 1; ca = c; rb = r; cb = c + 1; break;
    }
}

template <class R>
__device__ inline R excess_of(R diff, R crit)
{
    R ex = (R)fabs((double)diff) - crit;
    return ex > (R)0 ? ex : (R)0;
}

template <class R>
__global__ void __launch_bounds__(kBlock) relax_kernel(RelaxArgs<R> a)
{
    __shared__ R tile[kRelaxLH * kRelaxLW];
    __shared__ double smem[8];
    unsigned epoch = 0;
    const int ny = a.ny, nx = a.nx;
    int src = 0;
    int iteration = 0;
    int converged = 0;
    double last = 0.0;
    for (iteration = 1; iteration <= a.max_iterations; iteration++) {
        if (blockIdx.x == 0 && threadIdx.x == 0)
            __hip_atomic_store(&a.sync->slot[(iteration + 1) % 3], 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        double wg_max = 0.0;
        for (int t = blockIdx.x; t < a.tiles; t += gridDim.x) {
            const int r0 = (t / a.tiles_x) * kRelaxTH, c0 = (t % a.tiles_x) * kRelaxTW;
            const int R0 = r0 - kRelaxHalo, C0 = c0 - kRelaxHalo;
            // stage tile + halo (cells outside the grid are never touched)
            for (int k = threadIdx.x; k < kRelaxLH * kRelaxLW; k += blockDim.x) {
                int lr = k / kRelaxLW, lc = k - lr * kRelaxLW;
                int gr = R0 + lr, gc = C0 + lc;
                R v = (R)0;
                if (gr >= 0 && gr < ny && gc >= 0 && gc < nx) v = a.h[src][(int64_t)gr * nx + gc];
                tile[k] = v;
            }
            __syncthreads();
            R tmax = (R)0;
            for (int d = 0; d < 4; d++) {
                const R crit = a.crit[d];
                for (int phase = 0; phase < 2; phase++) {
                    // maximum excess over the edges this tile owns (pre-phase state)
                    for (int k = threadIdx.x; k < kRelaxTH * kRelaxTW; k += blockDim.x) {
                        int ir = k / kRelaxTW, ic = k - ir * kRelaxTW;
                        int r = r0 + ir, c = c0 + ic;
                        int ra, ca, rb, cb;
                        edge_cells(d, r, c, ra, ca, rb, cb);
                        if (r >= ny || c >= nx || ra >= ny || rb >= ny || ca >= nx || cb >= nx) continue;
                        R va = tile[(ra - R0) * kRelaxLW + (ca - C0)];
                        R vb = tile[(rb - R0) * kRelaxLW + (cb - C0)];
                        R ex = excess_of<R>(va - vb, crit);
                        tmax = tmax >= ex ? tmax : ex;
                    }
                    __syncthreads();
                    // red-black transfer on the active parity (disjoint edges)
                    for (int k = threadIdx.x; k < kRelaxLH * kRelaxLW; k += blockDim.x) {
                        int lr = k / kRelaxLW, lc = k - lr * kRelaxLW;
                        int r = R0 + lr, c = C0 + lc;
                        int coord = d == 1 ? r : c;
                        if (r < 0 || c < 0 || ((coord & 1) != phase)) continue;
                        int ra, ca, rb, cb;
                        edge_cells(d, r, c, ra, ca, rb, cb);
                        if (ra >= ny || rb >= ny || ca >= nx || cb >= nx) continue;
                        int la_r = ra - R0, la_c = ca - C0, lb_r = rb - R0, lb_c = cb - C0;
                        if (la_r < 0 || la_r >= kRelaxLH || lb_r < 0 || lb_r >= kRelaxLH ||
                            la_c < 0 || la_c >= kRelaxLW || lb_c < 0 || lb_c >= kRelaxLW)
                            continue;
                        R &va = tile[la_r * kRelaxLW + la_c];
                        R &vb = tile[lb_r * kRelaxLW + lb_c];
                        R diff = va - vb;
                        R ex = excess_of<R>(diff, crit);
                        R sgn = diff > (R)0 ? (R)1 : (diff < (R)0 ? (R)-1 : (R)0);
                        R tr = ((R)0.5 * ex) * sgn;
                        va = va - tr;
                        vb = vb + tr;
                    }
                    __syncthreads();
                }
            }
            wg_max = wg_max >= (double)tmax ? wg_max : (double)tmax;
            for (int k = threadIdx.x; k < kRelaxTH * kRelaxTW; k += blockDim.x) {
                int ir = k / kRelaxTW, ic = k - ir * kRelaxTW;
                int r = r0 + ir, c = c0 + ic;
                if (r < ny && c < nx)
                    a.h[src ^ 1][(int64_t)r * nx + c] = tile[(ir + kRelaxHalo) * kRelaxLW + ic + kRelaxHalo];
            }
            __syncthreads();
        }
        double m[1] = {wg_max};
        block_reduce<double, 1, 1>(m, smem);
        if (threadIdx.x == 0) atomic_max_nonneg(&a.sync->slot[iteration % 3], m[0]);
        grid_barrier(&a.sync->counter, epoch);
        src ^= 1;
        last = __longlong_as_double((long long)__hip_atomic_load(&a.sync->slot[iteration % 3], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
        if (last <= a.tolerance) { converged = 1; break; }
    }
    if (iteration > a.max_iterations) iteration = a.max_iterations;
    if (src == 1) {
        const int64_t n = (int64_t)ny * nx;
        for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
             i += (int64_t)gridDim.x * blockDim.x)
            a.h[0][i] = a.h[1][i];
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        a.out->iterations = iteration;
        a.out->converged = converged;
        a.out->max_excess_last = last;
    }
}

__global__ void zero_relax_sync(RelaxSync *s)
{
    if (threadIdx.x == 0) {
        s->counter = 0;
        s->slot[0] = s->slot[1] = s->slot[2] = 0ull;
    }
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
    a.out = (mv2_relax_result *)((char *)ctx->result.ptr + 2048);
    a.tiles_x = (nx + kRelaxTW - 1) / kRelaxTW;
    a.tiles = a.tiles_x * ((ny + kRelaxTH - 1) / kRelaxTH);
    int per_cu = 0;
    MV2_HIP(ctx, hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, relax_kernel<R>, kBlock, 0));
    if (per_cu < 1) return set_error(ctx, MV2_EDEVICE, "[MiniSlope] relaxation kernel cannot be resident");
    int grid = per_cu * ctx->cu_count;
    if (grid > a.tiles) grid = a.tiles;
    hipLaunchKernelGGL(zero_relax_sync, dim3(1), dim3(64), 0, ctx->stream, a.sync);
    MV2_HIP(ctx, hipGetLastError());
    void *args[] = {&a};
    MV2_HIP(ctx, hipLaunchCooperativeKernel((const void *)relax_kernel<R>, dim3(grid), dim3(kBlock), args, 0,
                                            ctx->stream));
    MV2_HIP(ctx, hipMemcpyAsync(ctx->host_result.ptr, a.out, sizeof(mv2_relax_result), hipMemcpyDeviceToHost,
                                ctx->stream));
    MV2_HIP(ctx, hipStreamSynchronize(ctx->stream));
    memcpy(out, ctx->host_result.ptr, sizeof(mv2_relax_result));
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
        if (reached[a] == 0 && reached[b] == 0) continue;
        double difference = height[a] - height[b];
        double excess = fabs(difference) - critical_difference;
        if (excess > 0.0) {
            double signed_excess = difference < 0.0 ? -excess : excess;
            double wa = weights[a], wb = weights[b];
            double ws = wa + wb;
            double ta = signed_excess * wb / ws;
            double tb = signed_excess * wa / ws;
            height[a] = height[a] - ta;
            height[b] = height[b] + tb;
            reached[a] = 1;
            reached[b] = 1;
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
        if (reached[a] != 0 || reached[b] != 0) {
            double excess = fabs(height[a] - height[b]) - critical_difference;
            if (excess > tolerance) atomicMax(&flags[tile_ids[e]], 1);
        }
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
    MV2_HIP(ctx, hipStreamSynchronize(ctx->stream));
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
