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
// full-array passes.
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

// cells of edge direction d anchored at (r, c): a and b as in the table above
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

// Per-thread work of the LDS phases has a fixed trip count: the tile is
// kRelaxLH x kRelaxLW = 3072 cells, the owned edges of one direction are the
// 1536 interior anchors, and one parity selects 1536 anchors of the halo tile.
// 512 threads (8 waves) per tile: the kernel is latency-bound on LDS and
// barriers, so more waves per tile hide more of it at the same total work.
constexpr int kRelaxBlock = 512;
constexpr int kRelaxStage = kRelaxLH * kRelaxLW / kRelaxBlock;           // 6
constexpr int kRelaxOwned = kRelaxTH * kRelaxTW / kRelaxBlock;           // 3
constexpr int kRelaxActive = kRelaxLH * kRelaxLW / 2 / kRelaxBlock;      // 3
static_assert(kRelaxLH * kRelaxLW % kRelaxBlock == 0 && kRelaxTH * kRelaxTW % kRelaxBlock == 0 &&
              kRelaxLH * kRelaxLW / 2 % kRelaxBlock == 0, "relaxation tile must split evenly over the block");
static_assert(kRelaxHalo % 2 == 0 && kRelaxTH % 2 == 0 && kRelaxTW % 2 == 0,
              "tile origins must be even so local and global parities agree");

// One relaxation iteration (slope_model.py:296-338) per launch; see the file
// comment.  Every phase first loads all the operands a thread needs, then
// computes, then stores: the edges of one parity are disjoint, so the loads
// of one thread's edges never depend on its own stores.
template <class R>
__global__ void __launch_bounds__(kRelaxBlock) relax_kernel(RelaxArgs<R> a, int k)
{
    __shared__ R tile[kRelaxLH * kRelaxLW];
    __shared__ double smem[8];
    const int ny = a.ny, nx = a.nx;
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
    const int r0 = (t / a.tiles_x) * kRelaxTH, c0 = (t % a.tiles_x) * kRelaxTW;
    const int R0 = r0 - kRelaxHalo, C0 = c0 - kRelaxHalo;
    // stage tile + halo in one round trip (cells outside the grid are never touched)
    {
        R v[kRelaxStage];
#pragma unroll
        for (int j = 0; j < kRelaxStage; j++) {
            const int e = threadIdx.x + j * kRelaxBlock;
            const int gr = R0 + e / kRelaxLW, gc = C0 + e % kRelaxLW;
            v[j] = (R)0;
            if (gr >= 0 && gr < ny && gc >= 0 && gc < nx) v[j] = a.h[src][gr * nx + gc];
        }
#pragma unroll
        for (int j = 0; j < kRelaxStage; j++) tile[threadIdx.x + j * kRelaxBlock] = v[j];
    }
    __syncthreads();
    R tmax = (R)0;
#pragma unroll 1
    for (int d = 0; d < 4; d++) {
        const R crit = a.crit[d];
#pragma unroll 1
        for (int phase = 0; phase < 2; phase++) {
            // maximum excess over every edge of direction d that this tile owns
            // (anchor in the interior), in the state before the sub-phase
            {
                R va[kRelaxOwned], vb[kRelaxOwned];
                bool ok[kRelaxOwned];
#pragma unroll
                for (int j = 0; j < kRelaxOwned; j++) {
                    const int e = threadIdx.x + j * kRelaxBlock;
                    const int r = r0 + e / kRelaxTW, c = c0 + e % kRelaxTW;
                    int ra, ca, rb, cb;
                    edge_cells(d, r, c, ra, ca, rb, cb);
                    ok[j] = r < ny && c < nx && ra < ny && rb < ny && ca < nx && cb < nx;
                    va[j] = tile[(ra - R0) * kRelaxLW + (ca - C0)];
                    vb[j] = tile[(rb - R0) * kRelaxLW + (cb - C0)];
                }
#pragma unroll
                for (int j = 0; j < kRelaxOwned; j++) {
                    const R ex = ok[j] ? excess_of<R>(va[j] - vb[j], crit) : (R)0;
                    tmax = tmax >= ex ? tmax : ex;
                }
            }
            __syncthreads();
            // red-black transfer on the active parity (disjoint edges); the
            // parity coordinate is c (d = 0, 2, 3) or r (d = 1)
            {
                int ia[kRelaxActive], ib[kRelaxActive];
                R va[kRelaxActive], vb[kRelaxActive];
                bool ok[kRelaxActive];
#pragma unroll
                for (int j = 0; j < kRelaxActive; j++) {
                    const int e = threadIdx.x + j * kRelaxBlock;
                    int lr, lc;
                    if (d == 1) {
                        lr = 2 * (e / kRelaxLW) + phase;
                        lc = e % kRelaxLW;
                    } else {
                        lr = e / (kRelaxLW / 2);
                        lc = 2 * (e % (kRelaxLW / 2)) + phase;
                    }
                    const int r = R0 + lr, c = C0 + lc;
                    int ra, ca, rb, cb;
                    edge_cells(d, r, c, ra, ca, rb, cb);
                    const int la_r = ra - R0, la_c = ca - C0, lb_r = rb - R0, lb_c = cb - C0;
                    ok[j] = r >= 0 && c >= 0 && ra < ny && rb < ny && ca < nx && cb < nx && la_r >= 0 &&
                            la_r < kRelaxLH && lb_r >= 0 && lb_r < kRelaxLH && la_c >= 0 && la_c < kRelaxLW &&
                            lb_c >= 0 && lb_c < kRelaxLW;
                    ia[j] = ok[j] ? la_r * kRelaxLW + la_c : 0;
                    ib[j] = ok[j] ? lb_r * kRelaxLW + lb_c : 0;
                    va[j] = tile[ia[j]];
                    vb[j] = tile[ib[j]];
                }
#pragma unroll
                for (int j = 0; j < kRelaxActive; j++) {
                    if (!ok[j]) continue;
                    const R diff = va[j] - vb[j];
                    const R ex = excess_of<R>(diff, crit);
                    const R sgn = diff > (R)0 ? (R)1 : (diff < (R)0 ? (R)-1 : (R)0);
                    const R tr = ((R)0.5 * ex) * sgn;
                    tile[ia[j]] = va[j] - tr;
                    tile[ib[j]] = vb[j] + tr;
                }
            }
            __syncthreads();
        }
    }
    // write the interior
#pragma unroll
    for (int j = 0; j < kRelaxOwned; j++) {
        const int e = threadIdx.x + j * kRelaxBlock;
        const int ir = e / kRelaxTW, ic = e % kRelaxTW;
        const int r = r0 + ir, c = c0 + ic;
        if (r < ny && c < nx) a.h[src ^ 1][r * nx + c] = tile[(ir + kRelaxHalo) * kRelaxLW + ic + kRelaxHalo];
    }
    double m[1] = {(double)tmax};
    block_reduce<double, 1, 1, kRelaxBlock / kWave>(m, smem);
    if (threadIdx.x == 0) spread_max(a.spread + (k % 3) * kSpreadWords, m[0]);
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
        MV2_HIP(ctx, hipStreamSynchronize(st));
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
