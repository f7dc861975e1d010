// On-device compact-frontier rounds: the residual projection loop of
// V2PhysicsCore._advance_gpu_residual_frontier (pkg/runtime/v2_physics_core.py:1172-1254)
// with no host round trip per colour batch.
//
// A round runs, for each direction d and parity p (batch k = 2d + p), the
// weighted transfer of every edge owned by a cell of an active tile whose
// coordinate along d has parity p (the batches of
// pkg/solvers/sparse_tile_slope.py:116-230, conflict-free, so their result
// does not depend on thread order).  A tile touched by a transfer (the 3x3
// cell halo of both ends) joins the active set for the batches after k.
// Then every edge owned by a cell of the round's final active set is
// scanned; the tiles owning an edge still above the tolerance form the next
// round's active set.  The loop ends when that set is empty (frontier_reached
// is then cleared) or the round budget is spent.
//
// Device state, per tile: stamp[t] = first batch of the current round that
// sees t as active (0: active from the round's start, kInactive: not
// active).  A transfer in batch k lowers the stamps around it to k + 1 with
// atomicMin, so batch k never sees the tiles it activates -- the host loop
// builds each batch from the set as it stood before the batch.  Round r
// reads control record r & 1 and its last launch writes record (r + 1) & 1,
// so no launch reads a record it writes.  The host enqueues rounds in
// chunks and reads one record per chunk; rounds after the loop ended return
// at once.
#include "mv2_frontier_edge.hpp"
#include "mv2_internal.hpp"

#include <climits>
#include <cmath>
#include <cstring>

namespace mv2 {

constexpr int32_t kInactive = INT_MAX;
constexpr int kFrontierCells = kBlock;    // cells of one tile per workgroup: one per thread
constexpr int kFrontierMaxChunk = 64;     // rounds enqueued per host check, at most

struct FrontierControl {
    int active;                   // tiles active at the start of the round
    int owners;                   // unstable owner tiles found by the round's scan
    int done;                     // 1: no further round runs
    int rounds;                   // rounds run
    int cleared;                  // 1: a round ended with an empty active set
    int pad;
    unsigned long long moved;     // transfers over all rounds
};
static_assert(sizeof(FrontierControl) == 32, "control record layout");
constexpr size_t kFrontierCtlBytes = 256;  // both records, padded

struct FrontierArgs {
    double *height;
    const double *weights;
    int32_t *reached;
    int32_t *flags;               // in: active tiles at the call; out: active tiles at the end
    int32_t *stamp, *owner;
    FrontierControl *ctl;         // [2]
    int rows, cols, ts, tiles_x, tile_count, nd, budget;
    int tchunks;                  // workgroups per tile of a transfer batch (cells of one parity)
    int schunks;                  // workgroups per tile of the scan (every cell)
    int di[4], dj[4];
    double crit[4];
    double tolerance;
};

__device__ inline bool round_runs(const FrontierControl &c) { return !c.done && c.active > 0; }

// Activate the tiles of the 3x3 cell halo around (row, col) from batch
// `stamp` on.  The workgroup's own tile is active already (stamp <= k), so
// its atomics -- most of them, all on one address -- are skipped.
__device__ inline void stamp_tile_halo(const FrontierArgs &a, int row, int col, int32_t stamp, int own)
{
    for (int orow = -1; orow < 2; orow++)
        for (int ocol = -1; ocol < 2; ocol++) {
            const int nr = row + orow, nc = col + ocol;
            if (nr >= 0 && nr < a.rows && nc >= 0 && nc < a.cols) {
                const int tile = (nr / a.ts) * a.tiles_x + nc / a.ts;
                if (tile != own) atomicMin(&a.stamp[tile], stamp);
            }
        }
}

// stamps from the caller's flags, active count into record 0 (zeroed by the host)
__global__ void __launch_bounds__(kBlock) frontier_begin_kernel(FrontierArgs a)
{
    const int t = blockIdx.x * kBlock + threadIdx.x;
    int on = 0;
    if (t < a.tile_count) {
        on = a.flags[t] != 0;
        a.stamp[t] = on ? 0 : kInactive;
        a.owner[t] = 0;
    }
    const int n = wave_reduce<0>(on);
    if ((threadIdx.x & (kWave - 1)) == 0 && n) atomicAdd(&a.ctl[0].active, n);
}

// batch k = 2 d + p of round r
__global__ void __launch_bounds__(kBlock) frontier_transfer_kernel(FrontierArgs a, int r, int k)
{
    // no control-record check: after the loop ended every stamp is kInactive
    const int t = blockIdx.x / a.tchunks, chunk = blockIdx.x - t * a.tchunks;
    if (__hip_atomic_load(&a.stamp[t], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) > k) return;
    const int d = k >> 1, p = k & 1;
    const int di = a.di[d], dj = a.dj[d];
    const double crit = a.crit[d];
    const int tr = t / a.tiles_x, tc = t - tr * a.tiles_x;
    const int r0 = tr * a.ts, c0 = tc * a.ts;
    // only the cells whose coordinate along the direction has parity p: every
    // other row of the tile (di != 0) or every other column
    const int half = (a.ts + 1) >> 1;
    const int l = chunk * kFrontierCells + threadIdx.x;
    int moved = 0;
    if (l < half * a.ts) {
        int li, lj;
        if (di) {
            const int m = l / a.ts;
            lj = l - m * a.ts;
            li = ((p - r0) & 1) + 2 * m;
        } else {
            li = l / half;
            lj = ((p - c0) & 1) + 2 * (l - li * half);
        }
        const int ai = r0 + li, aj = c0 + lj;
        const int bi = ai + di, bj = aj + dj;
        if (li < a.ts && lj < a.ts && ai < a.rows && aj < a.cols && bi >= 0 && bi < a.rows && bj >= 0 &&
            bj < a.cols &&
            frontier_edge_transfer(a.height, a.weights, a.reached, (int64_t)ai * a.cols + aj,
                                   (int64_t)bi * a.cols + bj, crit)) {
            moved = 1;
            stamp_tile_halo(a, ai, aj, k + 1, t);
            stamp_tile_halo(a, bi, bj, k + 1, t);
        }
    }
    const int total = wave_reduce<0>(moved);
    if ((threadIdx.x & (kWave - 1)) == 0 && total)
        atomicAdd(&a.ctl[r & 1].moved, (unsigned long long)total);
}

// scan of every edge owned by the round's final active set (stamp <= 2 nd)
__global__ void __launch_bounds__(kBlock) frontier_scan_round_kernel(FrontierArgs a, int r)
{
    const int t = blockIdx.x / a.schunks, chunk = blockIdx.x - t * a.schunks;
    if (a.stamp[t] > 2 * a.nd) return;
    const int tr = t / a.tiles_x, tc = t - tr * a.tiles_x;
    const int l = chunk * kFrontierCells + threadIdx.x;
    const int li = l / a.ts, lj = l - li * a.ts;
    const int ai = tr * a.ts + li, aj = tc * a.ts + lj;
    int unstable = 0;
    if (l < a.ts * a.ts && ai < a.rows && aj < a.cols) {
        for (int d = 0; d < a.nd; d++) {
            const int bi = ai + a.di[d], bj = aj + a.dj[d];
            if (bi < 0 || bi >= a.rows || bj < 0 || bj >= a.cols) continue;
            if (frontier_edge_unstable(a.height, a.reached, (int64_t)ai * a.cols + aj,
                                       (int64_t)bi * a.cols + bj, a.crit[d], a.tolerance)) {
                unstable = 1;
                break;
            }
        }
    }
    if (__syncthreads_or(unstable) && threadIdx.x == 0)
        if (atomicExch(&a.owner[t], 1) == 0) atomicAdd(&a.ctl[r & 1].owners, 1);
}

// next active set = the scan's owners; record (r + 1) & 1
__global__ void __launch_bounds__(kBlock) frontier_round_end_kernel(FrontierArgs a, int r)
{
    const FrontierControl c = load_coherent(&a.ctl[r & 1]);
    const bool ran = round_runs(c);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        FrontierControl n = c;
        n.owners = 0;
        if (ran) {
            n.rounds = c.rounds + 1;
            n.active = c.owners;
            n.cleared = c.owners == 0;
            n.done = c.owners == 0 || n.rounds >= a.budget;
        }
        a.ctl[(r + 1) & 1] = n;
    }
    if (!ran) return;
    const int t = blockIdx.x * kBlock + threadIdx.x;
    if (t < a.tile_count) {
        const int o = a.owner[t];
        a.stamp[t] = o ? 0 : kInactive;
        a.flags[t] = o;
        a.owner[t] = 0;
    }
}

}  // namespace mv2

using namespace mv2;

extern "C" int mv2_frontier_rounds(mv2_context *ctx, const mv2_frontier_round_config *cfg, double *height,
                                   const double *weights, int32_t *reached, int32_t *active_tile_flags,
                                   mv2_frontier_round_result *out)
{
    if (!ctx) return set_error(nullptr, MV2_EINVAL, "[WarpMiniSlope] null context");
    if (!cfg || !out || !height || !weights || !reached || !active_tile_flags)
        return set_error(ctx, MV2_EINVAL, "[WarpMiniSlope] null round argument");
    const int rows = cfg->rows, cols = cfg->cols, ts = cfg->tile_size;
    if (rows < 1 || cols < 1 || ts < 1)
        return set_error(ctx, MV2_EINVAL, "[WarpMiniSlope] invalid frontier shape");
    const int tiles_y = (rows + ts - 1) / ts, tiles_x = (cols + ts - 1) / ts;
    if (cfg->tiles_x != tiles_x || cfg->tile_count != tiles_x * tiles_y)
        return set_error(ctx, MV2_EINVAL, "[WarpMiniSlope] tile grid does not match the shape");
    if (cfg->n_directions < 1 || cfg->n_directions > 4)
        return set_error(ctx, MV2_EINVAL, "[WarpMiniSlope] 1 to 4 neighbour directions");
    for (int d = 0; d < cfg->n_directions; d++) {
        const int di = cfg->di[d], dj = cfg->dj[d];
        if (di < -1 || di > 1 || dj < -1 || dj > 1 || (di == 0 && dj == 0) ||
            !std::isfinite(cfg->critical_difference[d]))
            return set_error(ctx, MV2_EINVAL, "[WarpMiniSlope] invalid neighbour direction %d", d);
    }
    if (!std::isfinite(cfg->tolerance) || cfg->max_rounds < 0)
        return set_error(ctx, MV2_EINVAL, "[WarpMiniSlope] invalid tolerance/round budget");
    hipSetDevice(ctx->device);
    const int tile_count = cfg->tile_count;
    const size_t bytes = kFrontierCtlBytes + 2 * sizeof(int32_t) * (size_t)tile_count;
    if (int rc = ensure_scratch(ctx, ctx->frontier, bytes)) return rc;

    FrontierArgs a{};
    a.height = height;
    a.weights = weights;
    a.reached = reached;
    a.flags = active_tile_flags;
    a.ctl = (FrontierControl *)ctx->frontier.ptr;
    a.stamp = (int32_t *)((char *)ctx->frontier.ptr + kFrontierCtlBytes);
    a.owner = a.stamp + tile_count;
    a.rows = rows;
    a.cols = cols;
    a.ts = ts;
    a.tiles_x = tiles_x;
    a.tile_count = tile_count;
    a.tchunks = (((ts + 1) / 2) * ts + kFrontierCells - 1) / kFrontierCells;
    a.schunks = (ts * ts + kFrontierCells - 1) / kFrontierCells;
    a.nd = cfg->n_directions;
    a.budget = cfg->max_rounds;
    for (int d = 0; d < 4; d++) {
        a.di[d] = d < a.nd ? cfg->di[d] : 0;
        a.dj[d] = d < a.nd ? cfg->dj[d] : 0;
        a.crit[d] = d < a.nd ? cfg->critical_difference[d] : 0.0;
    }
    a.tolerance = cfg->tolerance;

    const int64_t transfer_grid = (int64_t)tile_count * a.tchunks, scan_grid = (int64_t)tile_count * a.schunks;
    if (scan_grid > INT_MAX || (int64_t)ts * ts > INT_MAX)
        return set_error(ctx, MV2_EINVAL, "[WarpMiniSlope] frontier grid too large");
    const int tile_grid = grid_for(tile_count);
    MV2_HIP(ctx, hipMemsetAsync(a.ctl, 0, kFrontierCtlBytes, ctx->stream));
    frontier_begin_kernel<<<tile_grid, kBlock, 0, ctx->stream>>>(a);
    MV2_HIP(ctx, hipGetLastError());

    FrontierControl rec{};
    int enqueued = 0;
    int chunk = ctx->frontier_predicted > 0 ? ctx->frontier_predicted : 2;
    for (;;) {
        const int n = a.budget - enqueued < chunk ? a.budget - enqueued : chunk;
        for (int i = 0; i < n; i++) {
            const int r = enqueued + i;
            for (int k = 0; k < 2 * a.nd; k++)
                frontier_transfer_kernel<<<(int)transfer_grid, kBlock, 0, ctx->stream>>>(a, r, k);
            frontier_scan_round_kernel<<<(int)scan_grid, kBlock, 0, ctx->stream>>>(a, r);
            frontier_round_end_kernel<<<tile_grid, kBlock, 0, ctx->stream>>>(a, r);
        }
        MV2_HIP(ctx, hipGetLastError());
        enqueued += n;
        MV2_HIP(ctx, hipMemcpyAsync(ctx->host_result.ptr, a.ctl + (enqueued & 1), sizeof(FrontierControl),
                                    hipMemcpyDeviceToHost, ctx->stream));
        if (int rc_w = stream_wait(ctx, ctx->stream)) return rc_w;
        memcpy(&rec, ctx->host_result.ptr, sizeof(FrontierControl));
        if (rec.done || rec.active == 0 || enqueued >= a.budget) break;
        chunk = chunk * 2 < kFrontierMaxChunk ? chunk * 2 : kFrontierMaxChunk;
    }
    ctx->frontier_predicted = rec.rounds + 1 < kFrontierMaxChunk ? rec.rounds + 1 : kFrontierMaxChunk;
    if (rec.cleared)
        MV2_HIP(ctx, hipMemsetAsync(reached, 0, sizeof(int32_t) * (size_t)rows * cols, ctx->stream));
    out->rounds = rec.rounds;
    out->converged = rec.active == 0;
    out->active_tiles = rec.active;
    out->cleared_reached = rec.cleared;
    out->moved_edges = (long long)rec.moved;
    return MV2_OK;
}
