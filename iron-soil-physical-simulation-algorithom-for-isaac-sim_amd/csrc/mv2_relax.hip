// Angle-of-repose relaxation (slope_model.relax_critical_slope,
// slope_model.py:273-343) and the compact colour-batch sweep
// (pkg/solvers/warp_compact_frontier.py:57-130).
//
// relax: a launch applies up to K iterations to LDS tiles (tile interior
// 44 x 48, halo 6K + 2; mv2_relax_tile.hpp) and writes the field once
// (ping-pong buffers).  The first launch of a call applies one iteration (a
// field at rest converges there), later FP32 launches two when the whole grid
// runs at once (one otherwise).  Each iteration's
// maximum excess over the owned edges is reduced into its own spread-maximum
// set; the next launch tests them in order against the tolerance on the
// device.  The reference returns the field of the first converged iteration:
// when that is the first of a launch's two, the following launch re-applies
// exactly that one iteration to the launch's input buffer (still intact) and
// ends the loop.  Launches after the loop's end return at once, so the host
// enqueues batches of launches and reads one result per batch.  A tile whose
// 3 x 3 tile neighbourhood did not change in the previous launch, which
// applied as many iterations as this one, is skipped: its field and maxima
// are already known.
#include "mv2_internal.hpp"
#include "mv2_frontier_edge.hpp"
#include "mv2_relax_tile.hpp"

#include <cmath>
#include <cstring>

namespace mv2 {

// Iteration state: launch k reads ctl[k & 1] (launch 0 starts from a fresh
// state and reads nothing) and its workgroup 0 writes ctl[(k + 1) & 1].
// Launch k reduces the maxima of its iterations into spread group g(k) (set
// 2 g + i for its iteration i), tests group g(k - 1) (the previous launch)
// against the tolerance and clears group g(k + 1).  g(k) = k % 3 for k > 0;
// launch 0 uses group 3 or 4 by the parity of the call, and every end pass of
// a call clears the other one for the next call, so a call needs no reset
// launch (only the first call, or one after an error, zeroes everything).
constexpr int kRelaxKMax = 2;                 // iterations per launch at most
constexpr int kRelaxGroups = 3;               // rotating groups of launches k > 0
static_assert((kRelaxGroups + 2) * kRelaxKMax <= kRelaxSpreadSets, "the relaxation's spread sets");
using RelaxFirst = RelaxGeom<44, 48, 640, 1, false>;      // FP32: a call's first launch
using RelaxPair = RelaxGeom<44, 48, 1024, 2, true>;       // FP32: later launches
using RelaxSingle64 = RelaxGeom<44, 48, 960, 1, true>;     // FP64: every launch (spill-free at 960)
static_assert(RelaxFirst::TH == RelaxPair::TH && RelaxFirst::TW == RelaxPair::TW &&
              RelaxFirst::TH == RelaxSingle64::TH && RelaxFirst::TW == RelaxSingle64::TW,
              "all launches of a call share one tile grid");
constexpr int kRelaxTileH = RelaxFirst::TH, kRelaxTileW = RelaxFirst::TW;

struct RelaxControl {
    int iterations;      // iterations applied to the field in buffer `src`
    int src;             // buffer holding the current field
    int done, converged;
    int n_prev;          // iterations of the previous launch whose maxima are untested
    int prev_src;        // input buffer of the previous launch
    int prev_iterations; // iterations applied before the previous launch
    int replay;          // > 0: re-apply this many iterations to `src`, then end the loop
    int marked_set;      // the maximum excess of iteration `mark` has been tested
    double last;         // maximum excess of the last applied iteration
    double marked;       // maximum excess of iteration `mark`
};
struct RelaxSync {
    RelaxControl ctl[2];
};

// the end pass's record in pinned host memory
struct RelaxHostOut {
    mv2_relax_result r;
    double marked;       // maximum excess of iteration `mark`, or of the last one if the loop ended before
};

template <class R>
struct RelaxArgs {
    R *h[2];
    int ny, nx;
    R crit[4];
    int max_iterations;
    double tolerance;
    RelaxSync *sync;
    unsigned long long *spread;   // (kRelaxGroups + 2) x kRelaxKMax spread-maximum sets
    int first_group;              // group of launch 0 (3 or 4); the end pass clears the other
    int mark;                     // iteration whose maximum excess is reported (0: none)
    RelaxHostOut *out;            // pinned host memory: the end pass writes the result there
    int tiles_x, tiles;
    int32_t *tchanged;            // [2][tiles]: the tile's interior changed in launch k (parity k & 1)
    int32_t *titers;              // [tiles]: iterations of the tile's last computed launch
    double *texcess;              // [tiles][kRelaxKMax]: its per-iteration maxima there
};

__device__ inline unsigned long long *relax_set(unsigned long long *spread, int group, int i)
{
    return spread + (size_t)(group * kRelaxKMax + i) * kSpreadWords;
}

template <class R>
__device__ inline int relax_group(const RelaxArgs<R> &a, int k)
{
    return k == 0 ? a.first_group : k % kRelaxGroups;
}

// this lane's slot of each spread set the previous launch reduced into, in
// every wave (loaded before the loop state, whose test reads them; every
// wave folds the 64 slots itself, so the test needs no barrier)
template <class R>
__device__ inline void relax_prev_slots(const RelaxArgs<R> &a, int k, double (&slot)[kRelaxKMax])
{
    const int g = relax_group(a, k > 0 ? k - 1 : 0);
    const int lane = threadIdx.x & (kWave - 1);
#pragma unroll
    for (int i = 0; i < kRelaxKMax; i++) {
        const double v = __longlong_as_double((long long)relax_set(a.spread, g, i)[lane * kSpreadStride]);
        slot[i] = k > 0 ? v : 0.0;
    }
}

// loop state for launch k: test the previous launch's iterations in order
// (slot: relax_prev_slots).  Each wave reaches the same decision from the
// same slots (a fixed DPP tree), with no barrier.
template <class R>
__device__ inline RelaxControl relax_resolve(const RelaxArgs<R> &a, int k, const double (&slot)[kRelaxKMax])
{
    RelaxControl C;
    if (k == 0) {
        C = RelaxControl{};
        return C;
    }
    C = load_launch_record(&a.sync->ctl[k & 1]);     // written by an earlier launch
    if (!C.done && C.replay == 0 && C.n_prev > 0) {
        int first = -1;
        double mfirst = 0.0, mlast = 0.0;
        for (int i = 0; i < C.n_prev; i++) {
            double m = 0.0;
#pragma unroll
            for (int j = 0; j < kRelaxKMax; j++)
                if (j == i) m = slot[j];
            m = wave_max(m);
            if (first < 0 && m <= a.tolerance) {
                first = i;
                mfirst = m;
            }
            mlast = m;
            if (C.prev_iterations + i + 1 == a.mark && (first < 0 || first == i)) {
                C.marked = m;
                C.marked_set = 1;
            }
        }
        if (first == C.n_prev - 1) {
            C.converged = 1;
            C.done = 1;
            C.last = mfirst;
        } else if (first >= 0) {
            // converged inside the previous launch: redo its first iterations
            C.src = C.prev_src;
            C.iterations = C.prev_iterations;
            C.replay = first + 1;
            C.last = mfirst;
        } else {
            C.last = mlast;
            if (C.iterations >= a.max_iterations) C.done = 1;
        }
        C.n_prev = 0;
    }
    return C;
}

template <class R, class G>
__global__ void __launch_bounds__(G::BLOCK) relax_kernel(RelaxArgs<R> a, int k)
{
    __shared__ R tile[G::kCells];
    __shared__ double smem[kRelaxKMax * (G::BLOCK / kWave)];
    __shared__ int wflag[G::BLOCK / kWave];
    RelaxSync *s = a.sync;
    const bool lead = blockIdx.x == 0 && threadIdx.x == 0;
    const int t = blockIdx.x;
    const int trow = t / a.tiles_x, tcol = t % a.tiles_x;
    const int r0 = trow * G::TH, c0 = tcol * G::TW;
    // Requests that do not depend on the loop state, issued together so that
    // they share one memory round trip: the previous launch's maxima, the
    // quiet-tile flags of the 3 x 3 tile neighbourhood and this tile's
    // iteration count there, and the tile itself from buffer k & 1 -- the
    // input of launch k whenever every launch before it applied iterations
    // (a replay reads the other buffer and reloads below).
    double slot[kRelaxKMax];
    relax_prev_slots(a, k, slot);
    // Every load of the prologue is unconditional (clamped indices, values
    // selected afterwards), so the prologue is one block whose waits can
    // count on loads returning in issue order: loads under branches are
    // waited for in full at the join.  Lanes 0..8 of every wave read the
    // changed flags of the 3 x 3 tile neighbourhood; each wave tests them
    // with a ballot (no barrier).
    const int lane = threadIdx.x & (kWave - 1);
    int qraw;
    bool qown;
    {
        const int ty = a.tiles / a.tiles_x;
        const int q = lane % 9;
        const int rr = trow + q / 3 - 1, cc = tcol + q % 3 - 1;
        qraw = a.tchanged[((k - 1) & 1) * a.tiles + min(max(rr, 0), ty - 1) * a.tiles_x + min(max(cc, 0), a.tiles_x - 1)];
        qown = lane < 9 && rr >= 0 && rr < ty && cc >= 0 && cc < a.tiles_x;
    }
    const int titer_raw = a.titers[t];
    R v[G::kStage];
    unsigned vmask = relax_tile_load<R, G>(a.h[k & 1], a.ny, a.nx, r0, c0, v);
    const RelaxControl C = relax_resolve(a, k, slot);
    if (C.done) {
        if (lead) s->ctl[(k + 1) & 1] = C;
        return;
    }
    const int left = a.max_iterations - C.iterations;
    const int n = C.replay > 0 ? C.replay : (G::K < left ? G::K : left);
    const int src = C.src;
    if (blockIdx.x == 0)
        for (int i = 0; i < kRelaxKMax; i++) spread_clear(relax_set(a.spread, relax_group(a, k + 1), i));
    if (lead) {
        RelaxControl nc = C;
        nc.prev_src = src;
        nc.prev_iterations = C.iterations;
        nc.iterations = C.iterations + n;
        nc.src = src ^ 1;
        nc.n_prev = n;
        if (C.replay > 0) {      // the re-applied iterations end at the converged one
            nc.replay = 0;
            nc.n_prev = 0;
            nc.done = 1;
            nc.converged = 1;
        }
        s->ctl[(k + 1) & 1] = nc;
    }
    unsigned long long *sets = relax_set(a.spread, relax_group(a, k), 0);
    // Quiet tiles: the iterations of a tile read its interior and a halo of
    // at most one tile, i.e. its 3 x 3 tile neighbourhood.  When none of
    // those tiles changed in launch k - 1, the tile's input equals the input
    // of its last computed launch; with as many iterations, this launch would
    // reproduce that launch's field (already in both buffers) and maxima.
    // Launch 0 and a replay compute every tile.
    if (k > 0 && C.replay == 0) {
        const bool qany = __ballot(qown && qraw != 0) != 0ull;
        if (!qany && titer_raw == n) {
            if (threadIdx.x == 0) {
                a.tchanged[(k & 1) * a.tiles + t] = 0;
                for (int i = 0; i < n; i++)
                    spread_max(sets + (size_t)i * kSpreadWords, a.texcess[(size_t)t * kRelaxKMax + i]);
            }
            return;
        }
    }
    bool changed = false;
    R tm[G::K];
#pragma unroll
    for (int i = 0; i < G::K; i++) tm[i] = (R)0;
    if (src != (k & 1)) vmask = relax_tile_load<R, G>(a.h[src], a.ny, a.nx, r0, c0, v);
    relax_tile_iterations_staged<R, G>(v, vmask, a.h[src ^ 1], a.ny, a.nx, a.crit, r0, c0, tile, n, tm, &changed);
    // one LDS round and one barrier: each wave's maxima and changed flag,
    // folded by the first wave (lane w reads wave w; the maximum is exact in
    // any order)
    constexpr int NW = G::BLOCK / kWave;
    static_assert(NW <= kWave, "the first wave folds one value per wave");
    const int wave = (int)(threadIdx.x / kWave);
    double m[G::K];
#pragma unroll
    for (int i = 0; i < G::K; i++) m[i] = wave_max((double)tm[i]);
    const bool wchanged = __ballot(changed) != 0ull;
    if (lane == 0) {
#pragma unroll
        for (int i = 0; i < G::K; i++) smem[i * NW + wave] = m[i];
        wflag[wave] = wchanged ? 1 : 0;
    }
    __syncthreads();
    if (wave == 0) {
#pragma unroll
        for (int i = 0; i < G::K; i++) m[i] = wave_max(lane < NW ? smem[i * NW + lane] : 0.0);
        const bool any = __ballot(lane < NW && wflag[lane] != 0) != 0ull;
        if (lane == 0) {
            for (int i = 0; i < n; i++) {
                spread_max(sets + (size_t)i * kSpreadWords, m[i]);
                a.texcess[(size_t)t * kRelaxKMax + i] = m[i];
            }
            a.titers[t] = n;
            a.tchanged[(k & 1) * a.tiles + t] = any ? 1 : 0;
        }
    }
}

// after launch k - 1: once the loop has ended, copy back when the field sits
// in scratch.  A loop still running (or with a replay pending) leaves both
// buffers alone: the next launches continue from them.  Writes the result to
// pinned host memory and clears the next call's first group (read by no
// launch of this call).
constexpr int kRelaxEndAhead = 2;    // copy-back cells per thread read before the loop state arrives

template <class R>
__global__ void __launch_bounds__(kBlock) relax_end_kernel(RelaxArgs<R> a, int k, unsigned *stamp_word,
                                                           unsigned stamp)
{
    // the copy-back's first cells are read with the maxima, before the loop
    // state says whether they are needed (the final field is in buffer 1)
    const int64_t n = (int64_t)a.ny * a.nx;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    const int64_t i0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    double slot[kRelaxKMax];
    relax_prev_slots(a, k, slot);
    R ahead[kRelaxEndAhead];
#pragma unroll
    for (int u = 0; u < kRelaxEndAhead; u++) ahead[u] = i0 + u * stride < n ? a.h[1][i0 + u * stride] : (R)0;
    const RelaxControl C = relax_resolve(a, k, slot);
    if (C.done && C.src == 1) {
#pragma unroll
        for (int u = 0; u < kRelaxEndAhead; u++)
            if (i0 + u * stride < n) a.h[0][i0 + u * stride] = ahead[u];
        for (int64_t i = i0 + kRelaxEndAhead * stride; i < n; i += stride) a.h[0][i] = a.h[1][i];
    }
    if (blockIdx.x == 0) {
        for (int i = 0; i < kRelaxKMax; i++) spread_clear(relax_set(a.spread, a.first_group == 3 ? 4 : 3, i));
        if (threadIdx.x == 0) {
            RelaxHostOut r{};
            r.r.iterations = C.iterations;
            r.r.converged = C.converged;
            r.r.max_excess_last = C.last;
            r.r.complete = C.done;
            r.marked = C.marked_set ? C.marked : C.last;
            *a.out = r;
            host_stamp(stamp_word, stamp);
        }
    }
}

__global__ void zero_relax_sync(RelaxSync *s, unsigned long long *spread)
{
    unsigned *p = reinterpret_cast<unsigned *>(s);
    for (int i = threadIdx.x; i < (int)(sizeof(RelaxSync) / 4); i += blockDim.x) p[i] = 0u;
    for (int i = threadIdx.x; i < (kRelaxGroups + 2) * kRelaxKMax * kSpread; i += blockDim.x)
        spread[(int64_t)i * kSpreadStride] = 0ull;
}

template <class R>
static int launch_relax(mv2_context *ctx, void *height, int ny, int nx, const double crit_diff[4],
                        int max_iterations, double tolerance, mv2_relax_result *out, int mark = 0,
                        double *marked_out = nullptr)
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
    static_assert(sizeof(RelaxSync) <= 1024, "relaxation loop state");
    if (int rc = ensure_scratch(ctx, ctx->spread, kSpreadScratchBytes)) return rc;
    a.spread = (unsigned long long *)ctx->spread.ptr + (size_t)kRelaxSpreadSet * kSpreadWords;
    a.mark = mark;
    a.out = (RelaxHostOut *)((char *)ctx->host_result.ptr + 2048);
    static_assert(2048 + sizeof(RelaxHostOut) <= 3072, "the relaxation result in the pinned block");
    a.first_group = 3 + (ctx->relax_calls & 1);
    a.tiles_x = (nx + kRelaxTileW - 1) / kRelaxTileW;
    a.tiles = a.tiles_x * ((ny + kRelaxTileH - 1) / kRelaxTileH);
    const size_t quiet = (sizeof(double) * kRelaxKMax + 3 * sizeof(int32_t)) * (size_t)a.tiles;
    if (int rc = ensure_scratch(ctx, ctx->relax_quiet, quiet)) return rc;
    a.texcess = (double *)ctx->relax_quiet.ptr;
    a.tchanged = (int32_t *)(a.texcess + (size_t)kRelaxKMax * a.tiles);
    a.titers = a.tchanged + 2 * (size_t)a.tiles;
    hipStream_t st = ctx->stream;
    const int gend = grid_for((int64_t)n, kBlock, 4 * ctx->cu_count);
    if (!ctx->relax_clean) hipLaunchKernelGGL(zero_relax_sync, dim3(1), dim3(kBlock), 0, st, a.sync, a.spread);
    ctx->relax_clean = 0;          // set again once this call has completed
    MV2_HIP(ctx, hipGetLastError());
    // Launches after the loop's end return at once.  Enqueue about as many as
    // the previous call needed (at most 32), then grow the batch while the
    // field has not converged.  Launch 0 applies one iteration; later FP32
    // launches two (FP64 one: its two-iteration tile would spill registers).
    // A field that converged in its first iteration last time gets one launch
    // (the end pass detects the convergence); otherwise one more than the
    // launches that worked, which covers a replay.
    int chunk = ctx->relax_predicted <= 1 ? 1 : ctx->relax_predicted + 1;
    if (chunk > 32) chunk = 32;
    int k = 0;
    const volatile RelaxHostOut *hout = a.out;
    const volatile mv2_relax_result *hres = &hout->r;
    // two iterations per launch only while every workgroup of a launch runs at
    // once (at most 2 per CU; 701^2: 240 tiles): a larger grid runs in rounds,
    // where the two-iteration tile's 1.4x larger LDS area costs more than the
    // halved launches save
    const bool pair = a.tiles <= 2 * (ctx->cu_count > 0 ? ctx->cu_count : 256);
    for (;;) {
        for (int j = 0; j < chunk; j++, k++) {
            if constexpr (sizeof(R) == 8) {
                hipLaunchKernelGGL((relax_kernel<R, RelaxSingle64>), dim3(a.tiles), dim3(RelaxSingle64::BLOCK), 0, st, a, k);
            } else {
                if (k == 0 || !pair)
                    hipLaunchKernelGGL((relax_kernel<R, RelaxFirst>), dim3(a.tiles), dim3(RelaxFirst::BLOCK), 0, st, a, k);
                else
                    hipLaunchKernelGGL((relax_kernel<R, RelaxPair>), dim3(a.tiles), dim3(RelaxPair::BLOCK), 0, st, a, k);
            }
        }
        const unsigned stamp = next_stamp(ctx, 1);
        hipLaunchKernelGGL(relax_end_kernel<R>, dim3(gend), dim3(kBlock), 0, st, a, k, stamp_words(ctx), stamp);
        MV2_HIP(ctx, hipGetLastError());
        if (int rc_w = stamp_wait(ctx, st, stamp, 1)) return rc_w;
        if (hres->complete) break;
        // each launch applies at least one iteration: max_iterations + 1
        // launches always end the loop (the last may be a replay)
        if (k > max_iterations + 1) return set_error(ctx, MV2_EDEVICE, "[MiniSlope] relaxation loop did not end");
        chunk = k / 2 > 8 ? k / 2 : 8;
    }
    *out = mv2_relax_result{};
    out->iterations = hres->iterations;
    out->converged = hres->converged;
    out->max_excess_last = hres->max_excess_last;
    out->complete = hres->complete;
    out->launches = k;
    if (marked_out) *marked_out = hout->marked;
    ctx->relax_clean = 1;
    ctx->relax_calls++;
    // launches that applied iterations: one, then pairs (FP32, small grids) or
    // ones; the next call enqueues one more, which covers a replay
    const int it = out->iterations;
    ctx->relax_predicted = (sizeof(R) == 8 || !pair) ? it : (it <= 1 ? 1 : 1 + it / 2);
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

extern "C" int mv2_relax_marked(mv2_context *ctx, int precision, void *height, int ny, int nx,
                                const double crit_diff[4], int max_iterations, double tolerance, int mark_iteration,
                                mv2_relax_result *out, double *max_excess_at_mark)
{
    if (!max_excess_at_mark) return set_error(ctx, MV2_EINVAL, "[MiniSlope] null argument");
    if (mark_iteration < 1) return set_error(ctx, MV2_EINVAL, "[MiniSlope] mark_iteration must be >= 1");
    if (!ctx) return set_error(nullptr, MV2_EINVAL, "[MiniSlope] null context");
    if (!height || !crit_diff || !out) return set_error(ctx, MV2_EINVAL, "[MiniSlope] null argument");
    if (ny < 2 || nx < 2)
        return set_error(ctx, MV2_EINVAL, "height must be a 2-D array with both dimensions >= 2");
    if (max_iterations < 1) return set_error(ctx, MV2_EINVAL, "[MiniSlope] max_iterations must be >= 1");
    if (precision != MV2_FP64 && precision != MV2_FP32)
        return set_error(ctx, MV2_EINVAL, "[MiniSlope] unknown precision %d", precision);
    hipSetDevice(ctx->device);
    return precision == MV2_FP64 ? launch_relax<double>(ctx, height, ny, nx, crit_diff, max_iterations, tolerance,
                                                        out, mark_iteration, max_excess_at_mark)
                                 : launch_relax<float>(ctx, height, ny, nx, crit_diff, max_iterations, tolerance,
                                                       out, mark_iteration, max_excess_at_mark);
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
