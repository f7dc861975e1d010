// Batched Resting -> Mobile exchange and batched failure labelling: the
// single-terrain contracts applied to every instance of the bound batch at
// once, each instance exactly as if it were handled alone.
//
// entrain: bulk_state_authority.py:104-127 (entrain) on (instance, cell) index
// lists.  Depths become whole quanta rounded down, so the donor limit holds
// and the volume moved per instance is exact (integer quanta summed with
// integer atomics: the order of arrival does not matter).
//
// label: the failure mask of gpu_large_avalanche.py:59-162 (cohesive start
// criterion on the gradient of H = b + h) and the connected components of
// :165-252 per instance: 32 x 32 tiles of every instance, union-find in LDS
// inside a tile, border unions in global memory, flatten with per-root
// counts.  Labels are instance-local linear indices, the root of a component
// is its minimum index -- the labels of the terrain labelled alone.  Region
// count, largest component, winner (smallest root among the largest) and the
// unstable cell count per instance come from integer atomics per instance.
#include "mv2_internal.hpp"
#include "mv2_batch.hpp"
#include "mv2_label_common.hpp"

#include <cmath>
#include <cstring>
#include <vector>

namespace mv2 {

// ---- entrainment ------------------------------------------------------------
// One thread per (instance, cell) request.  The quanta moved are summed per
// instance: the lanes of a wave that name the same instance (request lists are
// grouped by instance, so a wave usually holds one or two) add theirs with a
// wave reduction and one atomic, instead of one atomic per cell on the few
// counters of the batch.
__global__ void __launch_bounds__(kBlock) batch_entrain_kernel(const float *z_base, float *b, int32_t *h,
                                                               const int64_t *inst_off, const int32_t *inst,
                                                               const int32_t *cell, const double *req, int n,
                                                               unsigned long long *moved)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    int my_inst = -1;
    long long dq = 0;
    if (k < n) {
        my_inst = inst[k];
        const int64_t i = inst_off[my_inst] + cell[k];
        const double zb = z_base ? (double)z_base[i] : 0.0;
        const double available = fmax((double)b[i] - zb, 0.0);
        const double depth = fmin(fmax(req[k], 0.0), available);
        dq = entrain_quanta_f32(b[i], depth, 2147483647LL - (long long)h[i]);
        if (dq > 0) {
            b[i] = (float)((double)b[i] - (double)dq * kQuantum);   // exact
            h[i] = (int32_t)((long long)h[i] + dq);
        } else {
            dq = 0;
        }
    }
    // per distinct instance of the wave: sum its lanes' quanta, one atomic
    const int lane = threadIdx.x & (kWave - 1);
    bool pending = my_inst >= 0;
    for (;;) {
        const unsigned long long active = __ballot(pending);
        if (!active) break;
        const int leader = __ffsll((long long)active) - 1;
        const int target = __shfl(my_inst, leader);
        long long v = pending && my_inst == target ? dq : 0;
#pragma unroll
        for (int off = kWave / 2; off > 0; off >>= 1) v += __shfl_xor(v, off);
        if (lane == leader && v > 0) atomicAdd(&moved[target], (unsigned long long)v);
        if (my_inst == target) pending = false;
    }
}

__global__ void zero_u64(unsigned long long *p, int n)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = 0ull;
}

// ---- labelling ----------------------------------------------------------------
struct BatchLabelArgs {
    int count, tiles;
    const float *b;
    const int32_t *h;
    const float *z_base;                // nullptr: zero
    const int32_t *inst_ny, *inst_nx;
    const int64_t *inst_off;
    const int32_t *tile_inst, *tile_r0, *tile_c0;
    int32_t *labels;                    // packed, instance-local indices or -1
    int32_t *counts;                    // packed: cells per root
    int32_t *stats;                     // [count][4]: regions, largest, winner, unstable
    double dx, dy;
    mv2_avalanche_config cfg;
};

// the start criterion of gpu_large_avalanche.py:59-162 at one cell (FP32
// storage read into FP64, the same expressions as the single-terrain kernel)
__device__ inline bool failure_start(const BatchLabelArgs &a, const float *b, const int32_t *h, const float *zb,
                                     int ny, int nx, int row, int col)
{
    const int64_t index = (int64_t)row * nx + col;
    int64_t left = index, right = index, down = index, up = index;
    double denom_x = a.dx, denom_y = a.dy;
    if (col > 0) left = index - 1;
    if (col + 1 < nx) right = index + 1;
    if (col > 0 && col + 1 < nx) denom_x = 2.0 * a.dx;
    if (row > 0) down = index - nx;
    if (row + 1 < ny) up = index + nx;
    if (row > 0 && row + 1 < ny) denom_y = 2.0 * a.dy;
    auto rest = [&](int64_t i) { return (double)b[i]; };
    auto mob = [&](int64_t i) { return (double)h[i] * kQuantum; };
    const double gx = (rest(right) + mob(right) - rest(left) - mob(left)) / denom_x;
    const double gy = (rest(up) + mob(up) - rest(down) - mob(down)) / denom_y;
    const double magnitude = sqrt(gx * gx + gy * gy);
    const double z = zb ? (double)zb[index] : 0.0;
    const double layer_depth = fmin(fmax(rest(index) - z, 0.0), a.cfg.mobilization_depth);
    const double normalization = sqrt(1.0 + magnitude * magnitude);
    const double drive = a.cfg.density * a.cfg.gravity * layer_depth * magnitude / normalization;
    const double normal = a.cfg.density * a.cfg.gravity * layer_depth / normalization;
    const double start_margin = drive - (a.cfg.cohesion + normal * a.cfg.tan_start);
    return layer_depth > a.cfg.dry_tolerance && start_margin > 0.0;
}

__global__ void batch_label_reset_kernel(BatchLabelArgs a, int64_t cells)
{
    const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t nth = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = tid; i < cells; i += nth) a.counts[i] = 0;
    for (int64_t i = tid; i < (int64_t)a.count; i += nth) {
        a.stats[4 * i + 0] = 0;
        a.stats[4 * i + 1] = 0;
        a.stats[4 * i + 2] = 0x7fffffff;
        a.stats[4 * i + 3] = 0;
    }
}

// mask of the tile's cells, then the tile-local union-find (as ccl_local_kernel)
__global__ void __launch_bounds__(kBlock) batch_ccl_local_kernel(BatchLabelArgs a)
{
    __shared__ int lp[kCclCells];
    const int g = xcd_block();
    const int inst = a.tile_inst[g];
    const int ny = a.inst_ny[inst], nx = a.inst_nx[inst];
    const int64_t off = a.inst_off[inst];
    const float *b = a.b + off, *zb = a.z_base ? a.z_base + off : nullptr;
    const int32_t *h = a.h + off;
    int32_t *p = a.labels + off;
    const int tr0 = a.tile_r0[g], tc0 = a.tile_c0[g];
    const int connectivity = a.cfg.connectivity;
    bool masked[kCclPer];
#pragma unroll
    for (int j = 0; j < kCclPer; j++) {
        const int e = threadIdx.x + j * kBlock;
        const int gr = tr0 + e / kCcl, gc = tc0 + e % kCcl;
        masked[j] = gr < ny && gc < nx && failure_start(a, b, h, zb, ny, nx, gr, gc);
    }
#pragma unroll
    for (int j = 0; j < kCclPer; j++) lp[threadIdx.x + j * kBlock] = masked[j] ? threadIdx.x + j * kBlock : -1;
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kCclPer; j++) {
        if (!masked[j]) continue;
        const int e = threadIdx.x + j * kBlock;
        const int lr = e / kCcl, lc = e % kCcl;
        if (lc + 1 < kCcl && lds_load(lp, e + 1) >= 0) lds_unite(lp, e, e + 1);
        if (lr + 1 < kCcl) {
            if (lds_load(lp, e + kCcl) >= 0) lds_unite(lp, e, e + kCcl);
            if (connectivity == 8) {
                if (lc + 1 < kCcl && lds_load(lp, e + kCcl + 1) >= 0) lds_unite(lp, e, e + kCcl + 1);
                if (lc > 0 && lds_load(lp, e + kCcl - 1) >= 0) lds_unite(lp, e, e + kCcl - 1);
            }
        }
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kCclPer; j++) {
        const int e = threadIdx.x + j * kBlock;
        const int gr = tr0 + e / kCcl, gc = tc0 + e % kCcl;
        if (gr >= ny || gc >= nx) continue;
        int32_t v = -1;
        if (masked[j]) {
            const int r = lds_find(lp, e);
            v = (tr0 + r / kCcl) * nx + tc0 + r % kCcl;
        }
        p[gr * nx + gc] = v;
    }
}

// edges crossing the tile's right, bottom and left borders (as ccl_border_kernel)
__global__ void __launch_bounds__(kBlock) batch_ccl_border_kernel(BatchLabelArgs a)
{
    const int g = xcd_block();
    const int inst = a.tile_inst[g];
    const int ny = a.inst_ny[inst], nx = a.inst_nx[inst];
    int32_t *p = a.labels + a.inst_off[inst];
    const int tr0 = a.tile_r0[g], tc0 = a.tile_c0[g];
    const int connectivity = a.cfg.connectivity;
    for (int q = threadIdx.x; q < 3 * kCcl; q += blockDim.x) {
        int r, c;
        if (q < kCcl) { r = tr0 + q; c = tc0 + kCcl - 1; }
        else if (q < 2 * kCcl) { r = tr0 + kCcl - 1; c = tc0 + q - kCcl; }
        else { r = tr0 + q - 2 * kCcl; c = tc0; }
        if (r >= ny || c >= nx) continue;
        const int i = r * nx + c;
        if (load_parent(p, i) < 0) continue;
        const bool right = q < kCcl, bottom = q >= kCcl && q < 2 * kCcl, left = q >= 2 * kCcl;
        if (right && c + 1 < nx && load_parent(p, i + 1) >= 0) unite(p, i, i + 1);
        if (r + 1 < ny) {
            if (bottom && load_parent(p, i + nx) >= 0) unite(p, i, i + nx);
            if (connectivity == 8) {
                if ((right || bottom) && c + 1 < nx && load_parent(p, i + nx + 1) >= 0) unite(p, i, i + nx + 1);
                if ((left || bottom) && c > 0 && load_parent(p, i + nx - 1) >= 0) unite(p, i, i + nx - 1);
            }
        }
    }
}

// final labels and cells per root (LDS hash per tile, as ccl_flatten_kernel)
__global__ void __launch_bounds__(kBlock) batch_ccl_flatten_kernel(BatchLabelArgs a)
{
    __shared__ int keys[kCclHash];
    __shared__ int vals[kCclHash];
    const int g = xcd_block();
    const int inst = a.tile_inst[g];
    const int ny = a.inst_ny[inst], nx = a.inst_nx[inst];
    const int64_t off = a.inst_off[inst];
    int32_t *p = a.labels + off;
    int32_t *counts = a.counts + off;
    const int tr0 = a.tile_r0[g], tc0 = a.tile_c0[g];
    if (threadIdx.x < kCclHash) {
        keys[threadIdx.x] = -1;
        vals[threadIdx.x] = 0;
    }
    int root[kCclPer];
#pragma unroll
    for (int j = 0; j < kCclPer; j++) {
        const int e = threadIdx.x + j * kBlock;
        const int gr = tr0 + e / kCcl, gc = tc0 + e % kCcl;
        root[j] = -1;
        if (gr >= ny || gc >= nx) continue;
        int r = p[gr * nx + gc];
        if (r < 0) continue;
        while (true) {
            const int up = p[r];
            if (up == r) break;
            r = up;
        }
        root[j] = r;
    }
#pragma unroll
    for (int j = 0; j < kCclPer; j++) {
        if (root[j] < 0) continue;
        const int e = threadIdx.x + j * kBlock;
        p[(tr0 + e / kCcl) * nx + tc0 + e % kCcl] = root[j];
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kCclPer; j++) {
        const int r = root[j];
        if (r < 0) continue;
        int hsh = (int)(((unsigned)r * 2654435761u) >> 25) & (kCclHash - 1);
        bool done = false;
        for (int probe = 0; probe < kCclHash && !done; probe++) {
            const int prev = atomicCAS(&keys[hsh], -1, r);
            if (prev == -1 || prev == r) {
                atomicAdd(&vals[hsh], 1);
                done = true;
            }
            hsh = (hsh + 1) & (kCclHash - 1);
        }
        if (!done) atomicAdd(&counts[r], 1);
    }
    __syncthreads();
    if (threadIdx.x < kCclHash && keys[threadIdx.x] >= 0) atomicAdd(&counts[keys[threadIdx.x]], vals[threadIdx.x]);
}

// region count, unstable cells and largest component per instance
__global__ void __launch_bounds__(kBlock) batch_label_count_kernel(BatchLabelArgs a)
{
    __shared__ int smem[3 * (kBlock / kWave)];
    const int g = xcd_block();
    const int inst = a.tile_inst[g];
    const int ny = a.inst_ny[inst], nx = a.inst_nx[inst];
    const int32_t *counts = a.counts + a.inst_off[inst];
    const int tr0 = a.tile_r0[g], tc0 = a.tile_c0[g];
    int regions = 0, unstable = 0, largest = 0;
#pragma unroll
    for (int j = 0; j < kCclPer; j++) {
        const int e = threadIdx.x + j * kBlock;
        const int gr = tr0 + e / kCcl, gc = tc0 + e % kCcl;
        if (gr >= ny || gc >= nx) continue;
        const int c = counts[gr * nx + gc];
        if (c > 0) {
            regions += 1;
            unstable += c;
            largest = largest >= c ? largest : c;
        }
    }
    regions = wave_sum(regions);
    unstable = wave_sum(unstable);
    largest = wave_max(largest);
    const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
    if (lane == 0) {
        smem[wave] = regions;
        smem[kBlock / kWave + wave] = unstable;
        smem[2 * (kBlock / kWave) + wave] = largest;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        int rg = 0, un = 0, lg = 0;
        for (int w = 0; w < kBlock / kWave; w++) {
            rg += smem[w];
            un += smem[kBlock / kWave + w];
            lg = lg >= smem[2 * (kBlock / kWave) + w] ? lg : smem[2 * (kBlock / kWave) + w];
        }
        if (rg) atomicAdd(&a.stats[4 * inst + 0], rg);
        if (lg) atomicMax(&a.stats[4 * inst + 1], lg);
        if (un) atomicAdd(&a.stats[4 * inst + 3], un);
    }
}

// winner: the smallest root among the instance's largest components
__global__ void __launch_bounds__(kBlock) batch_label_winner_kernel(BatchLabelArgs a)
{
    const int g = xcd_block();
    const int inst = a.tile_inst[g];
    const int largest = a.stats[4 * inst + 1];
    if (largest <= 0) return;
    const int ny = a.inst_ny[inst], nx = a.inst_nx[inst];
    const int32_t *counts = a.counts + a.inst_off[inst];
    const int tr0 = a.tile_r0[g], tc0 = a.tile_c0[g];
    int best = 0x7fffffff;
#pragma unroll
    for (int j = 0; j < kCclPer; j++) {
        const int e = threadIdx.x + j * kBlock;
        const int gr = tr0 + e / kCcl, gc = tc0 + e % kCcl;
        if (gr >= ny || gc >= nx) continue;
        const int i = gr * nx + gc;
        if (counts[i] == largest) best = i < best ? i : best;
    }
    best = wave_reduce<2>(best);
    if ((threadIdx.x & (kWave - 1)) == 0 && best != 0x7fffffff) atomicMin(&a.stats[4 * inst + 2], best);
}

}  // namespace mv2

using namespace mv2;

// 32 x 32 labelling tiles of every instance (built on first use after a bind)
static int ensure_label_tiles(mv2_context *ctx, mv2_batch_state *bs)
{
    if (bs->ctiles > 0) return MV2_OK;
    std::vector<int32_t> ti, tr, tc;
    for (int i = 0; i < bs->count; i++) {
        const int tx = (bs->nx[i] + kCcl - 1) / kCcl, ty = (bs->ny[i] + kCcl - 1) / kCcl;
        for (int r = 0; r < ty; r++)
            for (int c = 0; c < tx; c++) {
                ti.push_back(i);
                tr.push_back(r * kCcl);
                tc.push_back(c * kCcl);
            }
    }
    const size_t T = ti.size();
    if (int rc = ensure_scratch(ctx, bs->ctab, sizeof(int32_t) * 3 * T)) return rc;
    int32_t *tab = (int32_t *)bs->ctab.ptr;
    MV2_HIP(ctx, hipMemcpy(tab, ti.data(), sizeof(int32_t) * T, hipMemcpyHostToDevice));
    MV2_HIP(ctx, hipMemcpy(tab + T, tr.data(), sizeof(int32_t) * T, hipMemcpyHostToDevice));
    MV2_HIP(ctx, hipMemcpy(tab + 2 * T, tc.data(), sizeof(int32_t) * T, hipMemcpyHostToDevice));
    bs->ctiles = (int)T;
    return MV2_OK;
}

extern "C" int mv2_batch_entrain(mv2_context *ctx, const float *z_base, const int32_t *inst, const int32_t *cell,
                                 const double *depth, int n, double *volume_out)
{
    if (!ctx) return set_error(nullptr, MV2_EINVAL, "[BatchedDeviceBulkState] null context");
    mv2_batch_state *bs = ctx->batch;
    if (!bs || !bs->count) return set_error(ctx, MV2_ESTATE, "[BatchedDeviceBulkState] bind a batch first");
    if (n < 0 || (n > 0 && (!inst || !cell || !depth)) || !volume_out)
        return set_error(ctx, MV2_EINVAL, "[BatchedDeviceBulkState] invalid entrainment request");
    hipSetDevice(ctx->device);
    const size_t cnt = (size_t)bs->count;
    if (int rc = ensure_scratch(ctx, bs->moved, sizeof(unsigned long long) * cnt)) return rc;
    unsigned long long *moved = (unsigned long long *)bs->moved.ptr;
    hipStream_t st = ctx->stream;
    hipLaunchKernelGGL(zero_u64, dim3(grid_for((int64_t)cnt, kBlock, 64)), dim3(kBlock), 0, st, moved, bs->count);
    if (n > 0)
        hipLaunchKernelGGL(batch_entrain_kernel, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, st, z_base, bs->b,
                           bs->h, batch_inst_tables(bs).off, inst, cell, depth, n, moved);
    MV2_HIP(ctx, hipGetLastError());
    if (int rc = ensure_scratch(ctx, ctx->host_result, sizeof(unsigned long long) * cnt + 4096, true)) return rc;
    MV2_HIP(ctx, hipMemcpyAsync(ctx->host_result.ptr, moved, sizeof(unsigned long long) * cnt, hipMemcpyDeviceToHost,
                                st));
    if (int rc_w = stream_wait(ctx, st)) return rc_w;
    const unsigned long long *hq = (const unsigned long long *)ctx->host_result.ptr;
    const double area = bs->dx * bs->dy;
    for (size_t i = 0; i < cnt; i++) volume_out[i] = ((double)hq[i] * kQuantum) * area;
    return MV2_OK;
}

extern "C" int mv2_batch_label(mv2_context *ctx, const float *z_base, const mv2_avalanche_config *cfg,
                               int32_t *labels, int32_t *stats_out)
{
    if (!ctx) return set_error(nullptr, MV2_EINVAL, "[BatchedLargeAvalanche] null context");
    mv2_batch_state *bs = ctx->batch;
    if (!bs || !bs->count) return set_error(ctx, MV2_ESTATE, "[BatchedLargeAvalanche] bind a batch first");
    if (!cfg || !labels || !stats_out) return set_error(ctx, MV2_EINVAL, "[BatchedLargeAvalanche] null argument");
    if (cfg->connectivity != 4 && cfg->connectivity != 8)
        return set_error(ctx, MV2_EINVAL, "[LargeAvalanche] connectivity must be 4 or 8");
    hipSetDevice(ctx->device);
    if (int rc = ensure_label_tiles(ctx, bs)) return rc;
    const size_t cnt = (size_t)bs->count;
    if (int rc = ensure_scratch(ctx, bs->label_counts, sizeof(int32_t) * (size_t)bs->cells)) return rc;
    if (int rc = ensure_scratch(ctx, bs->label_stats, sizeof(int32_t) * 4 * cnt)) return rc;
    BatchLabelArgs a;
    a.count = bs->count;
    a.tiles = bs->ctiles;
    a.b = bs->b;
    a.h = bs->h;
    a.z_base = z_base;
    const BatchInstTables it = batch_inst_tables(bs);
    a.inst_ny = it.ny;
    a.inst_nx = it.nx;
    a.inst_off = it.off;
    const int32_t *tab = (const int32_t *)bs->ctab.ptr;
    const size_t T = (size_t)bs->ctiles;
    a.tile_inst = tab;
    a.tile_r0 = tab + T;
    a.tile_c0 = tab + 2 * T;
    a.labels = labels;
    a.counts = (int32_t *)bs->label_counts.ptr;
    a.stats = (int32_t *)bs->label_stats.ptr;
    a.dx = bs->dx;
    a.dy = bs->dy;
    a.cfg = *cfg;
    hipStream_t st = ctx->stream;
    hipLaunchKernelGGL(batch_label_reset_kernel, dim3(grid_for(bs->cells, kBlock, 2048)), dim3(kBlock), 0, st, a,
                       (int64_t)bs->cells);
    hipLaunchKernelGGL(batch_ccl_local_kernel, dim3(bs->ctiles), dim3(kBlock), 0, st, a);
    hipLaunchKernelGGL(batch_ccl_border_kernel, dim3(bs->ctiles), dim3(kWave * 2), 0, st, a);
    hipLaunchKernelGGL(batch_ccl_flatten_kernel, dim3(bs->ctiles), dim3(kBlock), 0, st, a);
    hipLaunchKernelGGL(batch_label_count_kernel, dim3(bs->ctiles), dim3(kBlock), 0, st, a);
    hipLaunchKernelGGL(batch_label_winner_kernel, dim3(bs->ctiles), dim3(kBlock), 0, st, a);
    MV2_HIP(ctx, hipGetLastError());
    if (int rc = ensure_scratch(ctx, ctx->host_result, sizeof(int32_t) * 4 * cnt + 4096, true)) return rc;
    MV2_HIP(ctx, hipMemcpyAsync(ctx->host_result.ptr, a.stats, sizeof(int32_t) * 4 * cnt, hipMemcpyDeviceToHost, st));
    if (int rc_w = stream_wait(ctx, st)) return rc_w;
    const int32_t *hs = (const int32_t *)ctx->host_result.ptr;
    for (size_t i = 0; i < cnt; i++) {
        stats_out[4 * i + 0] = hs[4 * i + 0];
        stats_out[4 * i + 1] = hs[4 * i + 1];
        stats_out[4 * i + 2] = hs[4 * i + 2] == 0x7fffffff ? -1 : hs[4 * i + 2];
        stats_out[4 * i + 3] = hs[4 * i + 3];
    }
    return MV2_OK;
}
