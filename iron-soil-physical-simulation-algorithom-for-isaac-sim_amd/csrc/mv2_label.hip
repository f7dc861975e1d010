// LargeAvalanche failure mask, connected-component labelling, component
// statistics and Resting->Mobile activation (pkg/runtime/gpu_large_avalanche.py).
//
// Labelling is a single-pass lock-free union-find instead of the reference's
// host-driven "union_pass + compress until unchanged" loop (:576-599, up to
// max(ny, nx) launches each followed by a host sync): every masked cell
// unites itself with its E, S (and SE, SW for 8-connectivity) neighbours by
// hooking the larger root onto the smaller with compare-and-swap, retrying
// when another thread hooked first.  Parent indices only ever decrease, so a
// component's final root is its minimum linear index -- the same canonical
// label as the reference (root = min index, tie -> smallest root).  One
// flatten pass then writes every cell's root and counts cells per root.
// The same pass yields the component statistics: region count (cells that
// are their own root), unstable cells, and the largest component -- a
// tile's atomicAdd onto a root's count returns the count before it, so the
// tile that adds last for a root sees that root's final count, and the
// maximum over every (count seen, root) key is the final maximum.
//
// Floating-point sums over the winning component use per-workgroup partials
// combined in fixed order (the reference uses FP64 atomics).
#include "mv2_internal.hpp"
#include "mv2_label_common.hpp"

#include <cmath>
#include <cstring>

namespace mv2 {

constexpr int kLabelParts = 2048;  // workgroups of the statistics kernels (one cell per thread at 701^2)
constexpr int kWinF = 7;           // winner float statistics

struct LabelAux {
    int32_t diag_i[8];
    double diag_f[8];
};

struct InitArgs {
    const void *z_base, *resting, *mobile, *qx, *qy;
    int f32, ny, nx;
    double dx, dy;
    mv2_avalanche_config cfg;
    double area, f6;
    mv2_avalanche_fields f;
};

// a field value of the state's precision as double (FP32: float fields,
// int32 mobile quanta), known at compile time so that a cell's loads form one
// straight-line group
template <bool F32>
__device__ inline double ldf(const void *p, int64_t i)
{
    return F32 ? (double)((const float *)p)[i] : ((const double *)p)[i];
}
template <bool F32>
__device__ inline double ldm(const void *p, int64_t i)
{
    return F32 ? (double)((const int32_t *)p)[i] * kQuantum : ((const double *)p)[i];
}

// initialize (gpu_large_avalanche.py:59-162) of one cell: gradient, slope,
// cohesive start / stop margins, severity, the unstable flag and the latch
// release; zeroes the cell's per-root count.  Returns the unstable flag.
// Every value the cell needs is read first (one memory round trip for the
// group), then the arithmetic runs in the reference's order.
template <bool F32>
__device__ inline bool init_cell(const InitArgs &A, int64_t index)
{
    const int ny = A.ny, nx = A.nx;
    const double dx = A.dx, dy = A.dy, area = A.area, f6 = A.f6;
    const mv2_avalanche_config &cfg = A.cfg;
    const mv2_avalanche_fields &f = A.f;
    const int row = (int)index / nx, col = (int)index - row * nx;
    int64_t left = index, right = index, down = index, up = index;
    double denom_x = dx, denom_y = dy;
    if (col > 0) left = index - 1;
    if (col + 1 < nx) right = index + 1;
    if (col > 0 && col + 1 < nx) denom_x = 2.0 * dx;
    if (row > 0) down = index - nx;
    if (row + 1 < ny) up = index + nx;
    if (row > 0 && row + 1 < ny) denom_y = 2.0 * dy;
    const double rr = ldf<F32>(A.resting, right), rl = ldf<F32>(A.resting, left);
    const double ru = ldf<F32>(A.resting, up), rd = ldf<F32>(A.resting, down);
    const double mr = ldm<F32>(A.mobile, right), ml = ldm<F32>(A.mobile, left);
    const double mu = ldm<F32>(A.mobile, up), md = ldm<F32>(A.mobile, down);
    const double rest = ldf<F32>(A.resting, index), zb = ldf<F32>(A.z_base, index);
    const double m = ldm<F32>(A.mobile, index);
    const double qxv = ldf<F32>(A.qx, index), qyv = ldf<F32>(A.qy, index);
    const double exported = f.export_cumulative[index] - f.owned_export_baseline[index];
    f.counts[index] = 0;
    const double gx = (rr + mr - rl - ml) / denom_x;
    const double gy = (ru + mu - rd - md) / denom_y;
    const double magnitude = sqrt(gx * gx + gy * gy);
    f.grad_x[index] = gx;
    f.grad_y[index] = gy;
    f.slope[index] = atan(magnitude);
    const double layer_depth = fmin(fmax(rest - zb, 0.0), cfg.mobilization_depth);
    const double normalization = sqrt(1.0 + magnitude * magnitude);
    const double drive = cfg.density * cfg.gravity * layer_depth * magnitude / normalization;
    const double normal = cfg.density * cfg.gravity * layer_depth / normalization;
    const double start_margin = drive - (cfg.cohesion + normal * cfg.tan_start);
    const double stop_margin = drive - (cfg.cohesion + normal * cfg.tan_stop);
    const double denom = normal * (cfg.tan_start - cfg.tan_stop);
    f.severity[index] = clampd(stop_margin / (denom >= 1.0e-12 ? denom : 1.0e-12), 0.0, 1.0);
    const bool active = layer_depth > cfg.dry_tolerance && start_margin > 0.0;
    f.unstable[index] = active ? 1 : 0;
    double speed = 0.0;
    if (m > cfg.dry_tolerance) {
        const double vx = qxv / m, vy = qyv / m;
        speed = sqrt(vx * vx + vy * vy);
    }
    const double w = cell_weight(cfg.weight_mode, ny, nx, area, f6, row, col);
    const bool resolved = exported > cfg.dry_tolerance * w;
    if (cfg.release_settled_latches && speed <= cfg.activity_speed && (resolved || stop_margin <= 0.0))
        f.latch[index] = 0;
    return active;
}

// The failure mask fused into the tile-local union-find: one workgroup per
// 32 x 32 tile evaluates initialize for its cells (the same per-cell
// arithmetic as the reference kernel), unites the unstable ones in LDS and
// writes every unstable cell's tile-local root; workgroup 0 also zeroes the
// statistics record and the label spread sets for the launches that follow.
constexpr int kLocalBlock = kCclCells;   // one thread per cell of the 32 x 32 tile

template <bool F32>
__global__ void __launch_bounds__(kLocalBlock) avalanche_local_kernel(InitArgs A, int tiles_x, LabelAux *aux,
                                                                      int32_t n_plus_one, unsigned long long *spread)
{
    __shared__ int lp[kCclCells];
    const int ny = A.ny, nx = A.nx;
    if (blockIdx.x == 0) {
        for (int i = threadIdx.x; i < 3 * kSpread; i += blockDim.x) spread[(int64_t)i * kSpreadStride] = 0ull;
        if (threadIdx.x == 0) {
            for (int k = 0; k < 8; k++) { aux->diag_i[k] = 0; aux->diag_f[k] = 0.0; }
            aux->diag_i[2] = n_plus_one;
        }
    }
    const int t = blockIdx.x;
    const int tr0 = (t / tiles_x) * kCcl, tc0 = (t % tiles_x) * kCcl;
    const int e = threadIdx.x;
    const int lr = e / kCcl, lc = e % kCcl;
    const int gr = tr0 + lr, gc = tc0 + lc;
    const bool inside = gr < ny && gc < nx;
    const bool masked = inside && init_cell<F32>(A, (int64_t)gr * nx + gc);
    lp[e] = masked ? e : -1;
    __syncthreads();
    if (masked) {
        if (lc + 1 < kCcl && lds_load(lp, e + 1) >= 0) lds_unite(lp, e, e + 1);
        if (lr + 1 < kCcl) {
            if (lds_load(lp, e + kCcl) >= 0) lds_unite(lp, e, e + kCcl);
            if (A.cfg.connectivity == 8) {
                if (lc + 1 < kCcl && lds_load(lp, e + kCcl + 1) >= 0) lds_unite(lp, e, e + kCcl + 1);
                if (lc > 0 && lds_load(lp, e + kCcl - 1) >= 0) lds_unite(lp, e, e + kCcl - 1);
            }
        }
    }
    __syncthreads();
    if (!inside) return;
    int32_t out = -1;
    if (masked) {
        const int r = lds_find(lp, e);
        out = (tr0 + r / kCcl) * nx + tc0 + r % kCcl;
    }
    A.f.labels[(int64_t)gr * nx + gc] = out;
}

__global__ void init_from_mask(const int32_t *mask, int32_t *labels, int64_t n)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        labels[i] = mask[i] ? (int32_t)i : -1;
}

__global__ void __launch_bounds__(kBlock) ccl_local_kernel(int32_t *p, int ny, int nx, int connectivity, int tiles_x)
{
    __shared__ int lp[kCclCells];
    const int t = blockIdx.x;
    const int tr0 = (t / tiles_x) * kCcl, tc0 = (t % tiles_x) * kCcl;
    bool masked[kCclPer];
#pragma unroll
    for (int j = 0; j < kCclPer; j++) {
        const int e = threadIdx.x + j * kBlock;
        const int gr = tr0 + e / kCcl, gc = tc0 + e % kCcl;
        masked[j] = gr < ny && gc < nx && p[gr * nx + gc] >= 0;
    }
#pragma unroll
    for (int j = 0; j < kCclPer; j++) lp[threadIdx.x + j * kBlock] = masked[j] ? threadIdx.x + j * kBlock : -1;
    __syncthreads();
    // masked cells keep lp >= 0 and unmasked ones -1 throughout, so the
    // neighbour test may read lp while other threads hook roots
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
        if (!masked[j]) continue;
        const int e = threadIdx.x + j * kBlock;
        const int r = lds_find(lp, e);
        p[(tr0 + e / kCcl) * nx + tc0 + e % kCcl] = (tr0 + r / kCcl) * nx + tc0 + r % kCcl;
    }
}

// Union of the components holding a and b, starting from ancestors ra of a
// and rb of b (their parents as read by the caller): both chains are followed
// to their roots in lockstep, one pair of loads per step, then the larger root
// is hooked onto the smaller as in unite.  Parents only ever decrease, so a
// stale ancestor still leads to the current root.  (Pointing the two starting
// ancestors at the roots found shortened the flatten pass by what it cost
// here: 701^2, 10.8 + 6.2 against 9.8 + 7.0 us.)
__device__ inline void unite_from(int32_t *p, int32_t ra, int32_t rb)
{
    for (;;) {
        const int32_t na = load_parent(p, ra), nb = load_parent(p, rb);
        if (na == ra && nb == rb) break;
        ra = na;
        rb = nb;
    }
    while (ra != rb) {
        if (ra < rb) {
            int32_t expected = rb;
            if (__hip_atomic_compare_exchange_strong(&p[rb], &expected, ra, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                     __HIP_MEMORY_SCOPE_AGENT))
                return;
            rb = find_root(p, expected);
        } else {
            int32_t expected = ra;
            if (__hip_atomic_compare_exchange_strong(&p[ra], &expected, rb, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                     __HIP_MEMORY_SCOPE_AGENT))
                return;
            ra = find_root(p, expected);
        }
    }
}

// edges crossing tile boundaries: per tile 32 right-column cells (E, SE),
// 32 bottom-row cells (S, SE, SW) and 32 left-column cells (SW).  A cell reads
// its own parent and those of its neighbours across the border in one group
// (a neighbour outside the grid or not on this cell's edges reads the cell
// itself and is not used).
__global__ void __launch_bounds__(kBlock) ccl_border_kernel(int32_t *p, int ny, int nx, int connectivity,
                                                             int tiles_x, int tiles)
{
    const int total = tiles * 3 * kCcl;
    for (int idx = blockIdx.x * blockDim.x + threadIdx.x; idx < total; idx += gridDim.x * blockDim.x) {
        const int t = idx / (3 * kCcl), q = idx % (3 * kCcl);
        const int tr0 = (t / tiles_x) * kCcl, tc0 = (t % tiles_x) * kCcl;
        int r, c;
        if (q < kCcl) { r = tr0 + q; c = tc0 + kCcl - 1; }
        else if (q < 2 * kCcl) { r = tr0 + kCcl - 1; c = tc0 + q - kCcl; }
        else { r = tr0 + q - 2 * kCcl; c = tc0; }
        if (r >= ny || c >= nx) continue;
        const int i = r * nx + c;
        const bool right = q < kCcl, bottom = q >= kCcl && q < 2 * kCcl, left = q >= 2 * kCcl;
        const bool has_d = r + 1 < ny, eight = connectivity == 8;
        const bool use_e = right && c + 1 < nx;
        const bool use_s = bottom && has_d;
        const bool use_se = eight && (right || bottom) && has_d && c + 1 < nx;
        const bool use_sw = eight && (left || bottom) && has_d && c > 0;
        const int32_t pi = load_parent(p, i);
        const int32_t pe = load_parent(p, use_e ? i + 1 : i);
        const int32_t ps = load_parent(p, use_s ? i + nx : i);
        const int32_t pse = load_parent(p, use_se ? i + nx + 1 : i);
        const int32_t psw = load_parent(p, use_sw ? i + nx - 1 : i);
        if (pi < 0) continue;
        if (use_e && pe >= 0) unite_from(p, pi, pe);
        if (use_s && ps >= 0) unite_from(p, pi, ps);
        if (use_se && pse >= 0) unite_from(p, pi, pse);
        if (use_sw && psw >= 0) unite_from(p, pi, psw);
    }
}

// final labels (full path to the root) and cell counts per root; with
// `spread`, the component statistics of reduce_components + select_winner
// (:234-252) into spread sets 0, 1 and 2: region count, unstable cells and the
// key (count << 32) | (INT32_MAX - root), whose maximum is the largest
// component with the smallest root (the reference's tie rule).  Integer sums
// and maxima: the order of the atomics does not change them.
__device__ inline long long winner_key(int32_t count, int32_t root)
{
    return ((long long)count << 32) | (long long)(0x7fffffff - root);
}

__global__ void __launch_bounds__(kBlock) ccl_flatten_kernel(int32_t *p, int ny, int nx, int tiles_x, int32_t *counts,
                                                             unsigned long long *spread)
{
    __shared__ int keys[kCclHash];
    __shared__ int vals[kCclHash];
    const int t = blockIdx.x;
    const int tr0 = (t / tiles_x) * kCcl, tc0 = (t % tiles_x) * kCcl;
    if (counts && threadIdx.x < kCclHash) {
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
        const int i = gr * nx + gc;
        int r = p[i];
        if (r < 0) continue;
        while (true) {
            const int up = p[r];
            if (up == r) break;
            r = up;
        }
        root[j] = r;
    }
    // every read above observes the completed border pass; writing this
    // cell's own entry (to its root, an ancestor) cannot change another root
#pragma unroll
    for (int j = 0; j < kCclPer; j++) {
        if (root[j] < 0) continue;
        const int e = threadIdx.x + j * kBlock;
        p[(tr0 + e / kCcl) * nx + tc0 + e % kCcl] = root[j];
    }
    if (!counts) return;
    __syncthreads();
    long long regions = 0, unstable = 0, key = 0;
#pragma unroll
    for (int j = 0; j < kCclPer; j++) {
        const int r = root[j];
        if (r < 0) continue;
        const int e = threadIdx.x + j * kBlock;
        unstable += 1;
        regions += r == (tr0 + e / kCcl) * nx + tc0 + e % kCcl ? 1 : 0;
        int h = (int)(((unsigned)r * 2654435761u) >> 25) & (kCclHash - 1);
        bool done = false;
        for (int probe = 0; probe < kCclHash && !done; probe++) {
            const int prev = atomicCAS(&keys[h], -1, r);
            if (prev == -1 || prev == r) {
                atomicAdd(&vals[h], 1);
                done = true;
            }
            h = (h + 1) & (kCclHash - 1);
        }
        if (!done) {
            const int32_t before = atomicAdd(&counts[r], 1);
            key = key >= winner_key(before + 1, r) ? key : winner_key(before + 1, r);
        }
    }
    __syncthreads();
    if (threadIdx.x < kCclHash && keys[threadIdx.x] >= 0) {
        const int32_t r = keys[threadIdx.x], add = vals[threadIdx.x];
        const int32_t before = atomicAdd(&counts[r], add);
        key = key >= winner_key(before + add, r) ? key : winner_key(before + add, r);
    }
    if (!spread) return;
    regions = wave_sum(regions);
    unstable = wave_sum(unstable);
    key = wave_max(key);
    if ((threadIdx.x & (kWave - 1)) == 0) {
        const int slot = (blockIdx.x % kSpread) * kSpreadStride;
        if (regions) atomicAdd(&spread[slot], (unsigned long long)regions);
        if (unstable) atomicAdd(&spread[kSpreadWords + slot], (unsigned long long)unstable);
        if (key) atomicMax(&spread[2 * kSpreadWords + slot], (unsigned long long)key);
    }
}

__global__ void fill32(int32_t *p, int64_t n, int32_t v)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) p[i] = v;
}

// the flatten pass's statistics: regions, unstable cells, largest count and
// the winner root (n + 1 when there is no component); workgroup 0 records them
__device__ inline int32_t resolve_winner(const unsigned long long *spread, int64_t n, LabelAux *aux)
{
    __shared__ long long win;
    if (threadIdx.x < kWave) {
        const int slot = threadIdx.x * kSpreadStride;
        const long long regions = wave_sum((long long)spread[slot]);
        const long long unstable = wave_sum((long long)spread[kSpreadWords + slot]);
        const long long key = wave_max((long long)spread[2 * kSpreadWords + slot]);
        if (threadIdx.x == 0) {
            const int32_t largest = (int32_t)(key >> 32);
            const int32_t root = largest > 0 ? 0x7fffffff - (int32_t)(key & 0xffffffffll) : (int32_t)(n + 1);
            win = root;
            if (blockIdx.x == 0) {
                aux->diag_i[0] = (int32_t)regions;
                aux->diag_i[1] = largest;
                aux->diag_i[2] = root;
                aux->diag_i[3] = (int32_t)unstable;
            }
        }
    }
    __syncthreads();
    return (int32_t)win;
}

// aggregate + winner_diagnostics (:201-231, :255-281) restricted to the
// winner.  The first cell's label is read with the statistics the winner
// comes from, and a winner cell's fields are read together.
template <bool F32>
__global__ void __launch_bounds__(kBlock) winner_stats_kernel(
    const int32_t *labels, const void *z_base, const void *resting, const int32_t *latch,
    const int32_t *previous, const double *slope, const double *severity, int ny, int nx,
    mv2_avalanche_config cfg, double area, double f6, LabelAux *aux, const unsigned long long *spread,
    double *partials)
{
    __shared__ double smem[kWinF * 4];
    const int64_t n = (int64_t)ny * nx;
    const int64_t i0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    const int32_t lab0 = i0 < n ? labels[i0] : -1;
    const int32_t winner = resolve_winner(spread, n, aux);
    double s[5] = {0.0, 0.0, 0.0, 0.0, 0.0};   // sum es, sum et, area, volume, root excess
    double m[2] = {0.0, 0.0};                   // max es, max et
    int prev = 0;
    for (int64_t i = i0; i < n; i += stride) {
        const int32_t lab = i == i0 ? lab0 : labels[i];
        if (lab < 0 || lab != winner) continue;
        const int pv = previous ? previous[i] : 0;
        const double sl = slope[i], sev = severity[i];
        const int lt = latch[i];
        const double rs = ldf<F32>(resting, i), zb = ldf<F32>(z_base, i);
        const int r = (int)i / nx, c = (int)i - r * nx;
        const double w = cell_weight(cfg.weight_mode, ny, nx, area, f6, r, c);
        prev += pv;
        double es = fmax(sl - cfg.start_rad, 0.0);
        double et = fmax(sl - cfg.stop_rad, 0.0);
        s[0] += es;
        s[1] += et;
        m[0] = fmax(m[0], es);
        m[1] = fmax(m[1], et);
        double transfer = 0.0;
        if (lt == 0) transfer = fmin(fmax(rs - zb, 0.0), cfg.mobilization_depth * sev);
        s[2] += w;
        s[3] += transfer * w;
        s[4] += es;
    }
    block_reduce<double, 5, 0>(s, smem);
    block_reduce<double, 2, 1>(m, smem);
    const int ptot = wave_sum(prev);
    if ((threadIdx.x & (kWave - 1)) == 0 && ptot) atomicAdd(&aux->diag_i[4], ptot);
    if (threadIdx.x == 0) {
        // field-major partials: field f of workgroup g at [f * gridDim.x + g]
        const double vals[kWinF] = {s[0], m[0], s[1], m[1], s[2], s[3], s[4]};
        for (int f = 0; f < kWinF; f++) partials[(size_t)f * gridDim.x + blockIdx.x] = vals[f];
    }
}

constexpr int kFinalBlock = 1024;   // the fold of up to kLabelParts partials per field

__global__ void __launch_bounds__(kFinalBlock) winner_final_kernel(const double *partials, int nparts, LabelAux *aux,
                                                                    LabelAux *host, unsigned *stamp_word,
                                                                    unsigned stamp)
{
    constexpr int kNW = kFinalBlock / kWave;
    __shared__ double smem[kWinF * kNW];
    double s[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    double m[2] = {0.0, 0.0};
    for (int p = threadIdx.x; p < nparts; p += blockDim.x) {
        double pp[kWinF];
#pragma unroll
        for (int f = 0; f < kWinF; f++) pp[f] = partials[(size_t)f * nparts + p];
        s[0] += pp[0]; s[1] += pp[2]; s[2] += pp[4]; s[3] += pp[5]; s[4] += pp[6];
        m[0] = fmax(m[0], pp[1]);
        m[1] = fmax(m[1], pp[3]);
    }
    block_reduce<double, 5, 0, kNW>(s, smem);
    block_reduce<double, 2, 1, kNW>(m, smem);
    if (threadIdx.x == 0) {
        aux->diag_f[0] = s[0];
        aux->diag_f[1] = m[0];
        aux->diag_f[2] = s[1];
        aux->diag_f[3] = m[1];
        aux->diag_f[4] = s[2];
        aux->diag_f[5] = s[3];
        aux->diag_f[6] = s[4];
        aux->diag_f[7] = 0.0;
        // the statistics pass's counters (earlier launches) and these sums
        LabelAux r;
        for (int k = 0; k < 8; k++) r.diag_i[k] = aux->diag_i[k];
        for (int k = 0; k < 8; k++) r.diag_f[k] = aux->diag_f[k];
        *host = r;
        host_stamp(stamp_word, stamp);
    }
}

// update_previous_and_tiles (:284-306)
__global__ void __launch_bounds__(kBlock) previous_kernel(const int32_t *labels, int32_t *previous, int32_t *reached,
                                                          int32_t *tile_flags, int64_t n, int nx, int32_t winner,
                                                          int keep, int seed, int tile_size, int tiles_x)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        bool belongs = labels[i] >= 0 && labels[i] == winner;
        previous[i] = (belongs && keep) ? 1 : 0;
        if (belongs && seed) {
            reached[i] = 1;
            int row = (int)i / nx, col = (int)i - row * nx;
            atomicMax(&tile_flags[(row / tile_size) * tiles_x + col / tile_size], 1);
        }
    }
}

// mobilize (:309-376): zero-momentum Resting -> Mobile transfer on the winner
constexpr int kMobF = 5;
__global__ void __launch_bounds__(kBlock) mobilize_kernel(
    const void *z_base, void *resting, void *mobile, int f32, int ny, int nx, mv2_avalanche_config cfg,
    double area, double f6, mv2_avalanche_fields f, int32_t winner, double activation_time, double *partials)
{
    __shared__ double smem[kMobF * 4];
    double s[kMobF] = {0.0, 0.0, 0.0, 0.0, 0.0};   // total, new vol, re vol, new area, re area
    const int64_t n = (int64_t)ny * nx;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        if (f.labels[i] < 0 || f.labels[i] != winner || f.latch[i] != 0) continue;
        double rest = ld_field(resting, f32, i);
        double transfer = fmin(fmax(rest - ld_field(z_base, f32, i), 0.0), cfg.mobilization_depth * f.severity[i]);
        if (transfer <= cfg.dry_tolerance) continue;
        double mob = ld_mobile(mobile, f32, i);
        if (f32) {
            long long dq = (long long)floor(transfer * kInvQuantum);
            int32_t *mq = (int32_t *)mobile;
            long long room = 2147483647LL - (long long)mq[i];
            if (dq > room) dq = room;
            transfer = (double)dq * kQuantum;
            f.owned_surface[i] = rest + mob;
            ((float *)resting)[i] = (float)(rest - transfer);
            mq[i] = (int32_t)((long long)mq[i] + dq);
        } else {
            f.owned_surface[i] = rest + mob;
            ((double *)resting)[i] = rest - transfer;
            ((double *)mobile)[i] = mob + transfer;
        }
        f.owned_export_baseline[i] = f.export_cumulative[i];
        f.latch[i] = 1;
        const int r = (int)i / nx, c = (int)i - r * nx;
        const double w = cell_weight(cfg.weight_mode, ny, nx, area, f6, r, c);
        double weighted = transfer * w;
        int32_t prev = f.activation_count[i];
        f.activation_count[i] = prev + 1;
        if (prev == 0) {
            f.first_activation_time[i] = activation_time;
            s[1] += weighted;
            s[3] += w;
        } else {
            s[2] += weighted;
            s[4] += w;
        }
        f.last_activation_time[i] = activation_time;
        f.r2m_cumulative[i] = f.r2m_cumulative[i] + weighted;
        s[0] += weighted;
    }
    block_reduce<double, kMobF, 0>(s, smem);
    if (threadIdx.x == 0)
        for (int k = 0; k < kMobF; k++) partials[(size_t)k * gridDim.x + blockIdx.x] = s[k];
}

__global__ void __launch_bounds__(kBlock) mobilize_final_kernel(const double *partials, int nparts, double *out)
{
    __shared__ double smem[kMobF * 4];
    double s[kMobF] = {0.0, 0.0, 0.0, 0.0, 0.0};
    for (int p = threadIdx.x; p < nparts; p += blockDim.x)
        for (int k = 0; k < kMobF; k++) s[k] += partials[(size_t)k * nparts + p];
    block_reduce<double, kMobF, 0>(s, smem);
    if (threadIdx.x == 0)
        for (int k = 0; k < kMobF; k++) out[k] = s[k];
}

// reduce_persistence (:378-451): lifecycle scalars over every cell.  Per
// thread: 14 float sums/maxima (the reference's avalanche_persistence_float
// slots, maxima in slots 5 and 9 from zero) and 4 counts (exact in double);
// folded per workgroup field-parallel, then across workgroups in order.
constexpr int kPersF = 14, kPersV = 18;
constexpr unsigned kPersMax = (1u << 5) | (1u << 9);
__global__ void __launch_bounds__(kBlock) persistence_kernel(
    const void *initial, const void *resting, const void *mobile, const void *qx_, const void *qy_, int f32,
    int ny, int nx, double area, double f6, int weight_mode, mv2_avalanche_fields f, const double *m2r,
    double start_rad, double stop_rad, double dry, double activity_speed, int tile_size, int tiles_x,
    int32_t *tile_flags, double *partials)
{
    __shared__ double lds[kPersV * kBlock];
    double v[kPersV];
#pragma unroll
    for (int k = 0; k < kPersV; k++) v[k] = 0.0;
    const int64_t n = (int64_t)ny * nx;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int r = (int)i / nx, c = (int)i - r * nx;
        const double w = cell_weight(weight_mode, ny, nx, area, f6, r, c);
        const int32_t count = f.activation_count[i];
        const double h = ld_mobile(mobile, f32, i);
        double speed = 0.0;
        if (h > dry) {
            const double vx = ld_field(qx_, f32, i) / h, vy = ld_field(qy_, f32, i) / h;
            speed = sqrt(vx * vx + vy * vy);
            v[kPersF + 2] += 1.0;
            v[4] += h * w;
            v[5] = fmax(v[5], speed);
            if (speed > activity_speed) v[13] += h * w;
        }
        if (f.unstable[i] != 0 || (h > dry && speed > activity_speed))
            tile_flags[(r / tile_size) * tiles_x + c / tile_size] = 1;
        const double delta = ld_field(resting, f32, i) + h - ld_field(initial, f32, i);
        v[6] += delta * w;
        v[7] += fabs(delta) * w * 0.5;
        v[0] += f.r2m_cumulative[i];
        v[1] += m2r[i];
        if (count <= 0) continue;
        const double s = f.slope[i];
        v[kPersF + 0] += 1.0;
        v[2] += w;
        v[8] += s;
        v[9] = fmax(v[9], s);
        v[10] += s - stop_rad;
        v[11] += start_rad - s;
        if (count > 1) {
            v[kPersF + 1] += 1.0;
            v[3] += w;
        }
        if (h <= dry && s <= stop_rad && f.latch[i] == 0) {
            v[kPersF + 3] += 1.0;
            v[12] += w;
        }
    }
    block_fold<double, kPersV, kPersMax>(v, lds);
    if (threadIdx.x == 0)
        for (int k = 0; k < kPersV; k++) partials[(size_t)k * gridDim.x + blockIdx.x] = v[k];
}

// fold of the workgroup partials (field-parallel, workgroup order) and the
// count of flagged tiles
__global__ void __launch_bounds__(kBlock) persistence_final_kernel(const double *partials, int nparts,
                                                                   const int32_t *tile_flags, int tile_count,
                                                                   double *out)
{
    __shared__ double lds[(kPersV + 1) * kBlock];
    double v[kPersV + 1];
#pragma unroll
    for (int k = 0; k <= kPersV; k++) v[k] = 0.0;
    for (int p = threadIdx.x; p < nparts; p += kBlock)
        for (int k = 0; k < kPersV; k++) {
            const double x = partials[(size_t)k * nparts + p];
            v[k] = ((kPersMax >> k) & 1u) ? fmax(v[k], x) : v[k] + x;
        }
    for (int t = threadIdx.x; t < tile_count; t += kBlock) v[kPersV] += tile_flags[t] != 0 ? 1.0 : 0.0;
    block_fold<double, kPersV + 1, kPersMax>(v, lds);
    if (threadIdx.x == 0)
        for (int k = 0; k <= kPersV; k++) out[k] = v[k];
}

static int check_ctx(mv2_context *ctx)
{
    if (!ctx) return set_error(nullptr, MV2_EINVAL, "[GpuLargeAvalanche] null context");
    if (!ctx->bound) return set_error(ctx, MV2_ESTATE, "[GpuLargeAvalanche] bind state first");
    hipSetDevice(ctx->device);
    return MV2_OK;
}

// the border unions and the flattening (after the tile-local pass)
static int label_global(mv2_context *ctx, int32_t *labels, int ny, int nx, int connectivity, int32_t *counts,
                        unsigned long long *spread)
{
    const int tiles_x = (nx + kCcl - 1) / kCcl;
    const int tiles = tiles_x * ((ny + kCcl - 1) / kCcl);
    hipLaunchKernelGGL(ccl_border_kernel, dim3(grid_for((int64_t)tiles * 3 * kCcl)), dim3(kBlock), 0, ctx->stream,
                       labels, ny, nx, connectivity, tiles_x, tiles);
    hipLaunchKernelGGL(ccl_flatten_kernel, dim3(tiles), dim3(kBlock), 0, ctx->stream, labels, ny, nx, tiles_x,
                       counts, spread);
    MV2_HIP(ctx, hipGetLastError());
    return MV2_OK;
}

static int label_core(mv2_context *ctx, int32_t *labels, int ny, int nx, int connectivity, int32_t *counts)
{
    const int tiles_x = (nx + kCcl - 1) / kCcl;
    const int tiles = tiles_x * ((ny + kCcl - 1) / kCcl);
    hipLaunchKernelGGL(ccl_local_kernel, dim3(tiles), dim3(kBlock), 0, ctx->stream, labels, ny, nx, connectivity,
                       tiles_x);
    hipLaunchKernelGGL(ccl_border_kernel, dim3(grid_for((int64_t)tiles * 3 * kCcl)), dim3(kBlock), 0, ctx->stream,
                       labels, ny, nx, connectivity, tiles_x, tiles);
    hipLaunchKernelGGL(ccl_flatten_kernel, dim3(tiles), dim3(kBlock), 0, ctx->stream, labels, ny, nx, tiles_x,
                       counts, (unsigned long long *)nullptr);
    MV2_HIP(ctx, hipGetLastError());
    return MV2_OK;
}

}  // namespace mv2

using namespace mv2;

extern "C" int mv2_label_mask(mv2_context *ctx, const int32_t *mask, int32_t *labels, int ny, int nx,
                              int connectivity)
{
    if (!ctx) return set_error(nullptr, MV2_EINVAL, "[LargeAvalanche] null context");
    if (!mask || !labels || ny < 1 || nx < 1)
        return set_error(ctx, MV2_EINVAL, "[LargeAvalanche] connected mask must be 2-D");
    if (connectivity != 4 && connectivity != 8)
        return set_error(ctx, MV2_EINVAL, "[LargeAvalanche] connectivity must be 4 or 8");
    if ((int64_t)ny * nx >= ((int64_t)1 << 31))
        return set_error(ctx, MV2_EINVAL, "[LargeAvalanche] mask too large for int32 labels");
    hipSetDevice(ctx->device);
    const int64_t n = (int64_t)ny * nx;
    hipLaunchKernelGGL(init_from_mask, dim3(grid_for(n, kBlock, 4096)), dim3(kBlock), 0, ctx->stream, mask, labels, n);
    if (int rc = label_core(ctx, labels, ny, nx, connectivity, nullptr)) return rc;
    if (int rc_w = stream_wait(ctx, ctx->stream)) return rc_w;
    return MV2_OK;
}

extern "C" int mv2_avalanche_label(mv2_context *ctx, const mv2_avalanche_config *cfg,
                                   const mv2_avalanche_fields *f, mv2_label_result *out)
{
    if (int rc = check_ctx(ctx)) return rc;
    if (!cfg || !f || !out) return set_error(ctx, MV2_EINVAL, "[GpuLargeAvalanche] null argument");
    if (cfg->connectivity != 4 && cfg->connectivity != 8)
        return set_error(ctx, MV2_EINVAL, "[LargeAvalanche] connectivity must be 4 or 8");
    const int ny = ctx->ny, nx = ctx->nx;
    const int64_t n = (int64_t)ny * nx;
    const int f32 = ctx->precision == MV2_FP32;
    const double area = ctx->dx * ctx->dy, f6 = area / 6.0;
    if (int rc = ensure_scratch(ctx, ctx->label_aux, sizeof(LabelAux) + sizeof(double) * kWinF * kLabelParts)) return rc;
    LabelAux *aux = (LabelAux *)ctx->label_aux.ptr;
    double *partials = (double *)((char *)ctx->label_aux.ptr + sizeof(LabelAux));
    if (int rc = ensure_scratch(ctx, ctx->spread, kSpreadScratchBytes)) return rc;
    unsigned long long *spread = (unsigned long long *)ctx->spread.ptr + (size_t)kLabelSpreadSet * kSpreadWords;
    const InitArgs init{ctx->z_base, ctx->b_eff, ctx->mobile, ctx->qx, ctx->qy, f32, ny, nx, ctx->dx, ctx->dy,
                        *cfg, area, f6, *f};
    const int tiles_x = (nx + kCcl - 1) / kCcl;
    const int tiles = tiles_x * ((ny + kCcl - 1) / kCcl);
    if (f32)
        hipLaunchKernelGGL(avalanche_local_kernel<true>, dim3(tiles), dim3(kLocalBlock), 0, ctx->stream, init, tiles_x,
                           aux, (int32_t)(n + 1), spread);
    else
        hipLaunchKernelGGL(avalanche_local_kernel<false>, dim3(tiles), dim3(kLocalBlock), 0, ctx->stream, init,
                           tiles_x, aux, (int32_t)(n + 1), spread);
    if (int rc = label_global(ctx, f->labels, ny, nx, cfg->connectivity, f->counts, spread)) return rc;
    const int sgrid = grid_for(n, kBlock, kLabelParts);
    if (f32)
        hipLaunchKernelGGL(winner_stats_kernel<true>, dim3(sgrid), dim3(kBlock), 0, ctx->stream, f->labels,
                           ctx->z_base, ctx->b_eff, f->latch, f->previous, f->slope, f->severity, ny, nx, *cfg, area,
                           f6, aux, spread, partials);
    else
        hipLaunchKernelGGL(winner_stats_kernel<false>, dim3(sgrid), dim3(kBlock), 0, ctx->stream, f->labels,
                           ctx->z_base, ctx->b_eff, f->latch, f->previous, f->slope, f->severity, ny, nx, *cfg, area,
                           f6, aux, spread, partials);
    // the result goes straight to pinned host memory (no copy launch)
    LabelAux *hres = (LabelAux *)((char *)ctx->host_result.ptr + 2048);
    static_assert(2048 + sizeof(LabelAux) <= 4096, "the labelling result in the pinned block");
    const unsigned stamp = next_stamp(ctx, 1);
    hipLaunchKernelGGL(winner_final_kernel, dim3(1), dim3(kFinalBlock), 0, ctx->stream, partials, sgrid, aux, hres,
                       stamp_words(ctx), stamp);
    MV2_HIP(ctx, hipGetLastError());
    if (int rc_w = stamp_wait(ctx, ctx->stream, stamp, 1)) return rc_w;
    LabelAux host;
    memcpy(&host, (const void *)hres, sizeof(LabelAux));
    memcpy(out->diag_i, host.diag_i, sizeof(host.diag_i));
    memcpy(out->diag_f, host.diag_f, sizeof(host.diag_f));
    out->union_passes = 1;
    out->reserved = 0;
    return MV2_OK;
}

extern "C" int mv2_avalanche_previous(mv2_context *ctx, const mv2_avalanche_fields *f, int32_t winner,
                                      int keep_candidate, int seed_frontier, int32_t *frontier_reached,
                                      int32_t *tile_flags, int tile_size, int tiles_x, int tile_count)
{
    if (int rc = check_ctx(ctx)) return rc;
    if (!f || !frontier_reached || !tile_flags || tile_size < 1 || tile_count < 1)
        return set_error(ctx, MV2_EINVAL, "[GpuLargeAvalanche] invalid previous/tiles arguments");
    const int64_t n = (int64_t)ctx->ny * ctx->nx;
    hipLaunchKernelGGL(fill32, dim3(grid_for(tile_count)), dim3(kBlock), 0, ctx->stream, tile_flags,
                       (int64_t)tile_count, 0);
    hipLaunchKernelGGL(previous_kernel, dim3(grid_for(n, kBlock, 4096)), dim3(kBlock), 0, ctx->stream, f->labels,
                       f->previous, frontier_reached, tile_flags, n, ctx->nx, winner, keep_candidate, seed_frontier,
                       tile_size, tiles_x);
    MV2_HIP(ctx, hipGetLastError());
    return MV2_OK;
}

extern "C" int mv2_avalanche_mobilize(mv2_context *ctx, const mv2_avalanche_config *cfg,
                                      const mv2_avalanche_fields *f, int32_t winner, double activation_time,
                                      double diag_f[16])
{
    if (int rc = check_ctx(ctx)) return rc;
    if (!cfg || !f || !diag_f) return set_error(ctx, MV2_EINVAL, "[GpuLargeAvalanche] null argument");
    state_written(ctx);
    const int ny = ctx->ny, nx = ctx->nx;
    const int64_t n = (int64_t)ny * nx;
    const int f32 = ctx->precision == MV2_FP32;
    const double area = ctx->dx * ctx->dy, f6 = area / 6.0;
    if (int rc = ensure_scratch(ctx, ctx->label_aux, sizeof(LabelAux) + sizeof(double) * kWinF * kLabelParts)) return rc;
    double *partials = (double *)((char *)ctx->label_aux.ptr + sizeof(LabelAux));
    double *dres = (double *)((char *)ctx->result.ptr + 3584);
    const int sgrid = grid_for(n, kBlock, kLabelParts);
    hipLaunchKernelGGL(mobilize_kernel, dim3(sgrid), dim3(kBlock), 0, ctx->stream, ctx->z_base, ctx->b_eff,
                       ctx->mobile, f32, ny, nx, *cfg, area, f6, *f, winner, activation_time, partials);
    hipLaunchKernelGGL(mobilize_final_kernel, dim3(1), dim3(kBlock), 0, ctx->stream, partials, sgrid, dres);
    MV2_HIP(ctx, hipGetLastError());
    MV2_HIP(ctx, hipMemcpyAsync(ctx->host_result.ptr, dres, sizeof(double) * kMobF, hipMemcpyDeviceToHost, ctx->stream));
    if (int rc_w = stream_wait(ctx, ctx->stream)) return rc_w;
    double s[kMobF];
    memcpy(s, ctx->host_result.ptr, sizeof(s));
    for (int k = 0; k < 16; k++) diag_f[k] = 0.0;
    diag_f[7] = s[0];
    diag_f[10] = s[1];
    diag_f[11] = s[2];
    diag_f[12] = s[3];
    diag_f[13] = s[4];
    return MV2_OK;
}

extern "C" int mv2_avalanche_persistence(mv2_context *ctx, const mv2_avalanche_config *cfg,
                                         const mv2_avalanche_fields *f, const void *initial_b_eff,
                                         const double *m2r_cumulative, int tile_size, int tiles_x,
                                         int tile_count, int32_t *tile_flags, int32_t diag_i[4],
                                         double diag_f[14], int32_t *active_tile_count)
{
    if (int rc = check_ctx(ctx)) return rc;
    if (!cfg || !f || !initial_b_eff || !m2r_cumulative || !tile_flags || !diag_i || !diag_f ||
        !active_tile_count)
        return set_error(ctx, MV2_EINVAL, "[GpuLargeAvalanche] null argument");
    const int ny = ctx->ny, nx = ctx->nx;
    if (tile_size < 1 || tiles_x != (nx + tile_size - 1) / tile_size ||
        tile_count != tiles_x * ((ny + tile_size - 1) / tile_size))
        return set_error(ctx, MV2_EINVAL, "[GpuLargeAvalanche] tile grid does not match the state");
    const int64_t n = (int64_t)ny * nx;
    const int f32 = ctx->precision == MV2_FP32;
    const double area = ctx->dx * ctx->dy, f6 = area / 6.0;
    const size_t need = sizeof(LabelAux) + sizeof(double) * kPersV * kLabelParts;
    if (int rc = ensure_scratch(ctx, ctx->label_aux, need)) return rc;
    double *partials = (double *)((char *)ctx->label_aux.ptr + sizeof(LabelAux));
    double *dres = (double *)((char *)ctx->result.ptr + 1024);
    const int sgrid = grid_for(n, kBlock, kLabelParts);
    hipLaunchKernelGGL(fill32, dim3(grid_for(tile_count)), dim3(kBlock), 0, ctx->stream, tile_flags,
                       (int64_t)tile_count, 0);
    hipLaunchKernelGGL(persistence_kernel, dim3(sgrid), dim3(kBlock), 0, ctx->stream, initial_b_eff, ctx->b_eff,
                       ctx->mobile, ctx->qx, ctx->qy, f32, ny, nx, area, f6, cfg->weight_mode, *f, m2r_cumulative,
                       cfg->start_rad, cfg->stop_rad, cfg->dry_tolerance, cfg->activity_speed, tile_size, tiles_x,
                       tile_flags, partials);
    hipLaunchKernelGGL(persistence_final_kernel, dim3(1), dim3(kBlock), 0, ctx->stream, partials, sgrid,
                       tile_flags, tile_count, dres);
    MV2_HIP(ctx, hipGetLastError());
    MV2_HIP(ctx, hipMemcpyAsync(ctx->host_result.ptr, dres, sizeof(double) * (kPersV + 1), hipMemcpyDeviceToHost,
                                ctx->stream));
    if (int rc_w = stream_wait(ctx, ctx->stream)) return rc_w;
    double s[kPersV + 1];
    memcpy(s, ctx->host_result.ptr, sizeof(s));
    for (int k = 0; k < kPersF; k++) diag_f[k] = s[k];
    for (int k = 0; k < 4; k++) diag_i[k] = (int32_t)s[kPersF + k];
    *active_tile_count = (int32_t)s[kPersV];
    return MV2_OK;
}
