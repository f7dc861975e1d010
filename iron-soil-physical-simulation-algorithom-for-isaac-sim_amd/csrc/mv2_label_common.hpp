// Union-find machinery and field loads shared by the single-terrain labelling
// (mv2_label.hip) and the batched one (mv2_batch_ops.hip).
#pragma once

#include "mv2_internal.hpp"

namespace mv2 {

__device__ inline double ld_field(const void *p, int f32, int64_t i)
{
    return f32 ? (double)((const float *)p)[i] : ((const double *)p)[i];
}
__device__ inline double ld_mobile(const void *p, int f32, int64_t i)
{
    return f32 ? (double)((const int32_t *)p)[i] * kQuantum : ((const double *)p)[i];
}

__device__ inline double clampd(double x, double lo, double hi) { return x < lo ? lo : (x > hi ? hi : x); }

__device__ inline int32_t load_parent(int32_t *p, int32_t i)
{
    return __hip_atomic_load(&p[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// root of x with path halving; parents satisfy p[x] <= x, roots p[r] == r
__device__ inline int32_t find_root(int32_t *p, int32_t x)
{
    int32_t cur = load_parent(p, x);
    if (cur != x) {
        int32_t prev = x, next;
        while (cur > (next = load_parent(p, cur))) {
            __hip_atomic_store(&p[prev], next, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            prev = cur;
            cur = next;
        }
    }
    return cur;
}

__device__ inline void unite(int32_t *p, int32_t a, int32_t b)
{
    int32_t ra = find_root(p, a), rb = find_root(p, b);
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

// ---------------------------------------------------------------------------
// Tile-based union-find.  The grid is cut into 32 x 32 tiles, one workgroup
// each.  (1) ccl_local_kernel: union-find of the tile's cells in LDS with
// workgroup-scope compare-and-swap, then every masked cell gets the global
// index of its tile-local root (the minimum index of its tile-local
// component: row-major order inside a tile agrees with global row-major
// order).  (2) ccl_border_kernel: the edges that cross a tile boundary are
// united in the global parent array with the same larger-root-onto-smaller
// hooking, so the final root of a component is its minimum linear index.
// (3) ccl_flatten_kernel: every cell follows its parents to the root and the
// tile counts cells per root in an LDS table before one global atomicAdd per
// distinct root, so a component spanning many tiles costs one atomic per tile
// instead of one per cell.
// ---------------------------------------------------------------------------
constexpr int kCcl = 32;
constexpr int kCclCells = kCcl * kCcl;
constexpr int kCclPer = kCclCells / kBlock;
constexpr int kCclHash = 128;
static_assert(kCclCells % kBlock == 0, "tile must split evenly over the block");

__device__ inline int lds_load(int *p, int i)
{
    return __hip_atomic_load(&p[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// root of local cell x with path halving (parents satisfy lp[x] <= x)
__device__ inline int lds_find(int *lp, int x)
{
    int cur = lds_load(lp, x);
    if (cur != x) {
        int prev = x, next;
        while (cur > (next = lds_load(lp, cur))) {
            __hip_atomic_store(&lp[prev], next, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            prev = cur;
            cur = next;
        }
    }
    return cur;
}

__device__ inline void lds_unite(int *lp, int a, int b)
{
    int ra = lds_find(lp, a), rb = lds_find(lp, b);
    while (ra != rb) {
        if (ra < rb) {
            int expected = rb;
            if (__hip_atomic_compare_exchange_strong(&lp[rb], &expected, ra, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                     __HIP_MEMORY_SCOPE_WORKGROUP))
                return;
            rb = lds_find(lp, expected);
        } else {
            int expected = ra;
            if (__hip_atomic_compare_exchange_strong(&lp[ra], &expected, rb, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                     __HIP_MEMORY_SCOPE_WORKGROUP))
                return;
            ra = lds_find(lp, expected);
        }
    }
}

}  // namespace mv2
