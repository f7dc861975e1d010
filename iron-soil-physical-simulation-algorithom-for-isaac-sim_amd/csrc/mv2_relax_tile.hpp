// One iteration of the angle-of-repose relaxation (slope_model.py:296-338)
// on one tile, shared by the single-field relaxation (mv2_relax.hip) and the
// batched one (mv2_batch.hip).  The workgroup stages a tile plus an 8-cell
// halo in LDS and performs all 8 red-black sub-phases of the iteration there.
// A sub-phase moves information one cell along its direction, so after the 8
// sub-phases (two per direction, directions x, y and both diagonals) at most 6
// halo layers are stale; the interior and every edge it owns are exact.
//
// Directions and parities follow the reference's legacy H[x, y] convention
// behind MinimumSlopeAdapter (pkg/solvers/minimum_slope_adapter.py:170-179):
// on the canonical H[y, x] field with view index (r, c):
//   d0 legacy (1,0):  a = (r, c),   b = (r, c+1),   parity on c, dist dx
//   d1 legacy (0,1):  a = (r, c),   b = (r+1, c),   parity on r, dist dy
//   d2 legacy (1,1):  a = (r, c),   b = (r+1, c+1), parity on c, dist hypot
//   d3 legacy (1,-1): a = (r+1, c), b = (r, c+1),   parity on c, dist hypot
#pragma once

#include "mv2_internal.hpp"

namespace mv2 {

constexpr int kRelaxHalo = 8;
constexpr int kRelaxTH = 32;   // interior rows per tile
constexpr int kRelaxTW = 48;   // interior columns per tile
constexpr int kRelaxLH = kRelaxTH + 2 * kRelaxHalo;
constexpr int kRelaxLW = kRelaxTW + 2 * kRelaxHalo;

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

// One iteration on the tile with interior origin (r0, c0) of the field hs
// (ny x nx), written to hd; returns this thread's maximum excess over the edges
// the tile owns, taken before each sub-phase.  Every phase first loads all the
// operands a thread needs, then computes, then stores: the edges of one parity
// are disjoint, so the loads of one thread's edges never depend on its own
// stores.  All threads of the block must call it.  With `changed`, the thread
// also reports whether any interior cell it writes differs from its value in hs
// (compared with the staged copy, not re-read).
template <class R>
__device__ inline R relax_tile_iteration(const R *hs, R *hd, int ny, int nx, const R (&crits)[4], int r0, int c0,
                                         R *tile, bool *changed = nullptr)
{
    const int R0 = r0 - kRelaxHalo, C0 = c0 - kRelaxHalo;
    // stage tile + halo in one round trip (cells outside the grid are never touched)
    {
        R v[kRelaxStage];
#pragma unroll
        for (int j = 0; j < kRelaxStage; j++) {
            const int e = threadIdx.x + j * kRelaxBlock;
            const int gr = R0 + e / kRelaxLW, gc = C0 + e % kRelaxLW;
            v[j] = (R)0;
            if (gr >= 0 && gr < ny && gc >= 0 && gc < nx) v[j] = hs[gr * nx + gc];
        }
#pragma unroll
        for (int j = 0; j < kRelaxStage; j++) tile[threadIdx.x + j * kRelaxBlock] = v[j];
    }
    __syncthreads();
    // the interior cells this thread writes back, as staged (change tracking)
    R orig[kRelaxOwned];
    if (changed) {
#pragma unroll
        for (int j = 0; j < kRelaxOwned; j++) {
            const int e = threadIdx.x + j * kRelaxBlock;
            orig[j] = tile[(e / kRelaxTW + kRelaxHalo) * kRelaxLW + e % kRelaxTW + kRelaxHalo];
        }
    }
    R tmax = (R)0;
    // the 8 sub-phases unrolled: the direction and parity are compile-time
    // constants in each, so the edge geometry folds into the addressing
    // (701^2 relaxation to rest: 11.6 ms against 17.0 ms with rolled loops)
#pragma unroll
    for (int d = 0; d < 4; d++) {
        const R crit = crits[d];
#pragma unroll
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
        if (r < ny && c < nx) {
            const R v = tile[(ir + kRelaxHalo) * kRelaxLW + ic + kRelaxHalo];
            hd[r * nx + c] = v;
            if (changed && v != orig[j]) *changed = true;
        }
    }
    return tmax;
}

}  // namespace mv2
