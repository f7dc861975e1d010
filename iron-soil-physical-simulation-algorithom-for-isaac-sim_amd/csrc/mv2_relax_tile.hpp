// Iterations of the angle-of-repose relaxation (slope_model.py:296-338) on
// one LDS tile, shared by the single-field relaxation (mv2_relax.hip) and the
// batched one (mv2_batch.hip).  The workgroup stages a tile plus a halo in
// LDS and performs all 8 red-black sub-phases of each iteration there.
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

// ---- several iterations per launch ------------------------------------------
// A relaxation launch runs up to K iterations on an LDS tile before
// writing it back.  A sub-phase moves information one cell along its
// direction, so K iterations leave at most 6K halo layers stale; with a halo
// of 6K + 2 the interior and every edge it owns stay exact for all K
// iterations.  The tile has a NaN border (one ring before, two after the
// computed region) and cells outside the grid are staged as NaN: an edge with
// a NaN end has excess max(|NaN| - crit, 0) = +0 and moves nothing (the valid
// end is written back unchanged), exactly as a skipped edge, so the loops
// carry no bounds tests.
//
// Each thread takes kSlots active edges of a sub-phase and also loads the
// cell after the edge's second cell: the edge from that second cell onwards
// has the opposite parity, so one set of loads gives the maximum excess over
// the owned edges of both parities in the sub-phase's input state.
//   X lanes (d = 0, 2, 3; parity on the column): anchor (lr, 2q + p)
//   Y lanes (d = 1; parity on the row):          anchor (2 yr + p, lc)
// COLS: consecutive lanes of the X mapping walk down a column pair (LDS
// stride PW, odd: no bank conflicts) rather than along a row (stride 2).
template <int TH_, int TW_, int BLOCK_, int K_, bool COLS_>
struct RelaxGeom {
    static constexpr int TH = TH_, TW = TW_, BLOCK = BLOCK_, K = K_;
    static constexpr bool COLS = COLS_;
    static constexpr int H = 6 * K + 2;
    static constexpr int LH = TH + 2 * H, LW = TW + 2 * H;   // computed region
    static constexpr int PW = LW + 3, PH = LH + 3;           // NaN border: 1 before, 2 after
    static constexpr int kCells = PH * PW;
    static constexpr int kStage = (kCells + BLOCK - 1) / BLOCK;
    static constexpr int kEdges = LH * LW / 2;               // active edges of one sub-phase
    static constexpr int kSlots = (kEdges + BLOCK - 1) / BLOCK;
    static constexpr int kOwned = (TH * TW + BLOCK - 1) / BLOCK;
    static_assert(TH % 2 == 0 && TW % 2 == 0 && H % 2 == 0, "even tile origins keep local and global parities equal");
    static_assert(BLOCK % kWave == 0 && BLOCK <= 1024, "whole waves");
};

template <class R> __device__ inline R relax_fabs(R x);
template <> __device__ inline float relax_fabs<float>(float x) { return fabsf(x); }
template <> __device__ inline double relax_fabs<double>(double x) { return fabs(x); }
// maximum of two non-NaN values, or the non-NaN one
template <class R> __device__ inline R relax_fmax(R x, R y);
template <> __device__ inline float relax_fmax<float>(float x, float y) { return fmaxf(x, y); }
template <> __device__ inline double relax_fmax<double>(double x, double y) { return fmax(x, y); }
template <class R> __device__ inline R relax_nan();
template <> __device__ inline float relax_nan<float>() { return __builtin_nanf(""); }
template <> __device__ inline double relax_nan<double>() { return __builtin_nan(""); }

// this thread's share of the padded tile with interior origin (r0, c0), read
// from hs into registers; bit j of the returned mask is set where v[j] is a
// grid cell of the computed region (the staging writes NaN elsewhere).
// The loads are unconditional (indices clamped into the grid) and the NaN
// selection happens at the staging, in a later block: a select next to the
// load is turned into a load under a branch, and the compiler's waits after
// such branches fall back to waiting for every load in flight.
template <class R, class G>
__device__ inline unsigned relax_tile_load(const R *hs, int ny, int nx, int r0, int c0, R (&v)[G::kStage])
{
    constexpr int H = G::H, LH = G::LH, LW = G::LW, PW = G::PW, BLOCK = G::BLOCK;
    static_assert(G::kStage <= 32, "one validity bit per staged cell");
    const int R0 = r0 - H - 1, C0 = c0 - H - 1;         // global cell of padded (0, 0)
    unsigned mask = 0u;
#pragma unroll
    for (int j = 0; j < G::kStage; j++) {
        const int e = threadIdx.x + j * BLOCK;
        const int pr = e / PW, pc = e - pr * PW;
        const int gr = R0 + pr, gc = C0 + pc;
        const bool inner = pr >= 1 && pr <= LH && pc >= 1 && pc <= LW;
        const bool in = e < G::kCells && inner && gr >= 0 && gr < ny && gc >= 0 && gc < nx;
        const unsigned o = (unsigned)(min(max(gr, 0), ny - 1) * nx + min(max(gc, 0), nx - 1)) * (unsigned)sizeof(R);
        v[j] = *reinterpret_cast<const R *>(reinterpret_cast<const char *>(hs) + o);
        mask |= in ? 1u << j : 0u;
    }
    return mask;
}

// n (1 <= n <= G::K) iterations on the tile with interior origin (r0, c0) of
// the field hs (ny x nx), written to hd; tmaxk[i] = this thread's maximum
// excess over the owned edges in iteration i, taken before each sub-phase.
// Per edge (slope_model.py:318-331): diff = a - b, excess =
// max(|diff| - crit, 0), transfer (excess / 2) sgn(diff) -- written as
// diff < 0 ? -half : half, equal to it for every diff (half is +0 whenever
// diff is 0 or NaN).  All threads of the block must call it.  With `changed`,
// the thread reports whether an interior cell it writes differs from its
// staged value.
// The _staged form takes the tile as read into registers by relax_tile_load
// (the caller may issue those loads early).
template <class R, class G>
__device__ inline void relax_tile_iterations_staged(const R (&v)[G::kStage], unsigned vmask, R *hd, int ny, int nx,
                                                    const R (&crits)[4], int r0, int c0, R *tile, int n,
                                                    R (&tmaxk)[G::K], bool *changed = nullptr)
{
    constexpr int H = G::H, LH = G::LH, LW = G::LW, PW = G::PW, TH = G::TH, TW = G::TW, BLOCK = G::BLOCK;
#pragma unroll
    for (int j = 0; j < G::kStage; j++) {
        const int e = threadIdx.x + j * BLOCK;
        if (e < G::kCells) tile[e] = (vmask >> j) & 1u ? v[j] : relax_nan<R>();
    }
    // per-slot anchors and ownership (anchor in the interior) of the active
    // edge and of the opposite-parity edge continuing from its second cell
    int bx[G::kSlots], by[G::kSlots];
    bool valid[G::kSlots];
    bool xa[G::kSlots][2], x0[G::kSlots][2], x2[G::kSlots][2], x3[G::kSlots][2];
    bool ya[G::kSlots][2], yi[G::kSlots][2];
#pragma unroll
    for (int j = 0; j < G::kSlots; j++) {
        const int e0 = threadIdx.x + j * BLOCK;
        valid[j] = G::kEdges % BLOCK == 0 || e0 < G::kEdges;
        const int e = valid[j] ? e0 : 0;
        int lr, q;
        if constexpr (G::COLS) {
            q = e / LH;
            lr = e - q * LH;
        } else {
            lr = e / (LW / 2);
            q = e - lr * (LW / 2);
        }
        // an unused slot reads inside the tile and stores nothing
        bx[j] = valid[j] ? (lr + 1) * PW + 2 * q + 1 : PW + 1;
        const int yr = e / LW, lc = e - yr * LW;
        by[j] = valid[j] ? (2 * yr + 1) * PW + lc + 1 : PW + 1;
        auto in_r = [](int r) { return (unsigned)(r - H) < (unsigned)TH; };
        auto in_c = [](int c) { return (unsigned)(c - H) < (unsigned)TW; };
#pragma unroll
        for (int p = 0; p < 2; p++) {
            const bool ca = in_c(2 * q + p), ci = in_c(2 * q + p + 1);
            xa[j][p] = valid[j] && in_r(lr) && ca;
            x0[j][p] = valid[j] && in_r(lr) && ci;
            x2[j][p] = valid[j] && in_r(lr + 1) && ci;
            x3[j][p] = valid[j] && in_r(lr - 1) && ci;
            ya[j][p] = valid[j] && in_c(lc) && in_r(2 * yr + p);
            yi[j][p] = valid[j] && in_c(lc) && in_r(2 * yr + p + 1);
        }
    }
    __syncthreads();
    R orig[G::kOwned];
    if (changed) {
#pragma unroll
        for (int j = 0; j < G::kOwned; j++) {
            const int e = threadIdx.x + j * BLOCK;
            const int ir = e / TW, ic = e - ir * TW;
            orig[j] = e < TH * TW ? tile[(ir + H + 1) * PW + ic + H + 1] : (R)0;
        }
    }
#pragma unroll 1
    for (int it = 0; it < n; it++) {
        R tmax = (R)0;
#pragma unroll
        for (int d = 0; d < 4; d++) {
            const R crit = crits[d];
            // a, b and the cell after b along the direction, from the anchor
            const int oa = d == 3 ? PW : 0;
            const int ob = d == 0 ? 1 : d == 1 ? PW : d == 2 ? PW + 1 : 1;
            const int on = d == 0 ? 2 : d == 1 ? 2 * PW : d == 2 ? 2 * PW + 2 : 2 - PW;
#pragma unroll
            for (int p = 0; p < 2; p++) {
                R va[G::kSlots], vb[G::kSlots], vn[G::kSlots];
#pragma unroll
                for (int j = 0; j < G::kSlots; j++) {
                    const int base = d == 1 ? by[j] + p * PW : bx[j] + p;
                    va[j] = tile[base + oa];
                    vb[j] = tile[base + ob];
                    vn[j] = tile[base + on];
                }
                __syncthreads();      // every input of the sub-phase read before any transfer
#pragma unroll
                for (int j = 0; j < G::kSlots; j++) {
                    const bool own_a = d == 1 ? ya[j][p] : xa[j][p];
                    const bool own_i = d == 0 ? x0[j][p] : d == 1 ? yi[j][p] : d == 2 ? x2[j][p] : x3[j][p];
                    const R diff = va[j] - vb[j];
                    const R ex = relax_fmax<R>(relax_fabs<R>(diff) - crit, (R)0);
                    const R exi = relax_fmax<R>(relax_fabs<R>(vb[j] - vn[j]) - crit, (R)0);
                    tmax = relax_fmax<R>(tmax, own_a ? ex : (R)0);
                    tmax = relax_fmax<R>(tmax, own_i ? exi : (R)0);
                    const R half = (R)0.5 * ex;
                    const R tr = diff < (R)0 ? -half : half;
                    if (valid[j]) {
                        const int base = d == 1 ? by[j] + p * PW : bx[j] + p;
                        tile[base + oa] = va[j] - tr;
                        tile[base + ob] = vb[j] + tr;
                    }
                }
                __syncthreads();
            }
        }
#pragma unroll
        for (int k = 0; k < G::K; k++)
            if (k == it) tmaxk[k] = tmax;
    }
#pragma unroll
    for (int j = 0; j < G::kOwned; j++) {
        const int e = threadIdx.x + j * BLOCK;
        const int ir = e / TW, ic = e - ir * TW;
        const int r = r0 + ir, c = c0 + ic;
        if (e < TH * TW && r < ny && c < nx) {
            const R v = tile[(ir + H + 1) * PW + ic + H + 1];
            hd[r * nx + c] = v;
            if (changed && v != orig[j]) *changed = true;
        }
    }
}

template <class R, class G>
__device__ inline void relax_tile_iterations(const R *hs, R *hd, int ny, int nx, const R (&crits)[4], int r0,
                                             int c0, R *tile, int n, R (&tmaxk)[G::K], bool *changed = nullptr)
{
    R v[G::kStage];
    const unsigned vmask = relax_tile_load<R, G>(hs, ny, nx, r0, c0, v);
    relax_tile_iterations_staged<R, G>(v, vmask, hd, ny, nx, crits, r0, c0, tile, n, tmaxk, changed);
}

}  // namespace mv2
