// One edge of a compact colour batch (pkg/solvers/warp_compact_frontier.py:57-130),
// shared by the per-phase entry points and the on-device round scheduler so
// that both move exactly the same bits.
#pragma once

#include "mv2_internal.hpp"

namespace mv2 {

// Weighted transfer across edge (a, b) when either end has been reached and
// the height difference exceeds the critical difference: the excess moves so
// that w_a h_a + w_b h_b is unchanged; both ends become reached.  Returns
// true when material moved.
__device__ inline bool frontier_edge_transfer(double *height, const double *weights, int32_t *reached,
                                              int64_t a, int64_t b, double critical_difference)
{
    // the reach flags and heights are loaded together (one memory round trip
    // instead of two dependent ones); the heights are used only when reached
    const int32_t ra = reached[a], rb = reached[b];
    const double ha = height[a], hb = height[b];
    if (ra == 0 && rb == 0) return false;
    const double difference = ha - hb;
    const double excess = fabs(difference) - critical_difference;
    if (!(excess > 0.0)) return false;
    const double signed_excess = difference < 0.0 ? -excess : excess;
    const double wa = weights[a], wb = weights[b];
    const double ws = wa + wb;
    const double ta = signed_excess * wb / ws;
    const double tb = signed_excess * wa / ws;
    height[a] = ha - ta;
    height[b] = hb + tb;
    reached[a] = 1;
    reached[b] = 1;
    return true;
}

// The scan's test of one edge: reached at either end and an excess over the
// critical difference above the tolerance.
__device__ inline bool frontier_edge_unstable(const double *height, const int32_t *reached, int64_t a,
                                              int64_t b, double critical_difference, double tolerance)
{
    const int32_t ra = reached[a], rb = reached[b];
    const double ha = height[a], hb = height[b];
    if (ra == 0 && rb == 0) return false;
    const double excess = fabs(ha - hb) - critical_difference;
    return excess > tolerance;
}

}  // namespace mv2
