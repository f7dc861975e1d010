// State of the bound batch of terrains (one batch per context), shared by the
// batched step and relaxation (mv2_batch.hip) and the batched exchange and
// labelling (mv2_batch_ops.hip).
#pragma once

#include "mv2_internal.hpp"

#include <vector>

using mv2::Scratch;

struct mv2_batch_state {
    int count = 0;
    int64_t cells = 0;
    double dx = 0.0, dy = 0.0;
    float *b = nullptr;
    int32_t *h = nullptr;
    float *qx = nullptr, *qy = nullptr;
    std::vector<int32_t> ny, nx;
    std::vector<int64_t> off;
    int shape = 0;
    int tiles = 0;
    Scratch inst;                        // ny, nx, off, first tile, tile count
    Scratch tiles_tab;                   // tile -> instance, r0, c0
    Scratch alt_h, alt_qx, alt_qy;
    Scratch ctl, wave, stat, ledger, rec;
    int predicted = 0;
    // relaxation: 44 x 48 tiles of every instance, ping-pong field, loop state
    Scratch rtab;                        // first tile, tile count per instance; tile -> instance, r0, c0
    int rtiles = 0;
    Scratch relax_alt, relax_ctl, relax_spread, relax_out;
    int relax_predicted = 0;
    std::vector<int32_t> rfirst, rcount;  // host copy: first relaxation tile and tile count per instance
    Scratch relax_list, relax_list_host;  // tiles of the instances still relaxing (device / pinned staging)
    Scratch relax_quiet;                  // per tile: last excess (double), changed flags [2] (int32)
    // labelling: 32 x 32 tiles of every instance (built on first use), per-root
    // counts, per-instance statistics; entrainment: quanta moved per instance
    Scratch ctab;
    int ctiles = 0;
    Scratch label_counts, label_stats, moved;
    int frame = 0;                       // steps taken: the parity picks the loop-state copy
    int clean = -1;                      // parity whose loop state is known zeroed, -1 none
    // graphs of eight substep launches, one per parity
    hipGraphExec_t graph[2] = {nullptr, nullptr};
    hipStream_t capture = nullptr;
    hipEvent_t ev_sub0 = nullptr, ev_sub1 = nullptr;   // bracket the substep launches
    std::vector<unsigned char> graph_key[2];
};

// device table of the instances: ny, nx (int32 each), offset (int64)
struct BatchInstTables {
    const int32_t *ny, *nx;
    const int64_t *off;
};
inline BatchInstTables batch_inst_tables(const mv2_batch_state *bs)
{
    const int32_t *inst = (const int32_t *)bs->inst.ptr;
    const size_t n = (size_t)bs->count;
    return {inst, inst + n, (const int64_t *)(inst + 2 * n)};
}
