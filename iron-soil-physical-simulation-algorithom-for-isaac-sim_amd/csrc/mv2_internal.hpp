// Internal definitions shared by the libmv2 translation units (gfx950, wave64).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "mv2.h"

namespace mv2 {

// Mobile thickness quantum of the FP32 storage mode: 2^-24 m (6.0e-8 m).
// Mass moves between cells as whole quanta, so FP32 runs conserve mass exactly
// and int32 storage covers 0 .. 128 m of mobile depth (material pooling in a
// pit of a steep pile can exceed 8 m, the range of a 2^-28 m quantum).
constexpr double kQuantum = 1.0 / 16777216.0;
constexpr double kInvQuantum = 16777216.0;

constexpr int kBlock = 256;      // threads per workgroup for streaming kernels
constexpr int kWave = 64;

struct Scratch {
    void *ptr = nullptr;
    size_t bytes = 0;
};

// A captured launch sequence of mv2_step (begin, the predicted substeps, the
// end / fold passes).  `key` holds the bytes of every launch argument of the
// sequence, so a replay is valid exactly when a new call's arguments equal it.
struct StepGraph {
    std::vector<unsigned char> key;
    hipGraphExec_t exec = nullptr;
    unsigned long long last_use = 0;
};

}  // namespace mv2

struct mv2_batch_state;

struct mv2_context {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string last_error;
    int cu_count = 0;
    // bound state
    int bound = 0;
    int precision = MV2_FP64;
    int ny = 0, nx = 0;
    double dx = 0.0, dy = 0.0;
    void *z_base = nullptr, *b_eff = nullptr, *mobile = nullptr;
    void *qx = nullptr, *qy = nullptr;
    // scratch owned by the library (sized at bind / grown on demand, never per substep)
    mv2::Scratch alt_h, alt_qx, alt_qy;   // ping-pong state of the substep loop
    mv2::Scratch sync;                    // barrier counters, reduction slots, flags
    mv2::Scratch partials;                // per-workgroup partial sums
    mv2::Scratch result;                  // device-side record staging
    mv2::Scratch relax_alt;               // ping-pong height of the relaxation
    mv2::Scratch relax_quiet;             // relaxation tiles: last excess (double), changed flags [2] (int32)
    mv2::Scratch env_moved;               // envelope tiles: interior changed in launch k, [2][tiles] (int32)
    mv2::Scratch nn_mask;                 // warm start: cells moved by the first mass projection (uint8)
    mv2::Scratch host_result;             // pinned host staging
    mv2::Scratch host_stamps;             // pinned: completion stamps of a call's last launch
    unsigned stamp_seq = 0;               // last stamp handed to a launch
    hipEvent_t ev_wait = nullptr;         // marker that stream_wait polls
    mv2::Scratch label_aux;               // labelling partials
    mv2::Scratch nn_partials;             // weight-gradient partials of the surrogate
    mv2::Scratch nn_loss;                 // loss / projection partials
    mv2::Scratch step_diag;               // per-workgroup substep diagnostics
    mv2::Scratch spread;                  // spread-maximum slots of the step and the relaxation
    mv2::Scratch frontier;                // round scheduler: control records, tile stamps, owner flags
    hipEvent_t ev_start = nullptr, ev_stop = nullptr;   // device timing of step calls
    int timing = 0;                       // bracket each substep launch with events
    std::vector<hipEvent_t> sub_events;   // event pairs of the substep launches
    int v2_predicted = 0;                 // substep launches to enqueue up front
    int v2_frame = 0;                     // steps taken (the parity picks the loop-state block)
    int v2_clean = -1;                    // parity whose loop state is known zeroed, -1 none
    int v2_spec = 1;                      // FP32 steps may speculate (mv2_set_speculation)
    int v2_pred_ready = 0;                // the last step's end pass left a CFL prediction and no library
                                          // call has written mobile or momentum since
    long long spec_steps = 0;             // steps that speculated, and those that restarted
    long long spec_restarts = 0;
    int v2_last_sync = 0;                 // parity of the last completed step
    double v2_static[2] = {0.0, 0.0};     // static diagnostic of the last step's input: max |grad H|, max |q|
    const void *ext_x = nullptr;          // per-cell external acceleration (state precision), or null
    const void *ext_y = nullptr;
    double ext_max = 0.0;                 // max |external| over the cells (static exit test)
    int relax_predicted = 0;              // relaxation launches that worked in the last call
    int relax_clean = 0;                  // the relaxation's loop state and next first group are zero
    unsigned relax_calls = 0;             // completed relaxation calls (parity picks the first group)
    int frontier_predicted = 0;           // frontier rounds to enqueue up front
    int env_gen = 1;                      // generation of the envelope's launch-0 tile flags (last used)
    int env_predicted[2] = {0, 0};        // envelope launches the last call needed: batch, single field
    mv2_batch_state *batch = nullptr;     // batched terrains bound to this context
    double *noalloc_scratch = nullptr;    // device address of the allocation-free ledger's partials
    mv2::Scratch patch;                   // FailureZone activation: per-cell flags, partials, result
    int graphs = 0;                       // replay cached HIP graphs of mv2_step's launch sequence (opt-in)
    hipStream_t capture_stream = nullptr; // private stream the sequences are captured on
    std::vector<mv2::StepGraph> step_graphs;               // instantiated sequences (at most kStepGraphsMax)
    std::vector<std::vector<unsigned char>> step_seen;     // keys seen once (captured on the second sighting)
    unsigned long long graph_clock = 0;
    long long graph_replays = 0;          // mv2_step calls served by a replay
};

void mv2_batch_release(mv2_context *ctx);

namespace mv2 {

int set_error(mv2_context *ctx, int code, const char *fmt, ...);
int check_hip(mv2_context *ctx, hipError_t err, const char *what);
int ensure_scratch(mv2_context *ctx, Scratch &s, size_t bytes, bool host_pinned = false);
// poll a marker behind the stream's work until it is done (blocking sync after 2 ms)
int stream_wait(mv2_context *ctx, hipStream_t st);
// Completion stamps.  The last launch of a call writes its results to pinned
// host memory; every thread that wrote one then stamps its own word with the
// call's number (`host_stamp`), and the host spins on those words instead of
// on the stream.  next_stamp returns 0 (no stamps: the caller waits on the
// stream) when more words are asked for than exist.
constexpr int kStampWords = 256;
unsigned next_stamp(mv2_context *ctx, int words);
unsigned *stamp_words(mv2_context *ctx);
// wait until words 0 .. words - 1 carry `stamp` (stream_wait for stamp 0,
// after 2 ms, or on a device error)
int stamp_wait(mv2_context *ctx, hipStream_t st, unsigned stamp, int words);
// a library call wrote mobile thickness or momentum: the next step takes its
// CFL maximum from a begin pass instead of the previous step's prediction
inline void state_written(mv2_context *ctx) { ctx->v2_pred_ready = 0; }
// device address of the allocation-free ledger's file-scope partials (mv2_exchange.hip)
int noalloc_scratch_address(mv2_context *ctx, double **out);

#define MV2_HIP(ctx, expr)                                     \
    do {                                                       \
        hipError_t _e = (expr);                                \
        if (_e != hipSuccess) return mv2::check_hip(ctx, _e, #expr); \
    } while (0)

// ---- reductions (wave64, DPP) ----------------------------------------------------
// A wave reduction is a Hillis-Steele scan inside each 16-lane row (row_shr 1,
// 2, 4, 8), then row_bcast:15 and row_bcast:31 carry the row totals up to lane
// 63, which is read back as a scalar.  Lanes without a source keep the
// identity.  Each step is a VALU move with a fixed lane pattern, so the
// result is deterministic; it replaces LDS-crossbar shuffles, whose latency
// dominated multi-value block reductions.
template <int CTRL, int ROW_MASK>
__device__ inline unsigned dpp_u32(unsigned old, unsigned v)
{
    return (unsigned)__builtin_amdgcn_update_dpp((int)old, (int)v, CTRL, ROW_MASK, 0xf, false);
}

template <int CTRL, int ROW_MASK, class T>
__device__ inline T dpp_move(T old, T v)
{
    static_assert(sizeof(T) == 4 || sizeof(T) == 8, "32- or 64-bit values");
    if constexpr (sizeof(T) == 4) {
        return __builtin_bit_cast(T, dpp_u32<CTRL, ROW_MASK>(__builtin_bit_cast(unsigned, old),
                                                              __builtin_bit_cast(unsigned, v)));
    } else {
        struct W { unsigned lo, hi; };
        W o = __builtin_bit_cast(W, old), x = __builtin_bit_cast(W, v), r;
        r.lo = dpp_u32<CTRL, ROW_MASK>(o.lo, x.lo);
        r.hi = dpp_u32<CTRL, ROW_MASK>(o.hi, x.hi);
        return __builtin_bit_cast(T, r);
    }
}

template <int OP, class T>  // OP 0 sum, 1 max, 2 min
__device__ inline T red_op(T a, T b)
{
    if constexpr (OP == 0) return a + b;
    else if constexpr (OP == 1) return b > a ? b : a;
    else return b < a ? b : a;
}

template <class T> struct RedLimits;
template <> struct RedLimits<float> {
    __device__ static float lowest() { return -__builtin_huge_valf(); }
    __device__ static float highest() { return __builtin_huge_valf(); }
};
template <> struct RedLimits<double> {
    __device__ static double lowest() { return -__builtin_huge_val(); }
    __device__ static double highest() { return __builtin_huge_val(); }
};
template <> struct RedLimits<int> {
    __device__ static int lowest() { return (int)0x80000000u; }
    __device__ static int highest() { return 0x7fffffff; }
};
template <> struct RedLimits<long long> {
    __device__ static long long lowest() { return (long long)0x8000000000000000ull; }
    __device__ static long long highest() { return 0x7fffffffffffffffll; }
};

template <int OP, class T>
__device__ inline T wave_reduce(T v)
{
    const T id = OP == 0 ? (T)0 : (OP == 1 ? RedLimits<T>::lowest() : RedLimits<T>::highest());
    v = red_op<OP>(v, dpp_move<0x111, 0xf>(id, v));   // row_shr:1
    v = red_op<OP>(v, dpp_move<0x112, 0xf>(id, v));   // row_shr:2
    v = red_op<OP>(v, dpp_move<0x114, 0xf>(id, v));   // row_shr:4
    v = red_op<OP>(v, dpp_move<0x118, 0xf>(id, v));   // row_shr:8
    v = red_op<OP>(v, dpp_move<0x142, 0xa>(id, v));   // row_bcast:15 into rows 1, 3
    v = red_op<OP>(v, dpp_move<0x143, 0xc>(id, v));   // row_bcast:31 into rows 2, 3
    if constexpr (sizeof(T) == 4) {
        return __builtin_bit_cast(T, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), 63));
    } else {
        struct W { int lo, hi; };
        W w = __builtin_bit_cast(W, v);
        w.lo = __builtin_amdgcn_readlane(w.lo, 63);
        w.hi = __builtin_amdgcn_readlane(w.hi, 63);
        return __builtin_bit_cast(T, w);
    }
}
template <class T> __device__ inline T wave_sum(T v) { return wave_reduce<0>(v); }
template <class T> __device__ inline T wave_max(T v) { return wave_reduce<1>(v); }
template <class T> __device__ inline T wave_min(T v) { return wave_reduce<2>(v); }

// Block-wide reduction of NV values; result valid in thread 0.  NW is the
// number of waves of the block (blockDim.x / 64) and `smem` must hold NV * NW
// elements.  Waves are combined in fixed order, so the result is
// deterministic for a given launch shape.  Thread k < NV folds value k (a
// single thread folding every value serialises hundreds of dependent LDS
// reads on one wave).  All threads of the block must call it (full waves).
template <class T, int NV, int OP, int NW = kBlock / kWave>  // OP 0 sum, 1 max, 2 min
__device__ inline void block_reduce(T (&v)[NV], T *smem)
{
    static_assert(NV <= kWave, "one wave folds the values");
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = threadIdx.x / kWave;
#pragma unroll
    for (int k = 0; k < NV; k++) {
        const T x = wave_reduce<OP>(v[k]);
        if (lane == 0) smem[k * NW + wave] = x;
    }
    __syncthreads();
    if (threadIdx.x < NV) {
        const int k = threadIdx.x;
        T acc = smem[k * NW];
#pragma unroll
        for (int w = 1; w < NW; w++) acc = red_op<OP>(acc, smem[k * NW + w]);
        smem[k * NW] = acc;            // only thread k touches row k
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < NV; k++) v[k] = smem[k * NW];
    }
    __syncthreads();
}

// Block-wide reduction of NV values, field-parallel: every thread stores its
// values to LDS, wave w folds fields w, w + NW, ... across the block (lane l
// takes thread l of every wave) and one DPP reduction per field finishes it.
// Field f is a maximum if bit f of MAXM is set, a minimum if bit f of MINM is
// set, a sum otherwise.  `lds` holds NV * NW * 64 values.  Result in thread 0.
template <class T, int NV, unsigned MAXM = 0u, unsigned MINM = 0u, int NW = kBlock / kWave>
__device__ inline void block_fold(T (&v)[NV], T *lds)
{
    constexpr int NT = NW * kWave;
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = threadIdx.x / kWave;
#pragma unroll
    for (int f = 0; f < NV; f++) lds[f * NT + threadIdx.x] = v[f];
    __syncthreads();
    for (int f = wave; f < NV; f += NW) {
        const bool mx = (MAXM >> f) & 1u, mn = (MINM >> f) & 1u;
        T x = lds[f * NT + lane];
#pragma unroll
        for (int w = 1; w < NW; w++) {
            const T y = lds[f * NT + w * kWave + lane];
            x = mx ? red_op<1>(x, y) : (mn ? red_op<2>(x, y) : red_op<0>(x, y));
        }
        x = mx ? wave_reduce<1>(x) : (mn ? wave_reduce<2>(x) : wave_reduce<0>(x));
        if (lane == 0) lds[f * NT] = x;          // only this wave touches field f
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int f = 0; f < NV; f++) v[f] = lds[f * NT];
    }
    __syncthreads();
}

// Triangle-A-C vertex weight (bulk_state_authority.py:624-634), same
// accumulation order as the reference's four slice additions.
__host__ __device__ inline double vertex_weight(int ny, int nx, double f6, int r, int c)
{
    double w = 0.0;
    if (r < ny - 1 && c < nx - 1) w += 2.0 * f6;
    if (r < ny - 1 && c > 0) w += f6;
    if (r > 0 && c > 0) w += 2.0 * f6;
    if (r > 0 && c < nx - 1) w += f6;
    return w;
}

__device__ inline double cell_weight(int mode, int ny, int nx, double area, double f6, int r, int c)
{
    return mode ? vertex_weight(ny, nx, f6, r, c) : area;
}

// ---- FP32-storage exchange amounts -----------------------------------------------
// b_eff is float32 and mobile holds whole quanta Q = 2^-24 m.  Moving d metres
// between them keeps H_free = b_eff + mobile bit-invariant per cell only if d
// is a whole number of quanta AND b_eff -/+ d is exactly a float32.  These give
// the largest such d (in quanta) not above the (already donor-limited) request;
// the ledger then reports exactly the change of both reservoirs.

// spacing of float32 numbers in the binade of v
__host__ __device__ inline double f32_spacing(float v)
{
    const unsigned e = (__builtin_bit_cast(unsigned, v) >> 23) & 0xffu;
    return e == 0 ? ldexp(1.0, -149) : ldexp(1.0, (int)e - 150);
}

__host__ __device__ inline bool f32_exact(double r) { return (double)(float)r == r; }

// Resting -> Mobile: depth = min(max(request, 0), available) metres, room =
// quanta the cell's int32 can still take.  A multiple of max(Q, spacing(b))
// below b stays on the float grid (it only moves towards smaller binades).
__host__ __device__ inline long long entrain_quanta_f32(float b, double depth, long long room)
{
    const double g = fmax(kQuantum, f32_spacing(b));
    double d = floor(depth / g) * g;
    const double cap = floor((double)room * kQuantum / g) * g;
    if (d > cap) d = cap;
    if (!(d > 0.0) || !f32_exact((double)b - d)) return 0;
    return (long long)(d * kInvQuantum);
}

// Mobile -> Resting: up to `quanta` = min(floor(max(request, 0) / Q), mobile
// quanta).  b_eff + d is kept exact when it can be: the sum itself, or the
// largest float32 below it when that still differs from b_eff by whole quanta
// (always so for b_eff >= 0.5 m, whose float32 spacing is >= Q; a request
// below that spacing moves nothing).  Otherwise b_eff lies below 0.5 m off
// the quantum grid and the deposit would carry it into a binade whose floats
// are all on the grid: then as many quanta as keep b_eff in its own binade
// move exactly, and when that is none the whole amount moves with b_eff
// rounded to the nearest float32 (a change of H_free below half that binade's
// spacing, once per such crossing, after which b_eff is on the grid).
// Returns the quanta moved; *b_new receives the new b_eff.
__host__ __device__ inline long long deposit_quanta_f32(float b, long long quanta, float *b_new)
{
    *b_new = b;
    if (quanta <= 0) return 0;
    const double d0 = (double)quanta * kQuantum;
    const double r = (double)b + d0;
    if (f32_exact(r)) {
        *b_new = (float)r;
        return quanta;
    }
    float rf = (float)r;
    if ((double)rf > r) rf = nextafterf(rf, -__builtin_inff());
    const double d = (double)rf - (double)b;
    if (d > 0.0 && floor(d * kInvQuantum) == d * kInvQuantum) {
        *b_new = rf;
        return (long long)(d * kInvQuantum);
    }
    // b_eff on the quantum grid: no whole number of quanta up to the request
    // lands on the float grid (the request is below b's spacing)
    if (floor((double)b * kInvQuantum) == (double)b * kInvQuantum) return 0;
    // b_eff off the quantum grid (below 0.5 m): stay inside its binade (its
    // floats differ from b_eff by multiples of its spacing, which divides Q)
    const double sb = f32_spacing(b);
    const unsigned e = (__builtin_bit_cast(unsigned, b) >> 23) & 0xffu;
    const double top = ldexp(1.0, (int)e - 126) - sb;            // largest float of b's binade
    const double k = floor((top - (double)b) * kInvQuantum);
    if (k >= 1.0 && f32_exact((double)b + k * kQuantum)) {
        *b_new = (float)((double)b + k * kQuantum);
        return (long long)k;
    }
    *b_new = (float)r;                                           // the rare rounded crossing
    return quanta;
}

// Load a small record written by an earlier launch, word by word with
// device-coherent loads (never served from a stale per-CU cache).
template <class T>
__device__ inline T load_coherent(const T *p)
{
    static_assert(sizeof(T) % 4 == 0, "record size must be a multiple of 4 bytes");
    unsigned w[sizeof(T) / 4];
    const unsigned *src = reinterpret_cast<const unsigned *>(p);
#pragma unroll
    for (int i = 0; i < (int)(sizeof(T) / 4); i++)
        w[i] = __hip_atomic_load(src + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    T v;
    __builtin_memcpy(&v, w, sizeof(T));
    return v;
}

// atomic max of a non-negative double via its ordered bit pattern
__device__ inline void atomic_max_nonneg(unsigned long long *slot, double v)
{
    if (!(v >= 0.0)) return;
    atomicMax(slot, (unsigned long long)__double_as_longlong(v));
}

// Workgroups are dispatched round-robin over the 8 XCDs.  The renumbered
// index gives the workgroups of one XCD one contiguous range of [0, G): tiles
// numbered row-major then share their halo rows and columns (and the cache
// lines a 62-column strip straddles) inside one XCD's L2.
__device__ inline int xcd_block()
{
    const int G = gridDim.x, b = blockIdx.x;
    const int x = b % 8, j = b / 8;
    return x * (G / 8) + min(x, G % 8) + j;
}
// A record written by an earlier launch (and not by this one), read through
// the constant address space as 32-bit words: scalar loads, counted apart
// from the vector loads a launch issues ahead of them.  A plain load may
// become a vector load, whose wait would then also cover every vector load
// issued before it.
template <class T>
__device__ inline T load_launch_record(const T *p)
{
    static_assert(sizeof(T) % 4 == 0, "the record is read as 32-bit words");
    constexpr int kWords = (int)(sizeof(T) / 4);
    typedef const __attribute__((address_space(4))) unsigned ConstWord;
    ConstWord *src = (ConstWord *)p;
    unsigned w[kWords];
#pragma unroll
    for (int i = 0; i < kWords; i++) w[i] = src[i];
    T v;
    __builtin_memcpy(&v, w, sizeof(T));
    return v;
}

// ---- spread maxima -----------------------------------------------------------------
// Device-scope atomics on one address serialise (~12 ns each on MI355X, so a
// launch of ~1000 workgroups pays ~12 us for one global maximum).  A spread
// maximum gives every 64th workgroup the same slot, 128 bytes apart, so the
// atomics of one launch proceed in parallel; the consumer (a later launch)
// folds the 64 slots with one wave.  Max is exact, so the order of arrival
// does not change the result.
constexpr int kSpread = 64;
constexpr int kSpreadStride = 16;                    // u64 words between slots
constexpr int kSpreadWords = kSpread * kSpreadStride;
// the step uses sets 0..4 (even frames) and 16..20 (odd frames) and, for its
// speculation, 8..9 (true maximum of the input, by frame parity) and 10..11
// (predicted maximum, by the parity of the frame that reads it), the
// envelope 5..7, the labelling 12..14, the relaxation 21..30
constexpr int kSpecTrueSet = 8;
constexpr int kSpecPredSet = 10;
constexpr int kEnvSpreadSet = 5;
constexpr int kLabelSpreadSet = 12;
constexpr int kStepOddSpreadSet = 16;
constexpr int kRelaxSpreadSet = 21;                  // five launch groups of two iterations
constexpr int kRelaxSpreadSets = 10;
constexpr size_t kSpreadScratchBytes = (size_t)(kRelaxSpreadSet + kRelaxSpreadSets) * kSpreadWords * 8;

__device__ inline void spread_max(unsigned long long *slots, double v)
{
    atomic_max_nonneg(&slots[(blockIdx.x % kSpread) * kSpreadStride], v);
}

// maximum over the slots (written by earlier launches); every thread of the
// block calls it and receives the value
__device__ inline double spread_read_max(const unsigned long long *slots, double *smem1)
{
    if (threadIdx.x < kWave) {
        double v = __longlong_as_double((long long)slots[threadIdx.x * kSpreadStride]);
        v = wave_max(v);
        if (threadIdx.x == 0) *smem1 = v;
    }
    __syncthreads();
    const double r = *smem1;
    __syncthreads();
    return r;
}

// zero the slots (threads 0..63 of one workgroup)
__device__ inline void spread_clear(unsigned long long *slots)
{
    if (threadIdx.x < kSpread) slots[threadIdx.x * kSpreadStride] = 0ull;
}

// Completion stamp, written by a thread after its writes to pinned host
// memory: the system-scope release makes those writes visible to the host
// before the stamp (stamp 0: no stamp).  Device memory the launch writes is
// ordered for later work by the stream, not by the stamp.
__device__ inline void host_stamp(unsigned *word, unsigned stamp)
{
    if (stamp) __hip_atomic_store(word, stamp, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
}

// ---- launch helpers --------------------------------------------------------------
inline int grid_for(int64_t n, int block = kBlock, int cap = 1 << 20)
{
    int64_t g = (n + block - 1) / block;
    if (g < 1) g = 1;
    if (g > cap) g = cap;
    return (int)g;
}

}  // namespace mv2
