// Internal definitions shared by the libmv2 translation units (gfx950, wave64).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>

#include "mv2.h"

namespace mv2 {

// Mobile thickness quantum of the FP32 storage mode: 2^-28 m (3.7e-9 m).
// Mass moves between cells as whole quanta, so FP32 runs conserve mass exactly
// and int32 storage covers 0 .. 8 m of mobile depth.
constexpr double kQuantum = 1.0 / 268435456.0;
constexpr double kInvQuantum = 268435456.0;

constexpr int kBlock = 256;      // threads per workgroup for streaming kernels
constexpr int kWave = 64;

struct Scratch {
    void *ptr = nullptr;
    size_t bytes = 0;
};

}  // namespace mv2

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
    mv2::Scratch host_result;             // pinned host staging
    mv2::Scratch label_aux;               // labelling partials
    mv2::Scratch nn_partials;             // weight-gradient partials of the surrogate
    mv2::Scratch nn_loss;                 // loss / projection partials
    hipEvent_t ev_start = nullptr, ev_stop = nullptr;   // device timing of step launches
};

namespace mv2 {

int set_error(mv2_context *ctx, int code, const char *fmt, ...);
int check_hip(mv2_context *ctx, hipError_t err, const char *what);
int ensure_scratch(mv2_context *ctx, Scratch &s, size_t bytes, bool host_pinned = false);

#define MV2_HIP(ctx, expr)                                     \
    do {                                                       \
        hipError_t _e = (expr);                                \
        if (_e != hipSuccess) return mv2::check_hip(ctx, _e, #expr); \
    } while (0)

// ---- reductions (wave64) ------------------------------------------------------
template <class T>
__device__ inline T wave_sum(T v)
{
    for (int off = kWave / 2; off > 0; off >>= 1) v += __shfl_down(v, off, kWave);
    return v;
}
template <class T>
__device__ inline T wave_max(T v)
{
    for (int off = kWave / 2; off > 0; off >>= 1) {
        T o = __shfl_down(v, off, kWave);
        v = o > v ? o : v;
    }
    return v;
}
template <class T>
__device__ inline T wave_min(T v)
{
    for (int off = kWave / 2; off > 0; off >>= 1) {
        T o = __shfl_down(v, off, kWave);
        v = o < v ? o : v;
    }
    return v;
}

// Block-wide reduction of NV values; result valid in thread 0.  `smem` must
// hold NV * (blockDim/64) elements.  Combines waves in fixed order, so the
// result is deterministic for a given launch shape.
template <class T, int NV, int OP>  // OP 0 sum, 1 max, 2 min
__device__ inline void block_reduce(T (&v)[NV], T *smem)
{
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = threadIdx.x / kWave;
    const int nwaves = blockDim.x / kWave;
    for (int k = 0; k < NV; k++) {
        T x = v[k];
        if (OP == 0) x = wave_sum(x);
        else if (OP == 1) x = wave_max(x);
        else x = wave_min(x);
        if (lane == 0) smem[k * nwaves + wave] = x;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 0; k < NV; k++) {
            T acc = smem[k * nwaves];
            for (int w = 1; w < nwaves; w++) {
                T o = smem[k * nwaves + w];
                if (OP == 0) acc += o;
                else if (OP == 1) acc = o > acc ? o : acc;
                else acc = o < acc ? o : acc;
            }
            v[k] = acc;
        }
    }
    __syncthreads();
}

// ---- grid barrier for cooperatively launched (co-resident) grids ---------------
// `counter` starts at 0 for the launch; barrier k completes when the counter
// reaches k * gridDim.x.  One agent-scope release fence before arriving and one
// acquire fence after the wait make every XCD's writes visible.  The wait
// itself polls with relaxed loads: an acquire load invalidates the XCD's L2 on
// every poll, and a thousand spinning workgroups would keep the L2 of all
// other workgroups cold.
__device__ inline void grid_barrier(unsigned *counter, unsigned &epoch)
{
    __syncthreads();
    epoch += 1;
    if (threadIdx.x == 0) {
        const unsigned target = epoch * gridDim.x;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        __hip_atomic_fetch_add(counter, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        while (__hip_atomic_load(counter, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
            __builtin_amdgcn_s_sleep(2);
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
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

// atomic max of a non-negative double via its ordered bit pattern
__device__ inline void atomic_max_nonneg(unsigned long long *slot, double v)
{
    if (!(v >= 0.0)) return;
    atomicMax(slot, (unsigned long long)__double_as_longlong(v));
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
