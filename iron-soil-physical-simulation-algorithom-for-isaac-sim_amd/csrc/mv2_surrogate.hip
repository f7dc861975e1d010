// Learned avalanche warm start: a 5-layer dilated 3x3 conv-net (3 -> 32 -> 32
// -> 32 -> 32 -> 1 channels, dilations 1, 2, 4, 8, 16: receptive field 63
// cells) predicting the change of b_eff under the angle-of-repose relaxation,
// plus the kernels that train it and the batched patch relaxation that makes
// its training pairs.
//
// * Layer 1 (3 input channels) runs in FP32 on v_mfma_f32_32x32x2_f32:
//   terrain heights of a few metres need FP32 to resolve millimetre
//   differences, and its taps on the b_eff channel are constrained to sum to
//   zero, so the network's output is invariant (in exact arithmetic) to
//   adding a constant to the terrain.
// * Layers 2-5 run as implicit GEMMs on the matrix cores:
//   v_mfma_f32_32x32x16_bf16, BF16 activations/weights, FP32 accumulation.
//   One wave computes 32 consecutive pixels of one row x 32 output channels;
//   K = 9 taps x input channels, 16 per MFMA.  Activations are NHWC so each
//   lane's 8 K-values of one pixel are one 16-byte load.
// * Backward-data reuses the same MFMA kernel with flipped, transposed
//   weights; the weight gradient is an LDS-tiled FP32 reduction with
//   per-workgroup partials combined in fixed order (deterministic).
#include "mv2_internal.hpp"

#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>

namespace mv2 {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

__device__ inline __bf16 to_bf16(float v) { return (__bf16)v; }
__device__ inline float from_bf16(__bf16 v) { return (float)v; }

// ---------------------------------------------------------------------------
// input features: f0 = b_eff, f1 = h_mobile, f2 = |grad(b_eff + h_mobile)|
// (centred differences inside, one-sided at the edges), FP32 [B][H][W][4]
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(kBlock) features_kernel(const float *b, const float *h, int B, int H, int W,
                                                          float dx, float dy, float *feat)
{
    const int64_t total = (int64_t)B * H * W;
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < total;
         p += (int64_t)gridDim.x * blockDim.x) {
        const int64_t n = p / ((int64_t)H * W);
        const int rem = (int)(p - n * H * W);
        const int y = rem / W, x = rem - y * W;
        auto Hf = [&](int yy, int xx) {
            const int64_t i = n * H * W + (int64_t)yy * W + xx;
            return b[i] + (h ? h[i] : 0.0f);
        };
        const int xl = x > 0 ? x - 1 : 0, xr = x + 1 < W ? x + 1 : W - 1;
        const int yd = y > 0 ? y - 1 : 0, yu = y + 1 < H ? y + 1 : H - 1;
        const float gx = (Hf(y, xr) - Hf(y, xl)) / ((float)(xr - xl) * dx);
        const float gy = (Hf(yu, x) - Hf(yd, x)) / ((float)(yu - yd) * dy);
        feat[p * 4 + 0] = b[p];
        feat[p * 4 + 1] = h ? h[p] : 0.0f;
        feat[p * 4 + 2] = __builtin_sqrtf(gx * gx + gy * gy);
        feat[p * 4 + 3] = 0.0f;
    }
}

// ---------------------------------------------------------------------------
// Tiling shared by the conv layers: one wave computes 32 output channels x 32
// consecutive pixels of one row per tile (D = W x X: the weights are the MFMA
// A operand, the pixels the B operand), so each lane's accumulators hold one
// pixel and the channels 8i + 4*half + j (i, j < 4) -- four runs of four
// contiguous channels, stored as four 8-byte writes.  A wave walks
// kConvRowsPerWave rows; a workgroup (4 waves) covers kConvRows rows x 32
// pixels of one field.
// Grid: (ceil(W/32), ceil(H/kConvRows), B).
// ---------------------------------------------------------------------------
constexpr int kConvRowsPerWave = 8;
constexpr int kConvRows = 4 * kConvRowsPerWave;

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));

// Output tile store: lane (n, half) holds channels 8i + 4*half + (0..3) of
// pixel n as o[i].  Written as-is, each 8-byte store instruction would touch
// 16 bytes of every 64-byte pixel; instead the wave stages the 32 x 32 tile in
// its LDS slice (pixel stride 40 channels: conflict-free 8-byte writes, 16-byte
// aligned reads) and writes it back as 16 contiguous pixels (1 KB) per store
// instruction.  dst: the row's first pixel of the segment; valid: pixels of
// the segment inside the field.  Every lane of the wave takes part.
constexpr int kStageStride = 40;
constexpr int kStageElems = 32 * kStageStride;

__device__ inline void store_tile_bf16(__bf16 *stage, const bf16x4 (&o)[4], int lane, __bf16 *dst, int valid)
{
    const int n = lane & 31, h = lane >> 5;
#pragma unroll
    for (int i = 0; i < 4; i++) *reinterpret_cast<bf16x4 *>(stage + n * kStageStride + 8 * i + 4 * h) = o[i];
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
#pragma unroll
    for (int k = 0; k < 2; k++) {
        const int q = (lane >> 2) + 16 * k, c = lane & 3;
        const bf16x8 v = *reinterpret_cast<const bf16x8 *>(stage + q * kStageStride + 8 * c);
        if (q < valid) *reinterpret_cast<bf16x8 *>(dst + q * 32 + 8 * c) = v;
    }
    // the slice is written again by the wave's next row
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

// layer 1: FP32 3x3 conv over 3 feature channels, bias, ReLU -> BF16 NHWC32,
// on v_mfma_f32_32x32x2_f32 (K = 9 taps x 3 channels, padded to 28).
// w1: [9][32][4] FP32 (channel 3 unused).  Zero padding of the b_eff channel
// would break the constant-shift invariance at the domain edge: outside the
// grid it takes the centre value, the other channels 0.
__global__ void __launch_bounds__(kBlock) conv1_mfma_kernel(const float *feat, const float *w1, const float *b1,
                                                            int B, int H, int W, int dil, __bf16 *y)
{
    __shared__ __bf16 stage[4][kStageElems];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int m = lane & 31, half = lane >> 5;
    const int x0 = blockIdx.x * 32, x = x0 + m;
    const int n = blockIdx.z;
    const int ybeg = blockIdx.y * kConvRows + wave * kConvRowsPerWave;
    if (ybeg >= H) return;
    const int yend = min(ybeg + kConvRowsPerWave, H);
    // A operand of step j: w[t][co = m][c] with t*3 + c = 2j + half
    float wa[14];
#pragma unroll
    for (int j = 0; j < 14; j++) {
        const int kk = 2 * j + half;
        wa[j] = kk < 27 ? w1[((kk / 3) * 32 + m) * 4 + kk % 3] : 0.0f;
    }
    float bias[16];
#pragma unroll
    for (int r = 0; r < 16; r++) bias[r] = b1[(r / 4) * 8 + half * 4 + (r % 4)];
    const f32x4 *f4 = reinterpret_cast<const f32x4 *>(feat);
    const int64_t plane = (int64_t)H * W;
    const int xc = x < W ? x : W - 1;                 // lanes past the edge read a valid pixel
    for (int yr = ybeg; yr < yend; yr++) {
        // every load is issued unconditionally (clamped address) and masked
        // afterwards, so the nine loads of a row are in flight together
        f32x4 f[9];
#pragma unroll
        for (int t = 0; t < 9; t++) {
            const int yy = yr + (t / 3 - 1) * dil, xx = xc + (t % 3 - 1) * dil;
            const bool in = yy >= 0 && yy < H && xx >= 0 && xx < W;
            f[t] = f4[n * plane + (int64_t)(in ? yy : yr) * W + (in ? xx : xc)];
        }
#pragma unroll
        for (int t = 0; t < 9; t++) {
            if (t == 4) continue;
            const int yy = yr + (t / 3 - 1) * dil, xx = xc + (t % 3 - 1) * dil;
            const bool in = yy >= 0 && yy < H && xx >= 0 && xx < W;
            if (!in) f[t] = f32x4{f[4][0], 0.0f, 0.0f, 0.0f};
        }
        f32x16 acc;
#pragma unroll
        for (int r = 0; r < 16; r++) acc[r] = 0.0f;
#pragma unroll
        for (int j = 0; j < 14; j++) {
            const int ke = 2 * j, ko = 2 * j + 1;
            const float ve = f[ke / 3][ke % 3];
            const float vo = ko < 27 ? f[ko / 3][ko % 3] : 0.0f;
            acc = __builtin_amdgcn_mfma_f32_32x32x2f32(wa[j], half ? vo : ve, acc, 0, 0, 0);
        }
        bf16x4 o[4];
#pragma unroll
        for (int i = 0; i < 4; i++)
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const float v = acc[4 * i + j] + bias[4 * i + j];
                o[i][j] = to_bf16(v > 0.0f ? v : 0.0f);
            }
        store_tile_bf16(stage[wave], o, lane, y + (n * plane + (int64_t)yr * W + x0) * 32, W - x0);
    }
}

// ---------------------------------------------------------------------------
// MFMA implicit-GEMM 3x3 conv.  x: BF16 NHWC with CX = 16*KS channels;
// w: BF16 [9][32][CX] (output channel major, input channel innermost).
// mode 0: BF16 NHWC32 output (+bias, optional ReLU, optional multiply by
//         (mask > 0) where mask is BF16 NHWC32)
// mode 1: FP32 single-channel output = channel 0 (+bias[0])
// v_mfma_f32_32x32x16_bf16, K = 9 taps x CX, 16 per step.
//
// A wave walks a chain of kChain output rows spaced by the dilation D (rows
// base, base + D, ...): output row y reads input rows y - D, y, y + D, so
// consecutive rows of a chain share two of their three input rows, kept in a
// rotating three-row register window; each step loads one new input row (the
// 3 horizontal taps x KS vectors), issued after the MFMAs of the window row it
// replaces, a step and a half ahead of its use.  Rows of a field split into
// blocks of kChain * D rows, each holding D chains.  The weights (padded
// stride per (tap, channel) row: conflict-free operand reads) and the bias
// sit in LDS.  Grid: (ceil(W/32), ceil(chains/4), B).
// ---------------------------------------------------------------------------
constexpr int kChain = 8;

struct ConvArgs {
    const __bf16 *x;
    const __bf16 *w;
    const float *bias;
    const __bf16 *mask;
    void *y;
    int B, H, W, dil;
    int relu, mode;
};

inline int conv_chains(int H, int dil) { return ((H + kChain * dil - 1) / (kChain * dil)) * dil; }

template <int KS, bool MASK>
__global__ void __launch_bounds__(kBlock) conv_mfma_kernel(ConvArgs a)
{
    constexpr int CX = 16 * KS;
    constexpr int VPR = CX / 8;                    // bf16x8 vectors per (tap, co) row
    constexpr int WS = VPR + 1;                    // padded LDS stride of a (tap, co) row
    __shared__ bf16x8 wl[9 * 32 * WS];
    __shared__ float bl[32];
    __shared__ __bf16 stage[4][kStageElems];
    const bf16x8 *wg = reinterpret_cast<const bf16x8 *>(a.w);
    for (int i = threadIdx.x; i < 9 * 32 * VPR; i += blockDim.x) wl[(i / VPR) * WS + i % VPR] = wg[i];
    if (threadIdx.x < 32) bl[threadIdx.x] = a.bias ? a.bias[threadIdx.x] : 0.0f;
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int m = lane & 31, half = lane >> 5;
    const int x0 = blockIdx.x * 32, x = x0 + m;
    const int n = blockIdx.z;
    const int D = a.dil;
    const int g = blockIdx.y * 4 + wave;
    const int base = (g / D) * (kChain * D) + g % D;
    if (base >= a.H) return;
    const int64_t plane = (int64_t)a.H * a.W;
    const int xc = x < a.W ? x : a.W - 1;             // lanes past the edge read a valid pixel
    bool col_in[3];
    int col[3];
#pragma unroll
    for (int tx = 0; tx < 3; tx++) {
        const int xx = xc + (tx - 1) * D;
        col_in[tx] = xx >= 0 && xx < a.W;
        col[tx] = col_in[tx] ? xx : xc;
    }
    const __bf16 *img = a.x + n * plane * CX;
    const bf16x8 zero = {};
    // input row yy into a window slot: loads at clamped addresses; the zero
    // padding is applied where the slot is used (row_ok), so nothing waits
    // for a load before its MFMA needs it
    bf16x8 win[3][3][KS];
    bool row_ok[3];
    auto load_row = [&](int yy, int slot) {
        row_ok[slot] = yy >= 0 && yy < a.H;
        const int yc = row_ok[slot] ? yy : base;
#pragma unroll
        for (int tx = 0; tx < 3; tx++) {
            const bf16x8 *xp = reinterpret_cast<const bf16x8 *>(img + ((int64_t)yc * a.W + col[tx]) * CX);
#pragma unroll
            for (int ks = 0; ks < KS; ks++) win[slot][tx][ks] = xp[ks * 2 + half];
        }
    };
    load_row(base - D, 0);
    load_row(base, 1);
    load_row(base + D, 2);
    __builtin_amdgcn_sched_barrier(0);             // keep the window loads ahead of the first prefetch
#pragma unroll
    for (int j = 0; j < kChain; j++) {
        const int y = base + j * D;
        if (y >= a.H) break;
        f32x16 acc;
#pragma unroll
        for (int r = 0; r < 16; r++) acc[r] = 0.0f;
#pragma unroll
        for (int ty = 0; ty < 3; ty++) {
            const int slot = (j + ty) % 3;
#pragma unroll
            for (int tx = 0; tx < 3; tx++)
#pragma unroll
                for (int ks = 0; ks < KS; ks++)
                    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(
                        wl[((ty * 3 + tx) * 32 + m) * WS + ks * 2 + half],
                        row_ok[slot] && col_in[tx] ? win[slot][tx][ks] : zero, acc, 0, 0, 0);
            if (ty == 0 && j + 1 < kChain && y + D < a.H) {
                // the slot of input row y - D is free: next step's row y + 2D.
                // Loads may complete out of order, so a use of any pending
                // load waits for all of them: the previous prefetch is waited
                // for here, before the new one is issued, and the rest of the
                // step uses complete registers only
                __builtin_amdgcn_s_waitcnt(0x0f70);            // vmcnt(0)
                __builtin_amdgcn_sched_barrier(0);
                load_row(y + 2 * D, j % 3);
                __builtin_amdgcn_sched_barrier(0);
            }
        }
        // weight reads two MFMAs ahead of their use rather than all hoisted
        // (registers): lead of 2 KS reads, then one read per MFMA
        __builtin_amdgcn_sched_group_barrier(0x100, 2 * KS, 0);
#pragma unroll
        for (int i = 0; i < 7 * KS; i++) {
            __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);
            __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);
        }
        __builtin_amdgcn_sched_group_barrier(0x008, 2 * KS, 0);
        if (a.mode == 1) {
            // channel 0 is accumulator 0 of the lanes with half == 0
            if (half == 0 && x < a.W) reinterpret_cast<float *>(a.y)[n * plane + (int64_t)y * a.W + x] = acc[0] + bl[0];
            continue;
        }
        // the backward's mask loads are a separate instantiation: their wait
        // would also wait for the prefetched row
        const __bf16 *mp = a.mask + (n * plane + (int64_t)y * a.W + xc) * 32 + half * 4;
        bf16x4 o[4];
#pragma unroll
        for (int i = 0; i < 4; i++) {
            bf16x4 mk = {};
            if (MASK) mk = *reinterpret_cast<const bf16x4 *>(mp + 8 * i);
#pragma unroll
            for (int jj = 0; jj < 4; jj++) {
                float v = acc[4 * i + jj] + bl[8 * i + 4 * half + jj];
                if (a.relu) v = v > 0.0f ? v : 0.0f;
                if (MASK && !(from_bf16(mk[jj]) > 0.0f)) v = 0.0f;
                o[i][jj] = to_bf16(v);
            }
        }
        store_tile_bf16(stage[wave], o, lane,
                        reinterpret_cast<__bf16 *>(a.y) + (n * plane + (int64_t)y * a.W + x0) * 32, a.W - x0);
    }
}

// pack FP32 master weights of an MFMA layer into BF16 forward / backward forms.
// master w: [9][cout][cin] FP32 with cout, cin <= 32.
//   fwd  [9][32][CXF]: fwd[t][co][ci] = w[t][co][ci]
//   bwd  [9][32][CXB]: bwd[t][ci][co] = w[8 - t][co][ci]   (for dX = conv(dZ, bwd))
__global__ void pack_weights_kernel(const float *w, int cout, int cin, int cxf, int cxb, __bf16 *fwd,
                                    __bf16 *bwd)
{
    const int total = 9 * 32 * 32;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < total; i += gridDim.x * blockDim.x) {
        const int t = i / 1024, r = (i / 32) % 32, c = i % 32;
        if (fwd && c < cxf) fwd[(t * 32 + r) * cxf + c] = to_bf16(r < cout && c < cin ? w[(t * cout + r) * cin + c] : 0.0f);
        if (bwd && c < cxb)
            bwd[(t * 32 + r) * cxb + c] = to_bf16(c < cout && r < cin ? w[((8 - t) * cout + c) * cin + r] : 0.0f);
    }
}

// single-channel FP32 gradient -> BF16 NHWC16 (channel 0 = value, others 0)
__global__ void expand_grad_kernel(const float *g, int64_t npix, __bf16 *out16)
{
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < npix; p += (int64_t)gridDim.x * blockDim.x) {
        bf16x8 v0 = {}, v1 = {};
        v0[0] = to_bf16(g[p]);
        reinterpret_cast<bf16x8 *>(out16)[p * 2] = v0;
        reinterpret_cast<bf16x8 *>(out16)[p * 2 + 1] = v1;
    }
}

// ---------------------------------------------------------------------------
// weight gradient: dW[t][co][ci] = sum_p dZ[p][co] * X[p + off(t)][ci],
// db[co] = sum_p dZ[p][co].  dZ: BF16 NHWC32 (zmode 0) or FP32 channel-0
// (zmode 1, cout = 1); X: BF16 NHWC32 (xmode 0) or FP32 features stride 4
// (xmode 1, 3 channels and a zero pad slot: cin <= 4).  partials:
// [nchunks][9][32][32] (+ [nchunks][32] bias, tap 0 only); only the pairs
// co < cout, ci < cin are written and read.
//
// weight_grad_small_kernel: the first and last layers, whose cout x cin is
// 1 x 32 (zmode 1: FP32 dZ, BF16 NHWC32 X) or 32 x 4 (xmode 1: BF16 NHWC32 dZ,
// FP32 features: 3 channels and a zero pad slot).  Each thread takes pixels
// p0 + t, p0 + t + 256, ... of the workgroup's chunk, loads the pixel's
// 32-channel row as 16-byte vectors and accumulates its 32 (or 3 x 32 + 32
// bias) sums in registers; DPP wave sums and a fixed-order sum over the 4
// waves give the chunk's partials.
template <int XMODE>
__global__ void __launch_bounds__(kBlock) weight_grad_small_kernel(const void *dz, const void *xin, int B, int H,
                                                                   int W, int dil, int64_t chunk, float *partials,
                                                                   float *bias_partials)
{
    constexpr int kWaves = kBlock / kWave;
    constexpr int NV = XMODE ? 32 * 4 : 32 + 1;   // sums per thread: [co][ci 0..2, bias] or [ci], bias
    __shared__ float red[kWaves][NV];
    const int tap = blockIdx.y;
    const int dyo = (tap / 3 - 1) * dil, dxo = (tap % 3 - 1) * dil;
    const int64_t plane = (int64_t)H * W;
    const int64_t total = (int64_t)B * plane;
    const int64_t p0 = (int64_t)blockIdx.x * chunk;
    const int64_t p1 = p0 + chunk < total ? p0 + chunk : total;
    float acc[NV];
#pragma unroll
    for (int v = 0; v < NV; v++) acc[v] = 0.0f;
    for (int64_t p = p0 + threadIdx.x; p < p1; p += blockDim.x) {
        const int64_t n = p / plane;
        const int rem = (int)(p - n * plane);
        const int yy = rem / W + dyo, xx = rem % W + dxo;
        const bool inside = yy >= 0 && yy < H && xx >= 0 && xx < W;
        const int64_t q = n * plane + (int64_t)yy * W + xx;
        if constexpr (XMODE) {
            bf16x8 zr[4];
            const bf16x8 *zp = reinterpret_cast<const bf16x8 *>(reinterpret_cast<const __bf16 *>(dz) + p * 32);
#pragma unroll
            for (int c = 0; c < 4; c++) zr[c] = zp[c];
            const float *feat = reinterpret_cast<const float *>(xin);
            float f[3] = {0.0f, 0.0f, 0.0f};
            if (inside) {
                f[0] = feat[q * 4];
                f[1] = feat[q * 4 + 1];
                f[2] = feat[q * 4 + 2];
            } else {
                f[0] = feat[p * 4];     // replicate-padding of the b_eff channel (conv1_mfma_kernel)
            }
#pragma unroll
            for (int co = 0; co < 32; co++) {
                const float z = from_bf16(zr[co / 8][co % 8]);
                acc[co * 4 + 0] += z * f[0];
                acc[co * 4 + 1] += z * f[1];
                acc[co * 4 + 2] += z * f[2];
                acc[co * 4 + 3] += z;
            }
        } else {
            const float z = reinterpret_cast<const float *>(dz)[p];
            acc[32] += z;
            if (inside) {
                const bf16x8 *xp = reinterpret_cast<const bf16x8 *>(reinterpret_cast<const __bf16 *>(xin) + q * 32);
#pragma unroll
                for (int c = 0; c < 4; c++) {
                    const bf16x8 xr = xp[c];
#pragma unroll
                    for (int i = 0; i < 8; i++) acc[8 * c + i] += z * from_bf16(xr[i]);
                }
            }
        }
    }
    const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
#pragma unroll
    for (int v = 0; v < NV; v++) {
        const float t = wave_reduce<0>(acc[v]);
        if (lane == 0) red[wave][v] = t;
    }
    __syncthreads();
    float *dst = partials + ((int64_t)blockIdx.x * 9 + tap) * 1024;
    for (int v = threadIdx.x; v < NV; v += blockDim.x) {
        float t = red[0][v];
        for (int w = 1; w < kWaves; w++) t += red[w][v];
        if constexpr (XMODE) {
            const int co = v / 4, k = v % 4;
            if (k < 3) dst[co * 32 + k] = t;
            else {
                dst[co * 32 + 3] = 0.0f;          // the features' zero pad slot
                if (tap == 0) bias_partials[(int64_t)blockIdx.x * 32 + co] = t;
            }
        } else {
            if (v < 32) dst[v] = t;
            else if (tap == 0) bias_partials[(int64_t)blockIdx.x * 32] = t;
        }
    }
}

// The 32 -> 32 channel layers (dZ and X both BF16 NHWC32) on
// v_mfma_f32_32x32x16_bf16: per tap, dW_t (32 co x 32 ci) = A x B with
// A = dZ^T (32 co x 16 pixels) and B = X_t (16 pixels x 32 ci), the pixels of
// the chunk as the contraction.  Each wave stages 64-pixel slabs of dZ and of
// X shifted by the tap in its own LDS rows (pixel-major, as stored: 16-byte
// coalesced loads), then gathers its operands channel-wise: lane (m, half)
// takes the 8 pixels half*8 .. half*8+7 of channel m.  D row i = co, column
// j = ci: accumulator r of lane (j, half) is co = 8 (r / 4) + 4 half + r % 4.
// The 4 waves' sums (and the bias sums of tap 0) combine in fixed order into
// the same per-chunk partials as weight_grad_kernel.
constexpr int kGradMfmaSlab = 64;                 // pixels per wave and slab
constexpr int kGradRow = 40;                      // LDS row of a pixel: 32 channels + pad (bf16)

__global__ void __launch_bounds__(kBlock) weight_grad_mfma_kernel(const __bf16 *dz, const __bf16 *xin, int B, int H,
                                                                  int W, int dil, int64_t chunk, float *partials,
                                                                  float *bias_partials)
{
    constexpr int kWaves = kBlock / kWave;
    __shared__ __bf16 zs[kWaves][kGradMfmaSlab * kGradRow];
    __shared__ __bf16 xs[kWaves][kGradMfmaSlab * kGradRow];
    __shared__ float red[kWaves][kWave * 17];
    const int tap = blockIdx.y;
    const int dyo = (tap / 3 - 1) * dil, dxo = (tap % 3 - 1) * dil;
    const int64_t plane = (int64_t)H * W;
    const int64_t total = (int64_t)B * plane;
    const int64_t p0 = (int64_t)blockIdx.x * chunk;
    const int64_t p1 = p0 + chunk < total ? p0 + chunk : total;
    const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
    const int m = lane & 31, half = lane >> 5;
    __bf16 *zw = zs[wave], *xw = xs[wave];
    f32x16 acc;
#pragma unroll
    for (int r = 0; r < 16; r++) acc[r] = 0.0f;
    float bacc = 0.0f;
    const bf16x8 zero = {};
    // lane l brings pixel s + l of a slab (dZ) and its tap-shifted input pixel
    // (zeros outside the image or past the chunk)
    bf16x8 zv[4], xv[4];
    auto fetch = [&](int64_t s) {
        const int64_t p = s + lane;
#pragma unroll
        for (int c = 0; c < 4; c++) zv[c] = xv[c] = zero;
        if (p < p1) {
            const bf16x8 *zp = reinterpret_cast<const bf16x8 *>(dz + p * 32);
#pragma unroll
            for (int c = 0; c < 4; c++) zv[c] = zp[c];
            const int64_t n = p / plane;
            const int rem = (int)(p - n * plane);
            const int yy = rem / W + dyo, xx = rem % W + dxo;
            if (yy >= 0 && yy < H && xx >= 0 && xx < W) {
                const bf16x8 *xp = reinterpret_cast<const bf16x8 *>(xin + (n * plane + (int64_t)yy * W + xx) * 32);
#pragma unroll
                for (int c = 0; c < 4; c++) xv[c] = xp[c];
            }
        }
    };
    const int64_t step = (int64_t)kWaves * kGradMfmaSlab;
    int64_t s = p0 + (int64_t)wave * kGradMfmaSlab;
    if (s < p1) fetch(s);
    for (; s < p1; s += step) {
#pragma unroll
        for (int c = 0; c < 4; c++) {
            *reinterpret_cast<bf16x8 *>(zw + lane * kGradRow + 8 * c) = zv[c];
            *reinterpret_cast<bf16x8 *>(xw + lane * kGradRow + 8 * c) = xv[c];
        }
        // the next slab's loads fly while this one is multiplied
        if (s + step < p1) fetch(s + step);
        // the rows are this wave's own: its LDS operations complete in order
        __builtin_amdgcn_wave_barrier();
#pragma unroll
        for (int k = 0; k < kGradMfmaSlab / 16; k++) {
            bf16x8 a, b;
#pragma unroll
            for (int i = 0; i < 8; i++) {
                const int row = k * 16 + half * 8 + i;
                a[i] = zw[row * kGradRow + m];
                b[i] = xw[row * kGradRow + m];
            }
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
            if (tap == 0) {
#pragma unroll
                for (int i = 0; i < 8; i++) bacc += from_bf16(a[i]);
            }
        }
        __builtin_amdgcn_wave_barrier();
    }
    // fixed-order combination of the waves: 16 accumulators + the bias sum per lane
    float *rw = red[wave] + lane * 17;
#pragma unroll
    for (int r = 0; r < 16; r++) rw[r] = acc[r];
    rw[16] = bacc;
    __syncthreads();
    if (wave == 0) {
        float out[17];
#pragma unroll
        for (int r = 0; r < 17; r++) out[r] = red[0][lane * 17 + r];
        for (int w = 1; w < kWaves; w++)
#pragma unroll
            for (int r = 0; r < 17; r++) out[r] += red[w][lane * 17 + r];
        float *dst = partials + ((int64_t)blockIdx.x * 9 + tap) * 1024;
#pragma unroll
        for (int r = 0; r < 16; r++) dst[(8 * (r / 4) + 4 * half + r % 4) * 32 + m] = out[r];
        if (tap == 0) {
            // lanes m and m + 32 hold the two pixel halves of channel m
            const float other = __shfl_down(out[16], 32, kWave);
            if (half == 0) bias_partials[(int64_t)blockIdx.x * 32 + m] = out[16] + other;
        }
    }
}

// fixed-order reduction of the partials into dW [9][cout][cin] and db [cout]
__global__ void __launch_bounds__(kBlock) reduce_grad_kernel(const float *partials, const float *bias_partials,
                                                             int nchunks, int cout, int cin, float *dw, float *db)
{
    const int total = 9 * 32 * 32 + 32;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < total; i += gridDim.x * blockDim.x) {
        if (i < 9 * 1024) {
            const int t = i / 1024, co = (i / 32) % 32, ci = i % 32;
            if (co >= cout || ci >= cin) continue;
            float s = 0.0f;
            for (int c = 0; c < nchunks; c++) s += partials[((int64_t)c * 9 + t) * 1024 + co * 32 + ci];
            dw[(t * cout + co) * cin + ci] = s;
        } else {
            const int co = i - 9 * 1024;
            if (co >= cout) continue;
            float s = 0.0f;
            for (int c = 0; c < nchunks; c++) s += bias_partials[(int64_t)c * 32 + co];
            db[co] = s;
        }
    }
}

// MSE loss on the predicted b_eff change; grad = 2 (pred - target) / N,
// per-workgroup partial loss sums.  `weight` (optional) scales per pixel.
__global__ void __launch_bounds__(kBlock) mse_kernel(const float *pred, const float *target, int64_t n,
                                                     float *grad, double *partials)
{
    __shared__ double smem[4];
    double s[1] = {0.0};
    const float scale = 2.0f / (float)n;
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += (int64_t)gridDim.x * blockDim.x) {
        const float d = pred[p] - target[p];
        grad[p] = scale * d;
        s[0] += (double)d * (double)d;
    }
    block_reduce<double, 1, 0>(s, smem);
    if (threadIdx.x == 0) partials[blockIdx.x] = s[0];
}

// Adam step on FP32 master parameters; `zero_sum_stride` > 0 re-projects the
// b_eff-channel taps of layer 1 onto zero sum (per output channel).
// Adam (Kingma & Ba 2015, Algorithm 1); omb1 / omb2 = 1 - beta1 / 1 - beta2
// formed in double and rounded once
__global__ void adam_kernel(float *param, const float *grad, float *m, float *v, int n, float lr, float beta1,
                            float beta2, float omb1, float omb2, float eps, float bc1, float bc2)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const float g = grad[i];
        const float mi = beta1 * m[i] + omb1 * g;
        const float vi = beta2 * v[i] + omb2 * g * g;
        m[i] = mi;
        v[i] = vi;
        param[i] -= lr * (mi / bc1) / (__builtin_sqrtf(vi / bc2) + eps);
    }
}

// layer-1 weights [9][32][4]: b_eff channel (0) taps sum to zero per output
// channel; channel 3 stays zero
__global__ void project_w1_kernel(float *w1)
{
    const int co = threadIdx.x;
    if (co >= 32) return;
    float s = 0.0f;
    for (int t = 0; t < 9; t++) s += w1[(t * 32 + co) * 4];
    s /= 9.0f;
    for (int t = 0; t < 9; t++) {
        w1[(t * 32 + co) * 4] -= s;
        w1[(t * 32 + co) * 4 + 3] = 0.0f;
    }
}

// layer-1 input gradient is not needed; the layer-1 pre-activation gradient
// is dX1 * (Y1 > 0), produced by conv_mfma_kernel's mask path.

// ---------------------------------------------------------------------------
// batched patch relaxation: one workgroup relaxes one independent H x W
// patch (FP32, closed boundary, legacy direction/parity schedule) entirely
// in LDS until max excess <= tol or max_iterations.
// ---------------------------------------------------------------------------
constexpr int kPatchMax = 64 * 64;

__global__ void __launch_bounds__(kBlock) relax_patches_kernel(float *h, int npatch, int H, int W, float c0, float c1,
                                                               float c2, float c3, int max_iterations, float tol,
                                                               int *iterations)
{
    __shared__ float t[kPatchMax];
    __shared__ float red[8];
    __shared__ int done;
    const int patch = blockIdx.x;
    if (patch >= npatch) return;
    float *hp = h + (int64_t)patch * H * W;
    for (int i = threadIdx.x; i < H * W; i += blockDim.x) t[i] = hp[i];
    __syncthreads();
    const float crit[4] = {c0, c1, c2, c3};
    int it = 0;
    for (it = 1; it <= max_iterations; it++) {
        float mx = 0.0f;
        for (int d = 0; d < 4; d++) {
            for (int phase = 0; phase < 2; phase++) {
                // view index (r, c) over the edge set of direction d
                const int rows = (d == 0) ? H : H - 1;
                const int cols = (d == 1) ? W : W - 1;
                for (int e = threadIdx.x; e < rows * cols; e += blockDim.x) {
                    const int r = e / cols, c = e - r * cols;
                    int ia, ib;
                    if (d == 0) { ia = r * W + c; ib = ia + 1; }
                    else if (d == 1) { ia = r * W + c; ib = ia + W; }
                    else if (d == 2) { ia = r * W + c; ib = ia + W + 1; }
                    else { ia = (r + 1) * W + c; ib = r * W + c + 1; }
                    float ex = fabsf(t[ia] - t[ib]) - crit[d];
                    ex = ex > 0.0f ? ex : 0.0f;
                    mx = mx >= ex ? mx : ex;
                }
                __syncthreads();
                // red-black transfer: enumerate only the anchors of the active
                // parity (the coordinate is c for d = 0, 2, 3 and r for d = 1);
                // their edges are disjoint, so a thread loads the operands of
                // four edges before it stores any of them
                const int arows = (d == 1) ? (rows - phase + 1) / 2 : rows;
                const int acols = (d == 1) ? cols : (cols - phase + 1) / 2;
                const int nact = arows * acols;
                for (int e0 = threadIdx.x; e0 < nact; e0 += 4 * blockDim.x) {
                    int ia[4], ib[4];
                    float va[4], vb[4];
#pragma unroll
                    for (int u = 0; u < 4; u++) {
                        const int e = e0 + u * blockDim.x;
                        ia[u] = -1;
                        if (e >= nact) continue;
                        const int ar = e / acols, ac = e - ar * acols;
                        const int r = d == 1 ? 2 * ar + phase : ar;
                        const int c = d == 1 ? ac : 2 * ac + phase;
                        if (d == 0) { ia[u] = r * W + c; ib[u] = ia[u] + 1; }
                        else if (d == 1) { ia[u] = r * W + c; ib[u] = ia[u] + W; }
                        else if (d == 2) { ia[u] = r * W + c; ib[u] = ia[u] + W + 1; }
                        else { ia[u] = (r + 1) * W + c; ib[u] = r * W + c + 1; }
                        va[u] = t[ia[u]];
                        vb[u] = t[ib[u]];
                    }
#pragma unroll
                    for (int u = 0; u < 4; u++) {
                        if (ia[u] < 0) continue;
                        const float diff = va[u] - vb[u];
                        float ex = fabsf(diff) - crit[d];
                        ex = ex > 0.0f ? ex : 0.0f;
                        const float sgn = diff > 0.0f ? 1.0f : (diff < 0.0f ? -1.0f : 0.0f);
                        const float tr = (0.5f * ex) * sgn;
                        t[ia[u]] = va[u] - tr;
                        t[ib[u]] = vb[u] + tr;
                    }
                }
                __syncthreads();
            }
        }
        // wave max then workgroup max
        const float wm = wave_max(mx);
        if ((threadIdx.x & (kWave - 1)) == 0) red[threadIdx.x / kWave] = wm;
        __syncthreads();
        if (threadIdx.x == 0) {
            float g = red[0];
            for (int w = 1; w < (int)(blockDim.x / kWave); w++) g = red[w] > g ? red[w] : g;
            done = g <= tol;
        }
        __syncthreads();
        if (done) break;
    }
    for (int i = threadIdx.x; i < H * W; i += blockDim.x) hp[i] = t[i];
    if (threadIdx.x == 0 && iterations) iterations[patch] = it > max_iterations ? max_iterations : it;
}

// Slope-excess penalty on the warm-started surface z = pre + pred:
//   L_ex = weight / N * sum over edges e of max(|z_a - z_b| - c_e, 0)^2
// over the four legacy edge directions (a -> b = a + off_d: d0 (0,+1),
// d1 (+1,0), d2 (+1,+1), d3 (-1,+1) in (row, col); c_e = crit[d]).  Each
// thread owns one cell and gathers the gradient of the (up to) eight edges it
// belongs to, so the in-place update of grad needs no atomics; the loss adds
// each edge once, on its anchor cell a.
__global__ void __launch_bounds__(kBlock) excess_loss_kernel(const float *pred, const float *pre, int B, int H,
                                                             int W, float c0, float c1, float c2, float c3,
                                                             float weight, float *grad, double *partials)
{
    __shared__ double smem[4];
    const float crit[4] = {c0, c1, c2, c3};
    const int dr[4] = {0, 1, 1, -1}, dc[4] = {1, 0, 1, 1};
    const int64_t n = (int64_t)B * H * W;
    const float scale = weight / (float)n;
    double s[1] = {0.0};
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += (int64_t)gridDim.x * blockDim.x) {
        const int64_t img = p / ((int64_t)H * W);
        const int rem = (int)(p - img * H * W);
        const int r = rem / W, c = rem - r * W;
        const float zp = pre[p] + pred[p];
        float g = 0.0f;
#pragma unroll
        for (int d = 0; d < 4; d++) {
            // cell as a: partner b = p + off
            int rr = r + dr[d], cc = c + dc[d];
            if (rr >= 0 && rr < H && cc >= 0 && cc < W) {
                const int64_t q = p + (int64_t)dr[d] * W + dc[d];
                const float diff = zp - (pre[q] + pred[q]);
                const float ex = fabsf(diff) - crit[d];
                if (ex > 0.0f) {
                    g += 2.0f * ex * (diff > 0.0f ? 1.0f : -1.0f);
                    s[0] += (double)ex * (double)ex;
                }
            }
            // cell as b: partner a = p - off
            rr = r - dr[d]; cc = c - dc[d];
            if (rr >= 0 && rr < H && cc >= 0 && cc < W) {
                const int64_t q = p - (int64_t)dr[d] * W - dc[d];
                const float diff = (pre[q] + pred[q]) - zp;
                const float ex = fabsf(diff) - crit[d];
                if (ex > 0.0f) g -= 2.0f * ex * (diff > 0.0f ? 1.0f : -1.0f);
            }
        }
        grad[p] += scale * g;
    }
    block_reduce<double, 1, 0>(s, smem);
    if (threadIdx.x == 0) partials[blockIdx.x] = s[0];
}

// Mass projection of a change delta of b (delta = given array, or z - b for a
// candidate surface z): remove a shift s from the change over the cells it
// moves (|delta| > eps) and keep heights >= 0, with s such that the moved cells
// keep their total: first s = the mean change, then Newton steps on the mass
// the moved cells gain, s += gain(s) / (moved cells above zero).  gain is
// convex and decreasing in s, so the steps approach its zero from one side;
// without cells held at zero the first s already zeroes it (to rounding).
//
// Masked form (MASKED, from a surface z): the moved cells are a given set M
// (the cells the first projection moved) and every other cell takes z as it
// is; the shift over M balances the mass of the whole field, so the change
// z - b outside M is paid for by M.  The settle rounds use it: the envelope of
// the committed surface cuts the steps the first shift left at the rim of M,
// and the masked projection returns that mass to M, which leaves a step of a
// much smaller shift.
// Field-major partials: sum at [g], count at [gridDim.x + g].
template <bool FROM_SURFACE>
__device__ inline float change_at(const float *b, const float *src, int64_t p)
{
    return FROM_SURFACE ? src[p] - b[p] : src[p];
}

struct ProjectArgs {
    const float *src;
    int64_t n;
    float eps;
    const uint8_t *mask_in;     // MASKED: the moved cells
    uint8_t *mask_out;          // optional (unmasked): records the moved cells
};

template <bool FROM_SURFACE, bool MASKED>
__device__ inline bool moved_at(const ProjectArgs &a, int64_t base, int64_t p, float d)
{
    return MASKED ? a.mask_in[base + p] != 0 : fabsf(d) > a.eps;
}

// Segments: the field is `gridDim.y` independent segments of n cells each
// (segment s at offset s * n), each with its own shift.
template <bool FROM_SURFACE, bool MASKED>
__global__ void __launch_bounds__(kBlock) delta_stats_kernel(const float *b, ProjectArgs a, double *partials)
{
    __shared__ double smem[8];
    const int64_t base = (int64_t)blockIdx.y * a.n;
    b += base;
    const float *src = a.src + base;
    partials += (int64_t)blockIdx.y * 2 * gridDim.x;
    double s[2] = {0.0, 0.0};
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < a.n; p += (int64_t)gridDim.x * blockDim.x) {
        const float d = change_at<FROM_SURFACE>(b, src, p);
        const bool mv = moved_at<FROM_SURFACE, MASKED>(a, base, p, d);
        if (MASKED) {
            s[0] += (double)src[p] - (double)b[p];
            if (mv) s[1] += 1.0;
        } else if (mv) {
            s[0] += (double)d;
            s[1] += 1.0;
        }
    }
    block_reduce<double, 2, 0>(s, smem);
    if (threadIdx.x == 0) {
        partials[blockIdx.x] = s[0];
        partials[gridDim.x + blockIdx.x] = s[1];
    }
}

constexpr int kProjectParts = 1024;           // workgroups per segment of the projection passes at most
constexpr int kProjectPer = kProjectParts / kBlock;

// The shift of the current pass, folded by every workgroup from the previous
// pass's partials in one fixed order (thread t sums partials t, t + kBlock,
// ... from zero, then the block tree), so all workgroups of a segment hold the
// same bits.  correction == 0: the first shift, sum of the change / moved
// cells; otherwise the Newton step on the gained mass from the previous shift
// (kept in shifts[(correction - 1) * segments + segment]): shift += gain /
// cells above zero.  Workgroup 0 of the segment stores the result in slot
// `correction` for the next pass.  Every thread receives it.
__device__ inline float fold_shift(const double *partials, int nparts, float *shifts, int correction, double *smem)
{
    __shared__ float shift_lds;    // its own word: smem is reused by the caller's reduction
    const int seg = blockIdx.y, segments = gridDim.y;
    partials += (int64_t)seg * 2 * nparts;
    double v0[kProjectPer], v1[kProjectPer];
#pragma unroll
    for (int u = 0; u < kProjectPer; u++) {
        const int k = threadIdx.x + u * kBlock;
        v0[u] = k < nparts ? partials[k] : 0.0;
        v1[u] = k < nparts ? partials[nparts + k] : 0.0;
    }
    double s[2] = {0.0, 0.0};
#pragma unroll
    for (int u = 0; u < kProjectPer; u++) {
        s[0] += v0[u];
        s[1] += v1[u];
    }
    block_reduce<double, 2, 0>(s, smem);
    if (threadIdx.x == 0) {
        float shift;
        if (correction == 0) {
            shift = s[1] > 0.0 ? (float)(s[0] / s[1]) : 0.0f;
        } else {
            shift = shifts[(correction - 1) * segments + seg];
            if (s[1] > 0.0) shift = (float)((double)shift + s[0] / s[1]);
        }
        if (blockIdx.x == 0) shifts[correction * segments + seg] = shift;
        shift_lds = shift;
    }
    __syncthreads();
    return shift_lds;
}

// mass the field gains at the current shift (clamp at zero included; masked:
// with the fixed change outside M), and the moved cells that stay above zero
// (same arithmetic as apply_delta_kernel); the shift comes from the previous
// pass's partials (`correction` as in fold_shift)
template <bool FROM_SURFACE, bool MASKED>
__global__ void __launch_bounds__(kBlock) clamp_stats_kernel(const float *b, ProjectArgs a, float *shifts,
                                                             int correction, const double *partials_in,
                                                             double *partials)
{
    __shared__ double smem[8];
    const float shift = fold_shift(partials_in, gridDim.x, shifts, correction, smem);
    const int64_t base = (int64_t)blockIdx.y * a.n;
    b += base;
    const float *src = a.src + base;
    partials += (int64_t)blockIdx.y * 2 * gridDim.x;
    double s[2] = {0.0, 0.0};
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < a.n; p += (int64_t)gridDim.x * blockDim.x) {
        const float d = change_at<FROM_SURFACE>(b, src, p);
        if (moved_at<FROM_SURFACE, MASKED>(a, base, p, d)) {
            float v = b[p] + (d - shift);
            v = v > 0.0f ? v : 0.0f;
            s[0] += (double)v - (double)b[p];     // mass the cell gains at this shift
            if (v > 0.0f) s[1] += 1.0;
        } else if (MASKED) {
            s[0] += (double)src[p] - (double)b[p];
        }
    }
    block_reduce<double, 2, 0>(s, smem);
    if (threadIdx.x == 0) {
        partials[blockIdx.x] = s[0];
        partials[gridDim.x + blockIdx.x] = s[1];
    }
}

template <bool FROM_SURFACE, bool MASKED>
__global__ void __launch_bounds__(kBlock) apply_delta_kernel(float *b, ProjectArgs a, float *shifts, int correction,
                                                             const double *partials_in)
{
    __shared__ double smem[8];
    const float shift = fold_shift(partials_in, gridDim.x, shifts, correction, smem);
    const int64_t base = (int64_t)blockIdx.y * a.n;
    b += base;
    const float *src = a.src + base;
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < a.n; p += (int64_t)gridDim.x * blockDim.x) {
        const float d = change_at<FROM_SURFACE>(b, src, p);
        const bool mv = moved_at<FROM_SURFACE, MASKED>(a, base, p, d);
        if (mv) {
            const float v = b[p] + (d - shift);
            b[p] = v > 0.0f ? v : 0.0f;
        } else if (MASKED) {
            b[p] = src[p];
        }
        if (!MASKED && a.mask_out) a.mask_out[base + p] = mv ? 1 : 0;
    }
}

// ---------------------------------------------------------------------------
// Lipschitz envelope: the largest surface below z whose differences across
// every edge of the four legacy directions obey |z_a - z_b| <= crit[d], i.e.
// z(p) = min over q of (z(q) + shortest edge-weighted path length q -> p).
// It only lowers cells, by exactly the amount their slopes exceed the limit,
// so after it the exact relaxation finds no excess.
//
// Every update z(p) <- min(z(p), z(q) + crit) keeps z a valid upper bound of
// the envelope, and the fixed point is the minimum over paths of the path sums
// evaluated from the source outwards (float rounding is monotone), so any
// update order that reaches a fixed point gives the same bits.  The kernel
// uses that freedom for directional Gauss-Seidel passes: one wave owns a
// 64 x 64 region (a kEnvTile^2 tile plus a kEnvHalo ring) in LDS; a down pass
// walks the rows top to bottom with lane = column, carrying the updated row
// above in registers and taking the neighbouring columns by DPP, so one pass
// moves information along any downward path of the region; up, right (lane =
// row) and left passes follow.  An optimal path between two cells uses at most
// two adjacent move directions, and every such pair lies inside one pass's
// three moves, so one cycle of the four passes settles a region up to
// rounding.  One cycle per launch (kEnvCycles): further cycles rarely moved
// anything, and a launch lasts as long as its busiest region (digging, 64 x
// 256^2: 245 frames/s with one cycle, 232 with up to 2, 3 or 6, the same 4
// launches per frame).  Tiles exchange through HBM between launches; the next
// launch stops the loop when no interior cell of any tile changed (a halo of
// one cell already makes that a global fixed point).
// ---------------------------------------------------------------------------
constexpr int kEnvRegion = 64;                          // region side = wave width
constexpr int kEnvHalo = 2;
constexpr int kEnvTile = kEnvRegion - 2 * kEnvHalo;     // interior cells per side
constexpr int kEnvStride = kEnvRegion + 1;              // LDS row stride: column walks are conflict-free
constexpr int kEnvWaves = 2;                            // regions per workgroup
constexpr int kEnvMaxCycles = 6;                        // upper limit of MV2_ENV_CYCLES
constexpr int kEnvCycles = 1;                           // default cycles per region and launch
constexpr int kEnvLoadBatch = 16;                       // region rows whose loads are in flight together
static_assert(kEnvRegion % kEnvLoadBatch == 0, "whole batches of region rows");

struct EnvControl {
    int launches, done, pad[2];
};

__global__ void add_surface_kernel(const float *b, const float *delta, float *z, int64_t n)
{
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += (int64_t)gridDim.x * blockDim.x)
        z[p] = b[p] + delta[p];
}

__device__ inline float env_from_prev_lane(float v)     // lane l gets lane l - 1 (lane 0: +inf)
{
    return __int_as_float(__builtin_amdgcn_update_dpp(0x7f800000, __float_as_int(v), 0x138, 0xf, 0xf, false));
}
__device__ inline float env_from_next_lane(float v)     // lane l gets lane l + 1 (lane 63: +inf)
{
    return __int_as_float(__builtin_amdgcn_update_dpp(0x7f800000, __float_as_int(v), 0x130, 0xf, 0xf, false));
}

// one directional pass over the wave's region.  HORZ: lane = region row, the
// walk goes over columns (FWD: left to right); otherwise lane = region column
// and the walk goes over rows (FWD: top to bottom).  Each cell is tested
// against the three cells one step back along the walk, already updated in
// this pass (the lane's own previous value and its two neighbour lanes' by
// DPP): a pass carries information along its three forward moves, and the four
// passes of a cycle test every cell against all eight neighbours, so a cycle
// that changes nothing leaves a fixed point of the region.  valid_* are bit
// masks of the region rows/columns inside the grid; cut_above / cut_below mark
// region rows whose edge to the row above / below crosses a seam between
// stacked fields.  Returns whether this lane lowered a cell; *moved_interior is
// set when that cell is in the tile's interior.
template <bool HORZ, bool FWD>
__device__ inline bool env_pass(float *t, int lane, uint64_t valid_rows, uint64_t valid_cols, uint64_t cut_above,
                                uint64_t cut_below, float c0, float c1, float c2, float c3, bool *moved_interior)
{
    constexpr float INF = __builtin_huge_valf();
    auto at = [&](int s) -> float * { return HORZ ? &t[lane * kEnvStride + s] : &t[s * kEnvStride + lane]; };
    const bool lane_valid = ((HORZ ? valid_rows : valid_cols) >> lane) & 1;
    const bool lane_inner = lane >= kEnvHalo && lane < kEnvHalo + kEnvTile;
    // HORZ: seams cut the edges to the lanes above / below (lane = row)
    const bool lane_cut_above = HORZ && ((cut_above >> lane) & 1);
    const bool lane_cut_below = HORZ && ((cut_below >> lane) & 1);
    // limits of the moves from the previous step's lane l - 1 / l + 1:
    // vertical down (r-1, c-1) d2, (r-1, c+1) d3; up (r+1, c-1) d3, (r+1, c+1) d2;
    // horizontal right (r-1, c-1) d2, (r+1, c-1) d3; left (r-1, c+1) d3, (r+1, c+1) d2
    const float lim_back = HORZ ? c0 : c1;
    const float lim_p = FWD ? c2 : c3, lim_n = FWD ? c3 : c2;
    bool changed = false;
    float prev = INF;                                   // updated value one step back
    float cur = *at(FWD ? 0 : kEnvRegion - 1);
    for (int k = 0; k < kEnvRegion; k++) {
        const int s = FWD ? k : kEnvRegion - 1 - k;
        const float next = k + 1 < kEnvRegion ? *at(FWD ? s + 1 : s - 1) : INF;
        const float prev_p = env_from_prev_lane(prev), prev_n = env_from_next_lane(prev);
        float m = cur;
        if (HORZ) {
            m = fminf(m, prev + lim_back);
            if (!lane_cut_above) m = fminf(m, prev_p + lim_p);
            if (!lane_cut_below) m = fminf(m, prev_n + lim_n);
        } else if (!(((FWD ? cut_above : cut_below) >> s) & 1)) {
            m = fminf(m, prev + lim_back);
            m = fminf(m, prev_p + lim_p);
            m = fminf(m, prev_n + lim_n);
        }
        const bool step_valid = ((HORZ ? valid_cols : valid_rows) >> s) & 1;
        if (lane_valid && step_valid && m < cur) {
            *at(s) = m;
            changed = true;
            if (lane_inner && s >= kEnvHalo && s < kEnvHalo + kEnvTile) *moved_interior = true;
            cur = m;
        }
        prev = cur;
        cur = next;
    }
    // the next pass reads what other lanes wrote
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
    return changed;
}

__global__ void __launch_bounds__(kEnvWaves * kWave) envelope_kernel(float *z, int ny, int nx, float c0, float c1,
                                                                      float c2, float c3, int tiles_x, int tiles,
                                                                      EnvControl *ctl, unsigned long long *spread,
                                                                      int32_t *tile_moved, int k, int seam,
                                                                      int max_cycles, int first_gen)
{
    __shared__ float lds[kEnvWaves][kEnvRegion * kEnvStride];
    __shared__ double smem[1];
    __shared__ int changed;
    // launch k - 1 recorded whether any interior cell changed in spread set (k-1) % 3
    EnvControl C = ctl[k & 1];
    if (!C.done && k > 0) {
        const double last = spread_read_max(spread + ((k - 1) % 3) * kSpreadWords, smem);
        if (last <= 0.0) C.done = 1;
    }
    if (C.done) {
        if (blockIdx.x == 0 && threadIdx.x == 0) ctl[(k + 1) & 1] = C;
        return;
    }
    if (blockIdx.x == 0) spread_clear(spread + ((k + 1) % 3) * kSpreadWords);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        EnvControl nc = C;
        nc.launches = C.launches + 1;
        ctl[(k + 1) & 1] = nc;
    }
    if (threadIdx.x == 0) changed = 0;
    __syncthreads();
    const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
    const int tile = blockIdx.x * kEnvWaves + wave;
    bool moved = false;
    // Quiet tiles: a region reads its tile's 3 x 3 tile neighbourhood.  When
    // no interior of those tiles changed in launch k - 1, nothing wrote into
    // the region during it, so its input equals that of the last launch that
    // processed it, which changed nothing: it is skipped.  Launch 0 reads the
    // flags of envelope_prep_kernel instead (parity 1, the call's generation
    // number first_gen): a region with no cell above a neighbour plus its
    // limit in its tile neighbourhood changes nothing.  first_gen == 0: launch
    // 0 runs every region.
    bool quiet = false;
    if (tile < tiles && !(k == 0 && first_gen == 0)) {
        const int ty = tiles / tiles_x, trow = tile / tiles_x, tcol = tile % tiles_x;
        int flag = 0;
        if (lane < 9) {
            const int rr = trow + lane / 3 - 1, cc = tcol + lane % 3 - 1;
            if (rr >= 0 && rr < ty && cc >= 0 && cc < tiles_x) {
                const int32_t f = tile_moved[((k - 1) & 1) * tiles + rr * tiles_x + cc];
                flag = k == 0 ? (f == first_gen) : f;
            }
        }
        quiet = !__any(flag);
    }
    if (tile < tiles && !quiet) {
        float *t = lds[wave];
        const int R0 = (tile / tiles_x) * kEnvTile - kEnvHalo, C0 = (tile % tiles_x) * kEnvTile - kEnvHalo;
        // this lane as a region row and as a region column
        const int grow = R0 + lane, gcol = C0 + lane;
        const bool row_in = grow >= 0 && grow < ny, col_in = gcol >= 0 && gcol < nx;
        const uint64_t valid_rows = __ballot(row_in), valid_cols = __ballot(col_in);
        const uint64_t cut_above = __ballot(row_in && seam > 0 && grow % seam == 0);
        const uint64_t cut_below = __ballot(row_in && seam > 0 && (grow + 1) % seam == 0);
        // the region's rows in batches of loads issued back to back: every
        // lane reads an in-grid cell (row and column clamped into the grid)
        // and cells outside the grid become +inf afterwards
        const int gc_in = min(max(gcol, 0), nx - 1);
#pragma unroll 1
        for (int r0 = 0; r0 < kEnvRegion; r0 += kEnvLoadBatch) {
            float v[kEnvLoadBatch];
#pragma unroll
            for (int u = 0; u < kEnvLoadBatch; u++) {
                const int gr_in = min(max(R0 + r0 + u, 0), ny - 1);
                v[u] = z[(int64_t)gr_in * nx + gc_in];
            }
#pragma unroll
            for (int u = 0; u < kEnvLoadBatch; u++) {
                const int r = r0 + u;
                t[r * kEnvStride + lane] = ((valid_rows >> r) & 1) && col_in ? v[u] : __builtin_huge_valf();
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
        for (int cycle = 0; cycle < max_cycles; cycle++) {
            bool ch = env_pass<false, true>(t, lane, valid_rows, valid_cols, cut_above, cut_below, c0, c1, c2, c3,
                                            &moved);
            ch |= env_pass<false, false>(t, lane, valid_rows, valid_cols, cut_above, cut_below, c0, c1, c2, c3,
                                         &moved);
            ch |= env_pass<true, true>(t, lane, valid_rows, valid_cols, cut_above, cut_below, c0, c1, c2, c3, &moved);
            ch |= env_pass<true, false>(t, lane, valid_rows, valid_cols, cut_above, cut_below, c0, c1, c2, c3,
                                        &moved);
            if (!__any(ch)) break;
        }
        if (__any(moved)) {
            for (int r = kEnvHalo; r < kEnvHalo + kEnvTile; r++)
                if (((valid_rows >> r) & 1) && col_in && lane >= kEnvHalo && lane < kEnvHalo + kEnvTile)
                    z[(int64_t)(R0 + r) * nx + gcol] = t[r * kEnvStride + lane];
        }
    }
    if (tile < tiles) {                              // wave-uniform
        const bool wave_moved = __any(moved);
        if (lane == 0) tile_moved[(k & 1) * tiles + tile] = wave_moved ? 1 : 0;
    }
    if (moved) changed = 1;     // benign race: every writer stores 1
    __syncthreads();
    if (threadIdx.x == 0 && changed) spread_max(spread + (k % 3) * kSpreadWords, 1.0);
}

// The call's set-up in one launch: z = b + delta when delta is given (else z
// holds the surface already); the loop state and spread sets zeroed
// (workgroup 0); with gen != 0, launch 0's flags (parity 1) set to gen for
// every tile with a cell that an update would lower (z(q) + lim < z(p) for a
// neighbour q across an edge the envelope uses; same float arithmetic, the
// neighbours' z formed from b + delta as above).  Flags of other tiles keep
// older values, which differ from gen: the generations of a context never
// repeat 0 or 1, the values the envelope launches write.
__global__ void __launch_bounds__(kBlock) envelope_prep_kernel(const float *b, const float *delta, float *z, int ny,
                                                               int nx, float c0, float c1, float c2, float c3,
                                                               int tiles_x, int tiles, int seam, int32_t *flags,
                                                               int gen, EnvControl *ctl, unsigned long long *spread)
{
    if (blockIdx.x == 0) {
        if (threadIdx.x < 2) ctl[threadIdx.x] = EnvControl{0, 0, {0, 0}};
        for (int i = threadIdx.x; i < 3 * kSpread; i += blockDim.x) spread[(int64_t)i * kSpreadStride] = 0ull;
    }
    const int64_t n = (int64_t)ny * nx;
    for (int64_t p = (int64_t)xcd_block() * kBlock + threadIdx.x; p < n; p += (int64_t)gridDim.x * kBlock) {
        if (!gen) {
            if (delta) z[p] = b[p] + delta[p];
            continue;
        }
        const int r = (int)(p / nx), c = (int)(p - (int64_t)r * nx);
        const bool has_l = c > 0, has_r = c + 1 < nx;
        const bool up_ok = r > 0 && !(seam > 0 && r % seam == 0);
        const bool dn_ok = r + 1 < ny && !(seam > 0 && (r + 1) % seam == 0);
        // the cell and its 8 neighbours, read together (a neighbour outside
        // the grid reads the cell itself and is not tested): centre, left,
        // right, up-left, up, up-right, down-left, down, down-right
        const int64_t cl = has_l ? -1 : 0, cr = has_r ? 1 : 0;
        const int64_t ru = r > 0 ? -(int64_t)nx : 0, rd = r + 1 < ny ? (int64_t)nx : 0;
        const int64_t q[9] = {p, p + cl, p + cr, p + ru + cl, p + ru, p + ru + cr, p + rd + cl, p + rd, p + rd + cr};
        float zn[9];
        if (delta) {
            float bb[9], dd[9];
#pragma unroll
            for (int i = 0; i < 9; i++) {
                bb[i] = b[q[i]];
                dd[i] = delta[q[i]];
            }
#pragma unroll
            for (int i = 0; i < 9; i++) zn[i] = bb[i] + dd[i];
            z[p] = zn[0];
        } else {
#pragma unroll
            for (int i = 0; i < 9; i++) zn[i] = z[q[i]];
        }
        const float v = zn[0];
        bool low = false;
        low |= has_l && zn[1] + c0 < v;
        low |= has_r && zn[2] + c0 < v;
        low |= up_ok && zn[4] + c1 < v;
        low |= up_ok && has_l && zn[3] + c2 < v;
        low |= up_ok && has_r && zn[5] + c3 < v;
        low |= dn_ok && zn[7] + c1 < v;
        low |= dn_ok && has_r && zn[8] + c2 < v;
        low |= dn_ok && has_l && zn[6] + c3 < v;
        if (low) flags[tiles + (r / kEnvTile) * tiles_x + c / kEnvTile] = gen;   // every writer stores gen
    }
}

// the loop's state after launch k - 1, written to pinned host memory, then the
// completion stamp
__global__ void envelope_result_kernel(const EnvControl *ctl, const unsigned long long *spread, int k, int *out,
                                       unsigned *stamp_word, unsigned stamp)
{
    __shared__ double smem[1];
    EnvControl C = ctl[k & 1];
    if (!C.done && k > 0 && spread_read_max(spread + ((k - 1) % 3) * kSpreadWords, smem) <= 0.0) C.done = 1;
    if (threadIdx.x == 0) {
        out[0] = C.launches;
        out[1] = C.done;
        host_stamp(stamp_word, stamp);
    }
}

}  // namespace mv2

using namespace mv2;

static int check_plain(mv2_context *ctx)
{
    if (!ctx) return set_error(nullptr, MV2_EINVAL, "[Surrogate] null context");
    hipSetDevice(ctx->device);
    return MV2_OK;
}

extern "C" int mv2_nn_features(mv2_context *ctx, const float *b, const float *h, int B, int H, int W, double dx,
                               double dy, float *feat)
{
    if (int rc = check_plain(ctx)) return rc;
    const int64_t n = (int64_t)B * H * W;
    hipLaunchKernelGGL(features_kernel, dim3(grid_for(n, kBlock, 8192)), dim3(kBlock), 0, ctx->stream, b, h, B, H,
                       W, (float)dx, (float)dy, feat);
    MV2_HIP(ctx, hipGetLastError());
    return MV2_OK;
}

extern "C" int mv2_nn_conv1(mv2_context *ctx, const float *feat, const float *w1, const float *b1, int B, int H,
                            int W, int dil, void *y)
{
    if (int rc = check_plain(ctx)) return rc;
    if (B < 1 || H < 1 || W < 1 || B > 65535 || (H + kConvRows - 1) / kConvRows > 65535)
        return set_error(ctx, MV2_EINVAL, "[Surrogate] invalid activation shape");
    const dim3 grid((W + 31) / 32, (H + kConvRows - 1) / kConvRows, B);
    hipLaunchKernelGGL(conv1_mfma_kernel, grid, dim3(kBlock), 0, ctx->stream, feat, w1, b1, B, H, W, dil,
                       (__bf16 *)y);
    MV2_HIP(ctx, hipGetLastError());
    return MV2_OK;
}

extern "C" int mv2_nn_conv_mfma(mv2_context *ctx, const void *x, int cx, const void *w, const float *bias,
                                const void *mask, void *y, int B, int H, int W, int dil, int relu, int mode)
{
    if (int rc = check_plain(ctx)) return rc;
    if (cx != 16 && cx != 32) return set_error(ctx, MV2_EINVAL, "[Surrogate] input channels must be 16 or 32");
    if (B < 1 || H < 1 || W < 1 || B > 65535 || dil < 1 || dil > 4096 || (conv_chains(H, dil) + 3) / 4 > 65535)
        return set_error(ctx, MV2_EINVAL, "[Surrogate] invalid activation shape");
    ConvArgs a{(const __bf16 *)x, (const __bf16 *)w, bias, (const __bf16 *)mask, y, B, H, W, dil, relu, mode};
    const dim3 grid((W + 31) / 32, (conv_chains(H, dil) + 3) / 4, B);
    if (cx == 16 && mask) hipLaunchKernelGGL((conv_mfma_kernel<1, true>), grid, dim3(kBlock), 0, ctx->stream, a);
    else if (cx == 16) hipLaunchKernelGGL((conv_mfma_kernel<1, false>), grid, dim3(kBlock), 0, ctx->stream, a);
    else if (mask) hipLaunchKernelGGL((conv_mfma_kernel<2, true>), grid, dim3(kBlock), 0, ctx->stream, a);
    else hipLaunchKernelGGL((conv_mfma_kernel<2, false>), grid, dim3(kBlock), 0, ctx->stream, a);
    MV2_HIP(ctx, hipGetLastError());
    return MV2_OK;
}

extern "C" int mv2_nn_pack_weights(mv2_context *ctx, const float *w, int cout, int cin, int cxf, int cxb, void *fwd,
                                   void *bwd)
{
    if (int rc = check_plain(ctx)) return rc;
    hipLaunchKernelGGL(pack_weights_kernel, dim3(36), dim3(kBlock), 0, ctx->stream, w, cout, cin, cxf, cxb,
                       (__bf16 *)fwd, (__bf16 *)bwd);
    MV2_HIP(ctx, hipGetLastError());
    return MV2_OK;
}

extern "C" int mv2_nn_expand_grad(mv2_context *ctx, const float *g, int64_t npix, void *out16)
{
    if (int rc = check_plain(ctx)) return rc;
    hipLaunchKernelGGL(expand_grad_kernel, dim3(grid_for(npix, kBlock, 8192)), dim3(kBlock), 0, ctx->stream, g,
                       npix, (__bf16 *)out16);
    MV2_HIP(ctx, hipGetLastError());
    return MV2_OK;
}

extern "C" int mv2_nn_weight_grad(mv2_context *ctx, const void *dz, int zmode, const void *x, int xmode, int B,
                                  int H, int W, int dil, int cout, int cin, float *dw, float *db)
{
    if (int rc = check_plain(ctx)) return rc;
    if ((zmode != 0 && zmode != 1) || (xmode != 0 && xmode != 1) || (zmode == 1 && xmode == 1) || cout < 1 ||
        cout > 32 || cin < 1 || cin > 32 || (zmode == 1 && cout != 1) || (xmode == 1 && cin > 4) || B < 1 || H < 1 ||
        W < 1 || dil < 1)
        return set_error(ctx, MV2_EINVAL, "[WarmStartNet] weight gradient: unsupported layer shape");
    const int64_t total = (int64_t)B * H * W;
    const bool mfma = zmode == 0 && xmode == 0;
    int64_t chunk = mfma ? 4096 : 1024;         // the narrow layers: 4 pixels per thread
    int nchunks = (int)((total + chunk - 1) / chunk);
    if (nchunks > 2048) {
        nchunks = 2048;
        chunk = (total + nchunks - 1) / nchunks;
        chunk = (chunk + kGradMfmaSlab - 1) / kGradMfmaSlab * kGradMfmaSlab;
        nchunks = (int)((total + chunk - 1) / chunk);
    }
    const size_t need = sizeof(float) * ((size_t)nchunks * 9 * 1024 + (size_t)nchunks * 32);
    if (int rc = ensure_scratch(ctx, ctx->nn_partials, need)) return rc;
    float *parts = (float *)ctx->nn_partials.ptr;
    float *bparts = parts + (size_t)nchunks * 9 * 1024;
    if (zmode == 0 && xmode == 0)
        hipLaunchKernelGGL(weight_grad_mfma_kernel, dim3(nchunks, 9), dim3(kBlock), 0, ctx->stream,
                           (const __bf16 *)dz, (const __bf16 *)x, B, H, W, dil, chunk, parts, bparts);
    else if (xmode == 1)
        hipLaunchKernelGGL(weight_grad_small_kernel<1>, dim3(nchunks, 9), dim3(kBlock), 0, ctx->stream, dz, x, B, H,
                           W, dil, chunk, parts, bparts);
    else
        hipLaunchKernelGGL(weight_grad_small_kernel<0>, dim3(nchunks, 9), dim3(kBlock), 0, ctx->stream, dz, x, B, H,
                           W, dil, chunk, parts, bparts);
    hipLaunchKernelGGL(reduce_grad_kernel, dim3(37), dim3(kBlock), 0, ctx->stream, parts, bparts, nchunks, cout, cin,
                       dw, db);
    MV2_HIP(ctx, hipGetLastError());
    return MV2_OK;
}

extern "C" int mv2_nn_mse(mv2_context *ctx, const float *pred, const float *target, int64_t n, float *grad,
                          double *loss_out)
{
    if (int rc = check_plain(ctx)) return rc;
    const int grid = grid_for(n, kBlock, 1024);
    if (int rc = ensure_scratch(ctx, ctx->nn_loss, sizeof(double) * 2048)) return rc;
    double *parts = (double *)ctx->nn_loss.ptr;
    hipLaunchKernelGGL(mse_kernel, dim3(grid), dim3(kBlock), 0, ctx->stream, pred, target, n, grad, parts);
    MV2_HIP(ctx, hipGetLastError());
    if (loss_out) {
        std::vector<double> host(grid);
        MV2_HIP(ctx, hipMemcpyAsync(host.data(), parts, sizeof(double) * grid, hipMemcpyDeviceToHost, ctx->stream));
        if (int rc_w = stream_wait(ctx, ctx->stream)) return rc_w;
        double s = 0.0;
        for (double v : host) s += v;
        *loss_out = s / (double)n;
    }
    return MV2_OK;
}

extern "C" int mv2_nn_excess_loss(mv2_context *ctx, const float *pred, const float *pre, int B, int H, int W,
                                  const double crit_diff[4], double weight, float *grad, double *loss_out)
{
    if (int rc = check_plain(ctx)) return rc;
    if (!pred || !pre || !grad || !crit_diff || B < 1 || H < 1 || W < 1)
        return set_error(ctx, MV2_EINVAL, "[WarmStart] invalid excess-loss arguments");
    const int64_t n = (int64_t)B * H * W;
    const int grid = grid_for(n, kBlock, 1024);
    if (int rc = ensure_scratch(ctx, ctx->nn_loss, sizeof(double) * 2048)) return rc;
    double *parts = (double *)ctx->nn_loss.ptr;
    hipLaunchKernelGGL(excess_loss_kernel, dim3(grid), dim3(kBlock), 0, ctx->stream, pred, pre, B, H, W,
                       (float)crit_diff[0], (float)crit_diff[1], (float)crit_diff[2], (float)crit_diff[3],
                       (float)weight, grad, parts);
    MV2_HIP(ctx, hipGetLastError());
    if (loss_out) {
        std::vector<double> host(grid);
        MV2_HIP(ctx, hipMemcpyAsync(host.data(), parts, sizeof(double) * grid, hipMemcpyDeviceToHost, ctx->stream));
        if (int rc_w = stream_wait(ctx, ctx->stream)) return rc_w;
        double acc = 0.0;
        for (double v : host) acc += v;
        *loss_out = weight * acc / (double)n;
    }
    return MV2_OK;
}

extern "C" int mv2_nn_adam(mv2_context *ctx, float *param, const float *grad, float *m, float *v, int n, double lr,
                           double beta1, double beta2, double eps, int step)
{
    if (int rc = check_plain(ctx)) return rc;
    const float bc1 = (float)(1.0 - pow(beta1, step)), bc2 = (float)(1.0 - pow(beta2, step));
    hipLaunchKernelGGL(adam_kernel, dim3(grid_for(n)), dim3(kBlock), 0, ctx->stream, param, grad, m, v, n, (float)lr,
                       (float)beta1, (float)beta2, (float)(1.0 - beta1), (float)(1.0 - beta2), (float)eps, bc1, bc2);
    MV2_HIP(ctx, hipGetLastError());
    return MV2_OK;
}

extern "C" int mv2_nn_project_w1(mv2_context *ctx, float *w1)
{
    if (int rc = check_plain(ctx)) return rc;
    hipLaunchKernelGGL(project_w1_kernel, dim3(1), dim3(64), 0, ctx->stream, w1);
    MV2_HIP(ctx, hipGetLastError());
    return MV2_OK;
}

extern "C" int mv2_relax_patches(mv2_context *ctx, float *h, int npatch, int H, int W, const double crit_diff[4],
                                 int max_iterations, double tolerance, int *iterations)
{
    if (int rc = check_plain(ctx)) return rc;
    if (H < 2 || W < 2 || H * W > kPatchMax)
        return set_error(ctx, MV2_EINVAL, "[MiniSlope] patch must be between 2x2 and 4096 cells");
    if (npatch < 0) return set_error(ctx, MV2_EINVAL, "[MiniSlope] negative patch count");
    if (npatch == 0) return MV2_OK;
    hipLaunchKernelGGL(relax_patches_kernel, dim3(npatch), dim3(kBlock), 0, ctx->stream, h, npatch, H, W,
                       (float)crit_diff[0], (float)crit_diff[1], (float)crit_diff[2], (float)crit_diff[3],
                       max_iterations, (float)tolerance, iterations);
    MV2_HIP(ctx, hipGetLastError());
    return MV2_OK;
}

constexpr int kClampCorrections = 4;

// Mass projection in kClampCorrections + 2 launches: the change statistics,
// kClampCorrections clamp passes and the apply pass, each of which folds the
// shift it uses from the previous pass's partials (ping-pong partial buffers,
// one shift slot per pass) instead of a one-workgroup shift launch between
// the passes.
template <bool FROM_SURFACE, bool MASKED = false>
static int project_change(mv2_context *ctx, float *b, const float *src, int64_t n, double eps, int segments = 1,
                          const uint8_t *mask_in = nullptr, uint8_t *mask_out = nullptr)
{
    const int grid = grid_for(n, kBlock, kProjectParts);    // per segment: the same partials as one field
    const size_t parts_n = (size_t)2 * grid * segments;
    const size_t shift_n = (size_t)(kClampCorrections + 1) * segments;
    if (int rc = ensure_scratch(ctx, ctx->nn_loss, sizeof(double) * 2 * parts_n + sizeof(float) * shift_n + 64))
        return rc;
    double *parts[2] = {(double *)ctx->nn_loss.ptr, (double *)ctx->nn_loss.ptr + parts_n};
    float *shifts = (float *)(parts[1] + parts_n);
    const ProjectArgs a{src, n, (float)eps, mask_in, mask_out};
    hipLaunchKernelGGL((delta_stats_kernel<FROM_SURFACE, MASKED>), dim3(grid, segments), dim3(kBlock), 0, ctx->stream,
                       b, a, parts[0]);
    for (int c = 0; c < kClampCorrections; c++)
        hipLaunchKernelGGL((clamp_stats_kernel<FROM_SURFACE, MASKED>), dim3(grid, segments), dim3(kBlock), 0,
                           ctx->stream, b, a, shifts, c, parts[c & 1], parts[(c + 1) & 1]);
    hipLaunchKernelGGL((apply_delta_kernel<FROM_SURFACE, MASKED>), dim3(grid, segments), dim3(kBlock), 0, ctx->stream,
                       b, a, shifts, kClampCorrections, parts[kClampCorrections & 1]);
    MV2_HIP(ctx, hipGetLastError());
    return MV2_OK;
}

extern "C" int mv2_nn_apply_delta(mv2_context *ctx, float *b, const float *delta, int64_t n, double eps)
{
    if (int rc = check_plain(ctx)) return rc;
    if (!b || !delta || n < 1) return set_error(ctx, MV2_EINVAL, "[WarmStart] invalid apply_delta arguments");
    return project_change<false>(ctx, b, delta, n, eps);
}

extern "C" int mv2_nn_commit_surface(mv2_context *ctx, float *b, const float *z, int64_t n, double eps)
{
    if (int rc = check_plain(ctx)) return rc;
    if (!b || !z || n < 1) return set_error(ctx, MV2_EINVAL, "[WarmStart] invalid commit arguments");
    return project_change<true>(ctx, b, z, n, eps);
}

static int envelope_run(mv2_context *ctx, const float *b, const float *delta, float *z, int ny, int nx,
                        const double crit_diff[4], int max_launches, int *launches_out, int seam, bool prior);

extern "C" int mv2_envelope(mv2_context *ctx, const float *b, const float *delta, float *z, int ny, int nx,
                            const double crit_diff[4], int max_launches, int *launches_out)
{
    return envelope_run(ctx, b, delta, z, ny, nx, crit_diff, max_launches, launches_out, 0, true);
}

extern "C" int mv2_batch_envelope(mv2_context *ctx, const float *b, const float *delta, float *z, int count,
                                  int ny, int nx, const double crit_diff[4], int max_launches, int *launches_out)
{
    if (count < 1 || (int64_t)count * ny >= ((int64_t)1 << 31))
        return set_error(ctx, MV2_EINVAL, "[WarmStart] invalid batch of envelopes");
    return envelope_run(ctx, b, delta, z, count * ny, nx, crit_diff, max_launches, launches_out, ny, false);
}

extern "C" int mv2_batch_commit_surface(mv2_context *ctx, float *b, const float *z, int count, int64_t n, double eps)
{
    if (int rc = check_plain(ctx)) return rc;
    if (!b || !z || n < 1 || count < 1 || count > 65535)
        return set_error(ctx, MV2_EINVAL, "[WarmStart] invalid commit arguments");
    return project_change<true>(ctx, b, z, n, eps, count);
}

// Settle rounds: the commit, then (rounds - 1) times the Lipschitz envelope of
// the committed field and the masked projection onto the first commit's moved
// cells (stacked fields: the envelope does not cross their seams, every field
// keeps its own shift).  A round whose envelope changes nothing ends them.
static int commit_settled(mv2_context *ctx, float *b, float *z, int count, int ny, int nx, const double crit_diff[4],
                          double eps, int rounds, int max_launches, int *launches_out)
{
    if (int rc = check_plain(ctx)) return rc;
    if (!b || !z || !crit_diff || count < 1 || count > 65535 || ny < 1 || nx < 1 || rounds < 1 || max_launches < 1)
        return set_error(ctx, MV2_EINVAL, "[WarmStart] invalid settle arguments");
    const int64_t n = (int64_t)ny * nx;
    if (n * count >= ((int64_t)1 << 31)) return set_error(ctx, MV2_EINVAL, "[WarmStart] grid too large");
    if (int rc = ensure_scratch(ctx, ctx->nn_mask, (size_t)(n * count))) return rc;
    uint8_t *mask = (uint8_t *)ctx->nn_mask.ptr;
    if (int rc = project_change<true>(ctx, b, z, n, eps, count, nullptr, mask)) return rc;
    int total = 0;
    for (int r = 1; r < rounds; r++) {
        MV2_HIP(ctx, hipMemcpyAsync(z, b, sizeof(float) * (size_t)(n * count), hipMemcpyDeviceToDevice, ctx->stream));
        int l = 0;
        if (int rc = envelope_run(ctx, nullptr, nullptr, z, count * ny, nx, crit_diff, max_launches, &l, ny,
                                  count == 1))
            return rc;
        total += l;
        if (l <= 1) break;               // nothing moved: the field is its own envelope
        if (int rc = project_change<true, true>(ctx, b, z, n, eps, count, mask)) return rc;
    }
    if (launches_out) launches_out[0] = total;
    return MV2_OK;
}

extern "C" int mv2_nn_commit_settled(mv2_context *ctx, float *b, float *z, int ny, int nx, const double crit_diff[4],
                                     double eps, int rounds, int max_launches, int *launches_out)
{
    return commit_settled(ctx, b, z, 1, ny, nx, crit_diff, eps, rounds, max_launches, launches_out);
}

extern "C" int mv2_batch_commit_settled(mv2_context *ctx, float *b, float *z, int count, int ny, int nx,
                                        const double crit_diff[4], double eps, int rounds, int max_launches,
                                        int *launches_out)
{
    return commit_settled(ctx, b, z, count, ny, nx, crit_diff, eps, rounds, max_launches, launches_out);
}

static int envelope_run(mv2_context *ctx, const float *b, const float *delta, float *z, int ny, int nx,
                        const double crit_diff[4], int max_launches, int *launches_out, int seam, bool prior)
{
    if (int rc = check_plain(ctx)) return rc;
    if (!z || !crit_diff || ny < 1 || nx < 1 || max_launches < 1 || (delta && !b))
        return set_error(ctx, MV2_EINVAL, "[WarmStart] invalid envelope arguments");
    if ((int64_t)ny * nx >= ((int64_t)1 << 31)) return set_error(ctx, MV2_EINVAL, "[WarmStart] grid too large");
    for (int d = 0; d < 4; d++)
        if (!(crit_diff[d] >= 0.0)) return set_error(ctx, MV2_EINVAL, "[WarmStart] critical differences must be >= 0");
    const int64_t n = (int64_t)ny * nx;
    hipStream_t st = ctx->stream;
    if (int rc = ensure_scratch(ctx, ctx->spread, kSpreadScratchBytes)) return rc;
    unsigned long long *spread = (unsigned long long *)ctx->spread.ptr + (size_t)kEnvSpreadSet * kSpreadWords;
    EnvControl *ctl = (EnvControl *)((char *)ctx->sync.ptr + 2048);
    const int tiles_x = (nx + kEnvTile - 1) / kEnvTile;
    const int tiles = tiles_x * ((ny + kEnvTile - 1) / kEnvTile);
    const int blocks = (tiles + kEnvWaves - 1) / kEnvWaves;
    if (int rc = ensure_scratch(ctx, ctx->env_moved, 2 * sizeof(int32_t) * (size_t)tiles)) return rc;
    int32_t *tile_moved = (int32_t *)ctx->env_moved.ptr;
    // launch 0 of one field runs only the regions next to a cell that an update
    // would lower (flags from the prep pass: 701^2 learned relax-to-rest 0.50
    // against 0.52 ms); for a batch of fields, whose predictions leave such
    // cells in most regions, the flag test costs more than it skips (64 x
    // 256^2: 22 us per pass against no shorter first launch), so every region
    // runs (generation 0)
    int gen = 0;
    if (prior) {
        gen = ++ctx->env_gen;
        if (gen < 2) gen = ctx->env_gen = 2;       // 0 and 1 are the values the envelope launches write
    }
    // the prep pass reads b + delta at the neighbours of a cell while writing
    // z at the cell: with z aliasing delta that needs the surface in place first
    const bool alias = delta && z == delta && gen != 0;
    if (alias) hipLaunchKernelGGL(add_surface_kernel, dim3(grid_for(n, kBlock, 2048)), dim3(kBlock), 0, st, b, delta, z, n);
    hipLaunchKernelGGL(envelope_prep_kernel, dim3(grid_for(n, kBlock, 4096)), dim3(kBlock), 0, st, alias ? nullptr : b,
                       alias ? nullptr : delta, z, ny, nx, (float)crit_diff[0], (float)crit_diff[1],
                       (float)crit_diff[2], (float)crit_diff[3], tiles_x, tiles, seam, tile_moved, gen, ctl, spread);
    // cycles of the four passes per region and launch: more settle a region
    // in one launch, fewer bound the launch by its slowest region (a launch
    // lasts as long as its busiest region); MV2_ENV_CYCLES = 1..6 overrides
    int cycles = kEnvCycles;
    if (const char *e = getenv("MV2_ENV_CYCLES")) {
        const int forced = atoi(e);
        if (forced >= 1 && forced <= kEnvMaxCycles) cycles = forced;
    }
    // as many launches as the last call of the same kind needed (the result
    // pass after them finds the loop ended when the last one moved nothing),
    // then doubling rounds
    int &predicted = ctx->env_predicted[prior ? 1 : 0];
    int k = 0, chunk = predicted > 0 ? predicted : 4;
    int *hres = (int *)((char *)ctx->host_result.ptr + 3072);
    static_assert(3072 + 2 * sizeof(int) <= 4096, "the envelope result in the pinned block");
    for (;;) {
        if (chunk > max_launches - k) chunk = max_launches - k;
        for (int j = 0; j < chunk; j++, k++)
            hipLaunchKernelGGL(envelope_kernel, dim3(blocks), dim3(kEnvWaves * kWave), 0, st, z, ny, nx,
                               (float)crit_diff[0], (float)crit_diff[1], (float)crit_diff[2], (float)crit_diff[3],
                               tiles_x, tiles, ctl, spread, tile_moved, k, seam, cycles, gen);
        const unsigned stamp = next_stamp(ctx, 1);
        hipLaunchKernelGGL(envelope_result_kernel, dim3(1), dim3(kWave), 0, st, ctl, spread, k, hres,
                           stamp_words(ctx), stamp);
        MV2_HIP(ctx, hipGetLastError());
        if (int rc_w = stamp_wait(ctx, st, stamp, 1)) return rc_w;
        if (hres[1] || k >= max_launches) break;
        chunk = 2 * chunk;
    }
    if (hres[1]) predicted = hres[0];
    if (launches_out) launches_out[0] = hres[0];
    if (!hres[1]) return set_error(ctx, MV2_ENUMERIC, "[WarmStart] envelope did not converge in %d launches", max_launches);
    return MV2_OK;
}
