// Context lifecycle, error reporting and scratch ownership of libmv2.
//
// A context owns one HIP stream handle (borrowed from the caller or the null
// stream) and the scratch buffers of its kernels.  Field arrays stay owned by
// the caller (Python/torch); the library borrows their pointers at bind time,
// which is the ownership split of WarpRuntime + DeviceBulkState in the
// reference (pkg/performance/warp_backend.py:131-175,
// pkg/runtime/bulk_state_authority.py:543-621).
#include "mv2_internal.hpp"

#include <atomic>
#include <chrono>
#include <cstdarg>
#include <cstring>

namespace mv2 {

static thread_local std::string g_thread_error;

int set_error(mv2_context *ctx, int code, const char *fmt, ...)
{
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    if (ctx) ctx->last_error = buf;
    g_thread_error = buf;
    return code;
}

int check_hip(mv2_context *ctx, hipError_t err, const char *what)
{
    return set_error(ctx, MV2_EDEVICE, "HIP error %d (%s) in %s", (int)err, hipGetErrorString(err), what);
}

int ensure_scratch(mv2_context *ctx, Scratch &s, size_t bytes, bool host_pinned)
{
    if (bytes == 0) bytes = 16;
    if (s.ptr && s.bytes >= bytes) return MV2_OK;
    if (s.ptr) {
        // a grown buffer replaces the old one; work queued on the stream that
        // still reads it must finish first
        if (ctx->stream) hipStreamSynchronize(ctx->stream);
        else hipDeviceSynchronize();
        if (host_pinned) hipHostFree(s.ptr);
        else hipFree(s.ptr);
        s.ptr = nullptr;
        s.bytes = 0;
    }
    hipError_t e = host_pinned ? hipHostMalloc(&s.ptr, bytes, hipHostMallocDefault)
                               : hipMalloc(&s.ptr, bytes);
    if (e != hipSuccess) {
        s.ptr = nullptr;
        return check_hip(ctx, e, host_pinned ? "hipHostMalloc(scratch)" : "hipMalloc(scratch)");
    }
    s.bytes = bytes;
    return MV2_OK;
}

constexpr long long kSpinLimitNs = 2000000;

static long long elapsed_ns(std::chrono::steady_clock::time_point t0)
{
    return std::chrono::duration_cast<std::chrono::nanoseconds>(std::chrono::steady_clock::now() - t0).count();
}

// Wait for the work queued on a stream by polling a marker recorded behind
// it.  Measured on MI355X from the launch call of one small kernel to the
// return: 11.7 us polling the marker, 12.4 us in hipStreamSynchronize and
// 26 us polling hipStreamQuery.  After kSpinLimitNs the blocking synchronize
// takes over; errors of the stream come back either way.
int stream_wait(mv2_context *ctx, hipStream_t st)
{
    if (ctx->ev_wait && hipEventRecord(ctx->ev_wait, st) == hipSuccess) {
        const auto t0 = std::chrono::steady_clock::now();
        for (;;) {
            const hipError_t e = hipEventQuery(ctx->ev_wait);
            if (e == hipSuccess) return MV2_OK;
            if (e != hipErrorNotReady) break;            // the synchronize below reports it
            if (elapsed_ns(t0) > kSpinLimitNs) break;
        }
    }
    const hipError_t e = hipStreamSynchronize(st);
    return e == hipSuccess ? MV2_OK : check_hip(ctx, e, "hipStreamSynchronize");
}

unsigned next_stamp(mv2_context *ctx, int words)
{
    if (!ctx->host_stamps.ptr || words < 1 || words > kStampWords) return 0u;
    if (++ctx->stamp_seq == 0u) ctx->stamp_seq = 1u;    // 0 means "no stamps"
    return ctx->stamp_seq;
}

unsigned *stamp_words(mv2_context *ctx) { return (unsigned *)ctx->host_stamps.ptr; }

// Spin until every stamp word of the launch carries the call's number: the
// results written to pinned memory before the stamps are then visible
// (measured on MI355X from the launch call to the return: a pinned flag
// written by one more kernel 9.9 us, the marker 11.7 us).  A launch that has
// not stamped within kSpinLimitNs (a long call, or a device error that
// stopped the stream) is waited for on the stream, which also reports the
// error.
int stamp_wait(mv2_context *ctx, hipStream_t st, unsigned stamp, int words)
{
    if (!stamp) return stream_wait(ctx, st);
    const volatile unsigned *w = stamp_words(ctx);
    const auto t0 = std::chrono::steady_clock::now();
    int done = 0;
    for (;;) {
        while (done < words && w[done] == stamp) done++;
        if (done == words) {
            std::atomic_thread_fence(std::memory_order_acquire);
            return MV2_OK;
        }
        if (elapsed_ns(t0) > kSpinLimitNs) break;
    }
    return stream_wait(ctx, st);
}

static void release(Scratch &s, bool host_pinned = false)
{
    if (!s.ptr) return;
    if (host_pinned) hipHostFree(s.ptr);
    else hipFree(s.ptr);
    s.ptr = nullptr;
    s.bytes = 0;
}

}  // namespace mv2

using namespace mv2;

extern "C" int mv2_abi_version(void) { return MV2_ABI_VERSION; }

extern "C" double mv2_mobile_quantum(void) { return kQuantum; }

extern "C" int mv2_create(int device, mv2_context **out)
{
    if (!out) return set_error(nullptr, MV2_EINVAL, "[mv2_create] null output pointer");
    *out = nullptr;
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess || count <= 0)
        return set_error(nullptr, MV2_EDEVICE, "HIP backend unavailable: no GPU device (%s)",
                         hipGetErrorString(e));
    if (device < 0 || device >= count)
        return set_error(nullptr, MV2_EINVAL, "[mv2_create] device %d outside 0..%d", device, count - 1);
    e = hipSetDevice(device);
    if (e != hipSuccess) return check_hip(nullptr, e, "hipSetDevice");
    hipDeviceProp_t prop;
    e = hipGetDeviceProperties(&prop, device);
    if (e != hipSuccess) return check_hip(nullptr, e, "hipGetDeviceProperties");
    if (!prop.cooperativeLaunch)
        return set_error(nullptr, MV2_EDEVICE, "device %d lacks cooperative launch", device);
    mv2_context *ctx = new mv2_context();
    ctx->device = device;
    ctx->cu_count = prop.multiProcessorCount;
    if (int rc = ensure_scratch(ctx, ctx->sync, 4096)) { delete ctx; return rc; }
    if (int rc = ensure_scratch(ctx, ctx->result, 4096)) { delete ctx; return rc; }
    if (int rc = ensure_scratch(ctx, ctx->host_result, 4096, true)) { delete ctx; return rc; }
    if (int rc = ensure_scratch(ctx, ctx->host_stamps, sizeof(unsigned) * kStampWords, true)) {
        mv2_destroy(ctx);
        return rc;
    }
    memset(ctx->host_stamps.ptr, 0, sizeof(unsigned) * kStampWords);
    // with `device` current: the per-device address of the allocation-free scratch
    if (int rc = noalloc_scratch_address(ctx, &ctx->noalloc_scratch)) {
        mv2_destroy(ctx);
        return rc;
    }
    if (hipEventCreate(&ctx->ev_start) != hipSuccess || hipEventCreate(&ctx->ev_stop) != hipSuccess ||
        hipEventCreateWithFlags(&ctx->ev_wait, hipEventDisableTiming) != hipSuccess) {
        mv2_destroy(ctx);
        return set_error(nullptr, MV2_EDEVICE, "hipEventCreate failed");
    }
    *out = ctx;
    return MV2_OK;
}

extern "C" void mv2_destroy(mv2_context *ctx)
{
    if (!ctx) return;
    hipSetDevice(ctx->device);
    if (ctx->stream) hipStreamSynchronize(ctx->stream);
    else hipDeviceSynchronize();
    release(ctx->alt_h);
    release(ctx->alt_qx);
    release(ctx->alt_qy);
    release(ctx->sync);
    release(ctx->partials);
    release(ctx->result);
    release(ctx->relax_alt);
    release(ctx->relax_quiet);
    release(ctx->env_moved);
    release(ctx->nn_mask);
    release(ctx->label_aux);
    release(ctx->nn_partials);
    release(ctx->nn_loss);
    mv2_batch_release(ctx);
    release(ctx->step_diag);
    release(ctx->spread);
    release(ctx->frontier);
    release(ctx->patch);
    release(ctx->host_result, true);
    release(ctx->host_stamps, true);
    if (ctx->ev_start) hipEventDestroy(ctx->ev_start);
    if (ctx->ev_stop) hipEventDestroy(ctx->ev_stop);
    if (ctx->ev_wait) hipEventDestroy(ctx->ev_wait);
    for (hipEvent_t e : ctx->sub_events) hipEventDestroy(e);
    for (mv2::StepGraph &g : ctx->step_graphs) hipGraphExecDestroy(g.exec);
    if (ctx->capture_stream) hipStreamDestroy(ctx->capture_stream);
    delete ctx;
}

extern "C" const char *mv2_last_error(const mv2_context *ctx)
{
    if (ctx) return ctx->last_error.c_str();
    return g_thread_error.c_str();
}

extern "C" int mv2_set_stream(mv2_context *ctx, void *hip_stream)
{
    if (!ctx) return set_error(nullptr, MV2_EINVAL, "[mv2_set_stream] null context");
    ctx->stream = (hipStream_t)hip_stream;
    return MV2_OK;
}

extern "C" int mv2_synchronize(mv2_context *ctx)
{
    if (!ctx) return set_error(nullptr, MV2_EINVAL, "[mv2_synchronize] null context");
    hipSetDevice(ctx->device);
    MV2_HIP(ctx, ctx->stream ? hipStreamSynchronize(ctx->stream) : hipDeviceSynchronize());
    return MV2_OK;
}

extern "C" int mv2_bind(mv2_context *ctx, int precision, int ny, int nx, double dx, double dy,
                        void *z_base, void *b_eff, void *mobile, void *momentum_x, void *momentum_y)
{
    if (!ctx) return set_error(nullptr, MV2_EINVAL, "[MobileV2Production] null context");
    if (precision != MV2_FP64 && precision != MV2_FP32)
        return set_error(ctx, MV2_EINVAL, "[MobileV2Production] unknown precision %d", precision);
    if (ny < 2 || nx < 2)
        return set_error(ctx, MV2_EINVAL, "[TerrainGrid] nx and ny must be >= 2; nx=%d, ny=%d", nx, ny);
    if ((int64_t)ny * nx > (int64_t)1 << 30)
        return set_error(ctx, MV2_EINVAL, "[MobileV2Production] grid of %d x %d cells is too large", ny, nx);
    if (!(dx > 0.0) || !(dy > 0.0))
        return set_error(ctx, MV2_EINVAL, "[TerrainGrid] cell spacing must be positive metres; dx=%g, dy=%g", dx, dy);
    if (!z_base || !b_eff || !mobile || !momentum_x || !momentum_y)
        return set_error(ctx, MV2_EINVAL, "[MobileV2Production] missing state pointer");
    hipSetDevice(ctx->device);
    const size_t n = (size_t)ny * nx;
    const size_t es = precision == MV2_FP64 ? 8 : 4;
    if (int rc = ensure_scratch(ctx, ctx->alt_h, n * es)) return rc;
    if (int rc = ensure_scratch(ctx, ctx->alt_qx, n * es)) return rc;
    if (int rc = ensure_scratch(ctx, ctx->alt_qy, n * es)) return rc;
    ctx->precision = precision;
    ctx->ny = ny;
    ctx->nx = nx;
    ctx->dx = dx;
    ctx->dy = dy;
    ctx->z_base = z_base;
    ctx->b_eff = b_eff;
    ctx->mobile = mobile;
    ctx->qx = momentum_x;
    ctx->qy = momentum_y;
    ctx->bound = 1;
    state_written(ctx);
    return MV2_OK;
}
