// context.hip -- hip_context: device, stream, buffers, errors and timing.
// Replaces MetalContext (src/metal_context.h:14-100, src/metal_context.mm:13-60,448-461).
#include "hc_internal.h"

#include <cstring>

namespace hc {

int fail(hc_ctx* ctx, int code, const std::string& msg) {
    if (ctx) ctx->last_error = msg;
    return code;
}

int hip_fail(hc_ctx* ctx, hipError_t e, const char* what) {
    return fail(ctx, HC_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

int check_launch(hc_ctx* ctx, const char* name) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(ctx, e, name);
    return 0;
}

static hipEvent_t take_event(hc_ctx* ctx) {
    if (!ctx->event_pool.empty()) {
        hipEvent_t e = ctx->event_pool.back();
        ctx->event_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

KernelTimer::KernelTimer(hc_ctx* c, int k) : ctx(c), kernel(k) {
    if (ctx && (ctx->timing_mask >> k) & 1u) {
        a = take_event(ctx);
        b = take_event(ctx);
        // a launch that cannot be timed is simply not timed
        if (!a || !b || hipEventRecord(a, ctx->stream) != hipSuccess) {
            if (a) ctx->event_pool.push_back(a);
            if (b) ctx->event_pool.push_back(b);
            a = b = nullptr;
        }
    }
}

KernelTimer::~KernelTimer() {
    if (a && b) {
        if (hipEventRecord(b, ctx->stream) == hipSuccess) {
            ctx->pending.push_back({kernel, a, b});
        } else {
            ctx->event_pool.push_back(a);
            ctx->event_pool.push_back(b);
        }
    }
}

int probe_kernels_diffusion(hc_ctx* ctx);
int probe_kernels_cells(hc_ctx* ctx);
int probe_kernels_cells2(hc_ctx* ctx);
int probe_kernels_mirror(hc_ctx* ctx);
int probe_kernels_batch(hc_ctx* ctx);
int probe_kernels_motility(hc_ctx* ctx);
int probe_kernels_phenotype(hc_ctx* ctx);
int probe_kernels_neighbors(hc_ctx* ctx);

}  // namespace hc

using namespace hc;

extern "C" {

size_t hc_cells_bytes(uint32_t max_cells) {
    return (size_t)max_cells * (HC_NUM_F64_FIELDS * sizeof(double) + HC_NUM_U32_FIELDS * sizeof(uint32_t));
}

size_t hc_cells_field_offset(uint32_t field, uint32_t max_cells) {
    if (field < HC_NUM_F64_FIELDS) return (size_t)field * sizeof(double) * max_cells;
    return (size_t)HC_NUM_F64_FIELDS * sizeof(double) * max_cells +
           (size_t)(field - HC_NUM_F64_FIELDS) * sizeof(uint32_t) * max_cells;
}

size_t hc_rates_bytes(uint32_t max_cells, uint32_t n_substrates) {
    return (size_t)4 * max_cells * n_substrates * sizeof(double);
}

int hc_create(hc_ctx** out, int device_ordinal) {
    if (!out) return HC_ERR_ARG;
    *out = nullptr;
    int ndev = 0;
    hipError_t e = hipGetDeviceCount(&ndev);
    if (e != hipSuccess || ndev <= 0) {
        fprintf(stderr, "[hip_context] no HIP device: %s\n", hipGetErrorString(e));
        return HC_ERR_HIP;
    }
    if (device_ordinal < 0 || device_ordinal >= ndev) return HC_ERR_ARG;
    hc_ctx* ctx = new hc_ctx();
    ctx->device = device_ordinal;
    if ((e = hipSetDevice(device_ordinal)) != hipSuccess ||
        (e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking)) != hipSuccess ||
        (e = hipMalloc(&ctx->d_scalar, 256)) != hipSuccess) {
        fprintf(stderr, "[hip_context] setup failed: %s\n", hipGetErrorString(e));
        delete ctx;
        return HC_ERR_HIP;
    }
    *out = ctx;
    return 0;
}

void hc_destroy(hc_ctx* ctx) {
    if (!ctx) return;
    // teardown: nothing to report to, so failures are ignored
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    for (auto& p : ctx->pending) { (void)hipEventDestroy(p.a); (void)hipEventDestroy(p.b); }
    for (auto e : ctx->event_pool) (void)hipEventDestroy(e);
    if (ctx->d_scalar) (void)hipFree(ctx->d_scalar);
    if (ctx->d_list) (void)hipFree(ctx->d_list);
    if (ctx->d_query) (void)hipFree(ctx->d_query);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

int hc_load_kernels(hc_ctx* ctx, const char* code_object_or_null) {
    if (!ctx) return HC_ERR_ARG;
    if (code_object_or_null)
        return fail(ctx, HC_ERR_ARG, "hc_load_kernels: external code objects are not supported; pass NULL");
    hipError_t e = hipSetDevice(ctx->device);
    if (e != hipSuccess) return hip_fail(ctx, e, "hc_load_kernels: hipSetDevice");
    int r = probe_kernels_diffusion(ctx);
    if (r) return r;
    r = probe_kernels_cells(ctx);
    if (r) return r;
    r = probe_kernels_cells2(ctx);
    if (r) return r;
    r = probe_kernels_mirror(ctx);
    if (r) return r;
    r = totals_scratch_init(ctx);
    if (r) return r;
    r = probe_kernels_batch(ctx);
    if (r) return r;
    r = probe_kernels_motility(ctx);
    if (r) return r;
    r = probe_kernels_phenotype(ctx);
    if (r) return r;
    r = probe_kernels_neighbors(ctx);
    if (r) return r;
    ctx->kernels_loaded = true;
    return 0;
}

int hc_alloc(hc_ctx* ctx, size_t bytes, hc_buf** out) {
    if (!ctx || !out) return HC_ERR_ARG;
    *out = nullptr;
    hipError_t e0 = hipSetDevice(ctx->device);
    if (e0 != hipSuccess) return hip_fail(ctx, e0, "hc_alloc: hipSetDevice");
    hc_buf* b = new hc_buf();
    b->ctx = ctx;
    b->bytes = bytes;
    if (bytes) {
        hipError_t e = hipMalloc(&b->dptr, bytes);
        if (e != hipSuccess) {
            delete b;
            return fail(ctx, HC_ERR_NOMEM, std::string("hc_alloc(") + std::to_string(bytes) +
                                               " B): " + hipGetErrorString(e));
        }
        e = hipMemsetAsync(b->dptr, 0, bytes, ctx->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
        if (e != hipSuccess) {
            (void)hipFree(b->dptr);
            delete b;
            return hip_fail(ctx, e, "hc_alloc: zero fill");
        }
    }
    *out = b;
    return 0;
}

void hc_free(hc_buf* b) {
    if (!b) return;
    if (b->ctx) {
        // release: queued work may still use the buffer; failures have no caller to go to
        (void)hipSetDevice(b->ctx->device);
        (void)hipStreamSynchronize(b->ctx->stream);
    }
    if (b->dptr) (void)hipFree(b->dptr);
    if (b->host) (void)hipHostFree(b->host);
    delete b;
}

void* hc_host_mirror(hc_buf* b) {
    if (!b) return nullptr;
    if (!b->host && b->bytes) {
        if (hipHostMalloc(&b->host, b->bytes, hipHostMallocDefault) != hipSuccess) {
            b->host = nullptr;
            if (b->ctx) b->ctx->last_error = "hc_host_mirror: pinned allocation failed";
            return nullptr;
        }
        std::memset(b->host, 0, b->bytes);
    }
    return b->host;
}

void* hc_device_ptr(hc_buf* b) { return b ? b->dptr : nullptr; }
size_t hc_buf_size(const hc_buf* b) { return b ? b->bytes : 0; }

int hc_upload(hc_ctx* ctx, hc_buf* b, size_t off, size_t n) {
    if (!ctx || !b || off > b->bytes || n > b->bytes - off)
        return fail(ctx, HC_ERR_ARG, "hc_upload: range outside buffer");
    if (!n) return 0;
    if (!hc_host_mirror(b)) return fail(ctx, HC_ERR_NOMEM, "hc_upload: no host mirror");
    hipError_t e = hipMemcpyAsync((char*)b->dptr + off, (char*)b->host + off, n,
                                  hipMemcpyHostToDevice, ctx->stream);
    return e == hipSuccess ? 0 : hip_fail(ctx, e, "hc_upload");
}

int hc_download(hc_ctx* ctx, hc_buf* b, size_t off, size_t n) {
    if (!ctx || !b || off > b->bytes || n > b->bytes - off)
        return fail(ctx, HC_ERR_ARG, "hc_download: range outside buffer");
    if (!n) return 0;
    if (!hc_host_mirror(b)) return fail(ctx, HC_ERR_NOMEM, "hc_download: no host mirror");
    hipError_t e = hipMemcpyAsync((char*)b->host + off, (char*)b->dptr + off, n,
                                  hipMemcpyDeviceToHost, ctx->stream);
    return e == hipSuccess ? 0 : hip_fail(ctx, e, "hc_download");
}

int hc_wait(hc_ctx* ctx) {
    if (!ctx) return HC_ERR_ARG;
    hipError_t e = hipStreamSynchronize(ctx->stream);
    return e == hipSuccess ? 0 : hip_fail(ctx, e, "hc_wait");
}

const char* hc_last_error(const hc_ctx* ctx) { return ctx ? ctx->last_error.c_str() : "null context"; }

// internal: the stream the context dispatches on (used by the host step loop for graph capture)
hipStream_t hc_stream_of(hc_ctx* ctx) { return ctx ? ctx->stream : nullptr; }
// internal: dispatch on another stream from now on; returns the previous one
// (the host step loop runs the mechanics step on a second stream)
hipStream_t hc_stream_swap(hc_ctx* ctx, hipStream_t s) {
    if (!ctx) return nullptr;
    hipStream_t old = ctx->stream;
    ctx->stream = s;
    return old;
}

int hc_timing_enable(hc_ctx* ctx, uint32_t mask) {
    if (!ctx) return HC_ERR_ARG;
    ctx->timing_mask = mask;
    return 0;
}

static int drain_timing(hc_ctx* ctx) {
    hipError_t e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) return hip_fail(ctx, e, "timing: synchronize");
    for (auto& p : ctx->pending) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
            ctx->total_ms[p.kernel] += ms;
            ctx->launches[p.kernel] += 1;
        }
        ctx->event_pool.push_back(p.a);
        ctx->event_pool.push_back(p.b);
    }
    ctx->pending.clear();
    return 0;
}

int hc_timing_get(hc_ctx* ctx, int kernel, double* total_ms, int64_t* launches) {
    if (!ctx || kernel < 0 || kernel >= HC_K_COUNT) return HC_ERR_ARG;
    int r = drain_timing(ctx);
    if (r) return r;
    if (total_ms) *total_ms = ctx->total_ms[kernel];
    if (launches) *launches = ctx->launches[kernel];
    return 0;
}

int hc_timing_reset(hc_ctx* ctx) {
    if (!ctx) return HC_ERR_ARG;
    int r = drain_timing(ctx);
    for (int k = 0; k < HC_K_COUNT; k++) { ctx->total_ms[k] = 0; ctx->launches[k] = 0; }
    return r;
}

}  // extern "C"
