// Context, device memory helpers, error reporting and per-kernel event timing.
#include <cstdio>
#include <cstring>

#include "wf_internal.h"

namespace wf {

static thread_local std::string g_error;

wf_status set_error(wf_status code, const char* fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_error = buf;
    return code;
}

void clear_error() { g_error.clear(); }

wf_status workspace(wf_context* ctx, const char* name, size_t bytes, void** out) {
    auto& b = ctx->workspace[name];
    if (b.bytes < bytes) {
        if (b.ptr) {
            WF_HIP(hipStreamSynchronize(ctx->stream));
            WF_HIP(hipFree(b.ptr));
            b.ptr = nullptr;
            b.bytes = 0;
        }
        size_t want = bytes + bytes / 8 + 256;
        hipError_t e = hipMalloc(&b.ptr, want);
        if (e != hipSuccess) {
            b.ptr = nullptr;
            return set_error(WF_ERR_OUT_OF_MEMORY, "workspace '%s': hipMalloc(%zu) failed: %s",
                             name, want, hipGetErrorString(e));
        }
        b.bytes = want;
    }
    *out = b.ptr;
    return WF_OK;
}

wf_status pinned(wf_context* ctx, size_t bytes, void** out) {
    if (ctx->pinned_bytes < bytes) {
        if (ctx->pinned) WF_HIP(hipHostFree(ctx->pinned));
        ctx->pinned = nullptr;
        size_t want = bytes < 4096 ? 4096 : bytes;
        WF_HIP(hipHostMalloc(&ctx->pinned, want, hipHostMallocDefault));
        ctx->pinned_bytes = want;
    }
    *out = ctx->pinned;
    return WF_OK;
}

wf_status fault_error() {
    return set_error(WF_ERR_HIP, "a device-wide wait of the radix sort exceeded its bound (a "
                     "tile's predecessors never published); the results of the calls since the "
                     "last synchronisation are invalid");
}

wf_status check_fault(wf_context* ctx) {
    int32_t f = 0;
    WF_TRY(readback(ctx, &f, ctx->fault, sizeof(f)));
    if (f == 0) return WF_OK;
    WF_TRY(wf_fill_u32(ctx, ctx->fault, 0u, 1));
    WF_HIP(hipStreamSynchronize(ctx->stream));
    return fault_error();
}

wf_status readback(wf_context* ctx, void* host, const void* dev, size_t bytes) {
    void* p = nullptr;
    WF_TRY(pinned(ctx, bytes, &p));
    WF_HIP(hipMemcpyAsync(p, dev, bytes, hipMemcpyDeviceToHost, ctx->stream));
    WF_HIP(hipStreamSynchronize(ctx->stream));
    std::memcpy(host, p, bytes);
    return WF_OK;
}

static hipEvent_t take_event(wf_context* ctx) {
    if (!ctx->event_pool.empty()) {
        hipEvent_t e = ctx->event_pool.back();
        ctx->event_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

Timed::Timed(wf_context* c, const char* n, double algorithmic_bytes)
    : ctx(c), name(n), bytes(algorithmic_bytes) {
    if (!ctx || !ctx->profile) return;
    if (!ctx->profile_only.empty() && ctx->profile_only != n) return;
    a = take_event(ctx);
    b = take_event(ctx);
    if (a && b) (void)hipEventRecord(a, ctx->stream);
}

Timed::Timed(wf_context* c, const char* n, double bytes_per_unit, const Count& count)
    : Timed(c, n, count.dev ? bytes_per_unit : bytes_per_unit * count.n) {
    count_dev = count.dev;
}

Timed::Timed(wf_context* c, const char* n, double bytes_per_unit, const Count& count,
             double bytes_per_unit2, const Count& count2)
    : Timed(c, n, bytes_per_unit, count) {
    bytes2 = count2.dev ? bytes_per_unit2 : bytes_per_unit2 * count2.n;
    count_dev2 = count2.dev;
}

// A pinned slot receiving the device count at dev once the stream gets there (-1: none).
static int copy_count(wf_context* ctx, const int32_t* dev) {
    if (!dev) return -1;
    if (!ctx->count_slots &&
        hipHostMalloc((void**)&ctx->count_slots, sizeof(int32_t) * kCountSlots) != hipSuccess)
        ctx->count_slots = nullptr;
    if (!ctx->count_slots) return -1;
    const int slot = ctx->next_count_slot;
    ctx->next_count_slot = (slot + 1) % kCountSlots;
    (void)hipMemcpyAsync(ctx->count_slots + slot, dev, sizeof(int32_t), hipMemcpyDeviceToHost,
                         ctx->stream);
    return slot;
}

Timed::~Timed() {
    if (!ctx || !ctx->profile || !a || !b) return;
    (void)hipEventRecord(b, ctx->stream);
    const int slot = copy_count(ctx, count_dev);
    const int slot2 = copy_count(ctx, count_dev2);
    ctx->pending.push_back(PendingTiming{name, a, b, bytes, slot, bytes2, slot2});
    if (ctx->pending.size() >= kCountSlots / 2) harvest_timings(ctx);
}

wf_status harvest_timings(wf_context* ctx) {
    if (ctx->pending.empty()) return WF_OK;
    WF_HIP(hipStreamSynchronize(ctx->stream));
    for (auto& p : ctx->pending) {
        float ms = 0.0f;
        (void)hipEventElapsedTime(&ms, p.start, p.stop);
        auto& e = ctx->prof[p.name];
        e.ms += ms;
        e.launches += 1;
        e.bytes += p.count_slot >= 0 ? p.bytes * (double)ctx->count_slots[p.count_slot] : p.bytes;
        e.bytes += p.count_slot2 >= 0 ? p.bytes2 * (double)ctx->count_slots[p.count_slot2]
                                      : p.bytes2;
        ctx->event_pool.push_back(p.start);
        ctx->event_pool.push_back(p.stop);
    }
    ctx->pending.clear();
    return WF_OK;
}

}  // namespace wf

using namespace wf;

extern "C" {

const char* wf_last_error(void) { return g_error.c_str(); }

const char* wf_version(void) { return "wavefront-hip 0.1 (gfx950)"; }

wf_status wf_context_create(int32_t device, wf_context** out) {
    clear_error();
    WF_REQUIRE(out != nullptr, "wf_context_create: out is NULL");
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess || count == 0)
        return set_error(WF_ERR_HIP, "wf_context_create: no HIP device (%s)",
                         hipGetErrorString(e));
    WF_REQUIRE(device >= 0 && device < count, "wf_context_create: device %d of %d", device,
               count);
    WF_HIP(hipSetDevice(device));
    auto* ctx = new wf_context();
    ctx->device = device;
    hipError_t se = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
    if (se != hipSuccess) {
        delete ctx;
        return set_error(WF_ERR_HIP, "hipStreamCreate failed: %s", hipGetErrorString(se));
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess) ctx->cu_count = prop.multiProcessorCount;
    hipError_t fe = hipMalloc(&ctx->fault, sizeof(int32_t));
    if (fe == hipSuccess) fe = hipMemsetAsync(ctx->fault, 0, sizeof(int32_t), ctx->stream);
    if (fe == hipSuccess) fe = hipStreamSynchronize(ctx->stream);
    if (fe != hipSuccess) {
        if (ctx->fault) (void)hipFree(ctx->fault);
        (void)hipStreamDestroy(ctx->stream);
        delete ctx;
        return set_error(WF_ERR_HIP, "wf_context_create: fault word: %s", hipGetErrorString(fe));
    }
    ctx->fault_owned = true;
    const wf_status ss = wf::fixed_scratch_init(ctx);
    if (ss != WF_OK) {
        (void)hipFree(ctx->fault);
        (void)hipStreamDestroy(ctx->stream);
        delete ctx;
        return ss;
    }
    *out = ctx;
    return WF_OK;
}

wf_status wf_context_destroy(wf_context* ctx) {
    if (!ctx) return WF_OK;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    for (auto& p : ctx->pending) {
        (void)hipEventDestroy(p.start);
        (void)hipEventDestroy(p.stop);
    }
    for (auto e : ctx->event_pool) (void)hipEventDestroy(e);
    for (auto& kv : ctx->workspace)
        if (kv.second.ptr) (void)hipFree(kv.second.ptr);
    if (ctx->pinned) (void)hipHostFree(ctx->pinned);
    if (ctx->count_slots) (void)hipHostFree(ctx->count_slots);
    if (ctx->fault_owned && ctx->fault) (void)hipFree(ctx->fault);
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return WF_OK;
}

wf_status wf_context_stream(wf_context* ctx, void** out_stream) {
    WF_REQUIRE(ctx && out_stream, "wf_context_stream: NULL argument");
    *out_stream = (void*)ctx->stream;
    return WF_OK;
}

wf_status wf_device_alloc(wf_context* ctx, size_t bytes, void** out) {
    WF_REQUIRE(ctx && out, "wf_device_alloc: NULL argument");
    *out = nullptr;
    if (bytes == 0) return WF_OK;
    hipError_t e = hipMalloc(out, bytes);
    if (e != hipSuccess) {
        *out = nullptr;
        return set_error(WF_ERR_OUT_OF_MEMORY, "wf_device_alloc(%zu): %s", bytes,
                         hipGetErrorString(e));
    }
    return WF_OK;
}

wf_status wf_device_free(wf_context* ctx, void* ptr) {
    WF_REQUIRE(ctx, "wf_device_free: NULL context");
    if (!ptr) return WF_OK;
    WF_HIP(hipStreamSynchronize(ctx->stream));
    WF_HIP(hipFree(ptr));
    return WF_OK;
}

wf_status wf_memcpy_htod(wf_context* ctx, void* dst, const void* src, size_t bytes) {
    WF_REQUIRE(ctx, "wf_memcpy_htod: NULL context");
    if (bytes == 0) return WF_OK;
    WF_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
    WF_HIP(hipStreamSynchronize(ctx->stream));
    return WF_OK;
}

wf_status wf_memcpy_dtoh(wf_context* ctx, void* dst, const void* src, size_t bytes) {
    WF_REQUIRE(ctx, "wf_memcpy_dtoh: NULL context");
    if (bytes == 0) return WF_OK;
    WF_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
    WF_HIP(hipStreamSynchronize(ctx->stream));
    return WF_OK;
}

wf_status wf_memcpy_dtod(wf_context* ctx, void* dst, const void* src, size_t bytes) {
    WF_REQUIRE(ctx, "wf_memcpy_dtod: NULL context");
    if (bytes == 0) return WF_OK;
    WF_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, ctx->stream));
    return WF_OK;
}

wf_status wf_synchronize(wf_context* ctx) {
    WF_REQUIRE(ctx, "wf_synchronize: NULL context");
    return check_fault(ctx);
}

wf_status wf_profile_enable(wf_context* ctx, int32_t enable) {
    WF_REQUIRE(ctx, "wf_profile_enable: NULL context");
    WF_TRY(harvest_timings(ctx));
    ctx->prof.clear();
    ctx->profile = enable != 0;
    return WF_OK;
}

wf_status wf_profile_query(wf_context* ctx, const char* kernel, double* total_ms,
                           int64_t* launches, double* algorithmic_bytes) {
    WF_REQUIRE(ctx && kernel, "wf_profile_query: NULL argument");
    WF_TRY(harvest_timings(ctx));
    auto it = ctx->prof.find(kernel);
    ProfEntry e;
    if (it != ctx->prof.end()) e = it->second;
    if (total_ms) *total_ms = e.ms;
    if (launches) *launches = e.launches;
    if (algorithmic_bytes) *algorithmic_bytes = e.bytes;
    return WF_OK;
}

wf_status wf_context_set_tuning(wf_context* ctx, const wf_tuning* t) {
    clear_error();
    WF_REQUIRE(ctx && t, "wf_context_set_tuning: NULL argument");
    WF_REQUIRE(t->refill_lanes >= 1 && t->refill_lanes <= 64,
               "wf_context_set_tuning: refill_lanes %d not in [1, 64]", t->refill_lanes);
    WF_REQUIRE(t->trace_blocks_per_cu >= 1 && t->trace_blocks_per_cu <= 64,
               "wf_context_set_tuning: trace_blocks_per_cu %d not in [1, 64]", t->trace_blocks_per_cu);
    WF_REQUIRE(t->bvh_builder == WF_BVH_PLOC || t->bvh_builder == WF_BVH_LBVH,
               "wf_context_set_tuning: unknown bvh_builder %d", t->bvh_builder);
    WF_REQUIRE(t->ploc_radius >= 1 && t->ploc_radius <= 32,
               "wf_context_set_tuning: ploc_radius %d not in [1, 32]", t->ploc_radius);
    WF_REQUIRE(t->bvh_width == 4 || t->bvh_width == 8,
               "wf_context_set_tuning: bvh_width %d is not 4 or 8", t->bvh_width);
    WF_REQUIRE(t->radix_bits == 6 || t->radix_bits == 8 || t->radix_bits == 11,
               "wf_context_set_tuning: radix_bits %d is not 6, 8 or 11", t->radix_bits);
    WF_REQUIRE(t->shadow_blocks_per_cu >= 0 && t->shadow_blocks_per_cu <= 64,
               "wf_context_set_tuning: shadow_blocks_per_cu %d not in [0, 64]",
               t->shadow_blocks_per_cu);
    WF_REQUIRE(t->shadow_refill_lanes >= 0 && t->shadow_refill_lanes <= 64,
               "wf_context_set_tuning: shadow_refill_lanes %d not in [0, 64]",
               t->shadow_refill_lanes);
    WF_REQUIRE(t->ray_key_bits >= 1 && t->ray_key_bits <= 33,
               "wf_context_set_tuning: ray_key_bits %d not in [1, 33]", t->ray_key_bits);
    ctx->tuning = *t;
    return WF_OK;
}

wf_status wf_context_get_tuning(wf_context* ctx, wf_tuning* t) {
    clear_error();
    WF_REQUIRE(ctx && t, "wf_context_get_tuning: NULL argument");
    *t = ctx->tuning;
    return WF_OK;
}

wf_status wf_profile_filter(wf_context* ctx, const char* kernel) {
    WF_REQUIRE(ctx, "wf_profile_filter: NULL context");
    WF_TRY(harvest_timings(ctx));
    ctx->prof.clear();
    ctx->profile_only = kernel ? kernel : "";
    return WF_OK;
}

wf_status wf_profile_names(wf_context* ctx, char* buf, int32_t cap) {
    WF_REQUIRE(ctx && buf && cap > 0, "wf_profile_names: bad argument");
    WF_TRY(harvest_timings(ctx));
    std::string all;
    for (auto& kv : ctx->prof) {
        all += kv.first;
        all += '\n';
    }
    std::strncpy(buf, all.c_str(), (size_t)cap - 1);
    buf[cap - 1] = '\0';
    return WF_OK;
}

}  // extern "C"
