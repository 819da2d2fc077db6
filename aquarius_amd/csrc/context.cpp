// Context: device, stream, HBM budget, scratch pool, operand cache, timing.
// Replaces the role of Aquarius's Arena (src/util/distributed.hpp:28-74): the
// place every tensor of one job lives.  One context = one HIP stream; nothing
// here is global, so contexts on one device are independent.
#include "internal.hpp"

#include <atomic>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>

namespace aqamd {

static thread_local std::string g_last_error;

void set_last_error(const std::string& msg) { g_last_error = msg; }

size_t ctx_in_use(const aq_ctx* ctx) { return ctx->tensor_bytes + ctx->pool_bytes + ctx->cache_bytes; }

// Stream-ordered device memory: allocations and frees are ordered on the
// context's stream, so one context releasing memory never synchronises the
// device under other contexts (concurrent serving).
void* dmalloc(aq_ctx* ctx, size_t bytes) {
    void* p = nullptr;
    if (hipMallocAsync(&p, bytes, ctx->stream) != hipSuccess) {
        (void)hipGetLastError();
        return nullptr;
    }
    return p;
}

void dfree(aq_ctx* ctx, void* p) {
    if (p) {
        (void)hipFreeAsync(p, ctx->stream);
        ctx->mem_gen += 1;
    }
}

// Return the memory the device's default pool holds but no allocation uses
// (blocks whose stream-ordered frees have completed, on any stream) to the
// driver.  The pool keeps freed blocks between synchronisations (release
// threshold at its maximum), which is what makes repeated allocation cheap
// within a job; across jobs of different sizes those cached blocks can be
// the wrong shapes for the next request, so they are trimmed when a job
// ends and before an allocation is retried.
static void trim_device_pool(aq_ctx* ctx) {
    hipMemPool_t pool;
    if (hipDeviceGetDefaultMemPool(&pool, ctx->device) == hipSuccess) (void)hipMemPoolTrimTo(pool, 0);
    (void)hipGetLastError();
}

static void note_peak(aq_ctx* ctx) { ctx->peak_bytes = std::max(ctx->peak_bytes, ctx_in_use(ctx)); }

static void pool_release_all(aq_ctx* ctx) {
    for (auto& kv : ctx->pool_free) {
        for (void* p : kv.second) {
            dfree(ctx, p);
            ctx->pool_bytes -= kv.first;
        }
    }
    ctx->pool_free.clear();
}

void evict_cache(aq_ctx* ctx, size_t need) {
    // least recently used first; copies pinned by the running operation stay
    auto it = ctx->cache_lru.end();
    while (it != ctx->cache_lru.begin() && ctx_in_use(ctx) + need > ctx->budget) {
        --it;
        auto ce = ctx->cache.find(*it);
        if (ce == ctx->cache.end() || ce->second.pins > 0) continue;
        dfree(ctx, ce->second.ptr);
        ctx->cache_bytes -= ce->second.bytes;
        ctx->cache.erase(ce);
        it = ctx->cache_lru.erase(it);
    }
}

void invalidate_cache(aq_ctx* ctx) {
    for (auto& kv : ctx->cache) kv.second.valid = false;
}

// AQ_ALLOC_TRACE=1: one line on stderr per allocation that needed the retry path
static bool alloc_trace() {
    static const bool on = [] {
        const char* e = getenv("AQ_ALLOC_TRACE");
        return e && e[0] == '1';
    }();
    return on;
}

static void* raw_alloc(aq_ctx* ctx, size_t bytes) {
    // under graph capture an allocation would become part of the graph;
    // the capturing caller abandons the capture and runs eagerly instead
    AQ_CHECK(!ctx->capture, AQ_ERR_UNSUPPORTED, "device allocation during graph capture");
    if (ctx_in_use(ctx) + bytes > ctx->budget) {
        pool_release_all(ctx);
        evict_cache(ctx, bytes);
    }
    AQ_CHECK(ctx_in_use(ctx) + bytes <= ctx->budget, AQ_ERR_OOM,
             "HBM budget exceeded: need " + std::to_string(bytes) + " B, in use " +
                 std::to_string(ctx_in_use(ctx)) + " of " + std::to_string(ctx->budget));
    void* p = dmalloc(ctx, bytes);
    if (!p) {
        // The device pool holds freed blocks of other sizes (and other contexts'
        // memory): hand the unused blocks back to the driver and retry before
        // giving up any cached operand copy.
        const auto t0 = std::chrono::steady_clock::now();
        pool_release_all(ctx);
        AQ_HIP(hipStreamSynchronize(ctx->stream));
        trim_device_pool(ctx);
        p = dmalloc(ctx, bytes);
        if (!p) {
            evict_cache(ctx, ctx->budget);   // drop every cached copy
            AQ_HIP(hipStreamSynchronize(ctx->stream));
            trim_device_pool(ctx);
            p = dmalloc(ctx, bytes);
        }
        if (alloc_trace())
            fprintf(stderr, "aq_alloc retry: %zu B, %s, %.1f ms\n", bytes, p ? "served" : "failed",
                    std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count());
        if (!p) throw Error(AQ_ERR_OOM, "device allocation of " + std::to_string(bytes) + " B failed");
    }
    return p;
}

static size_t round_up(size_t b) {
    const size_t g = b >= (1u << 20) ? (size_t(1) << 20) : 256;
    return (b + g - 1) / g * g;
}

void* dev_alloc(aq_ctx* ctx, size_t bytes) {
    bytes = round_up(bytes);
    auto it = ctx->pool_free.find(bytes);
    if (it != ctx->pool_free.end() && !it->second.empty()) {
        void* p = it->second.back();
        it->second.pop_back();
        return p;
    }
    void* p = raw_alloc(ctx, bytes);
    ctx->pool_bytes += bytes;
    note_peak(ctx);
    return p;
}

void dev_free(aq_ctx* ctx, void* p, size_t bytes) {
    // stream-ordered reuse: later work on the same stream may reuse p
    ctx->pool_free[round_up(bytes)].push_back(p);
}

double* tensor_alloc(aq_ctx* ctx, size_t bytes) {
    bytes = std::max<size_t>(round_up(bytes), 256);
    void* p = raw_alloc(ctx, bytes);
    ctx->tensor_bytes += bytes;
    note_peak(ctx);
    return static_cast<double*>(p);
}

void tensor_free(aq_ctx* ctx, double* p, size_t bytes) {
    bytes = std::max<size_t>(round_up(bytes), 256);
    dfree(ctx, p);
    ctx->tensor_bytes -= bytes;
}

void note_cache_alloc(aq_ctx* ctx) { note_peak(ctx); }

// ---------------------------------------------------------------- timing

TimedLaunch::TimedLaunch(aq_ctx* c, const char* name, double flops, double bytes) : ctx(c), on(c->timing) {
    if (!on) return;
    rec.name = name;
    rec.flops = flops;
    rec.bytes = bytes;
    auto take = [&]() {
        if (!ctx->event_pool.empty()) {
            hipEvent_t e = ctx->event_pool.back();
            ctx->event_pool.pop_back();
            return e;
        }
        hipEvent_t e;
        AQ_HIP(hipEventCreate(&e));
        return e;
    };
    rec.start = take();
    rec.stop = take();
    AQ_HIP(hipEventRecord(rec.start, ctx->stream));
}

TimedLaunch::~TimedLaunch() {
    if (!on) return;
    if (hipEventRecord(rec.stop, ctx->stream) == hipSuccess) ctx->pending.push_back(rec);
}

void resolve_timing(aq_ctx* ctx) {
    for (auto& p : ctx->pending) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, p.start, p.stop) == hipSuccess) {
            auto& c = ctx->timing_classes[p.name];
            c.launches += 1;
            c.total_ms += ms;
            c.flops += p.flops;
            c.bytes += p.bytes;
        }
        ctx->event_pool.push_back(p.start);
        ctx->event_pool.push_back(p.stop);
    }
    ctx->pending.clear();
}

}  // namespace aqamd

uint64_t new_tensor_id() {
    static std::atomic<uint64_t> next{1};
    return next.fetch_add(1, std::memory_order_relaxed);
}

using namespace aqamd;

extern "C" {

const char* aq_last_error(void) { return g_last_error.c_str(); }

int aq_version(void) { return 100; }

int aq_ctx_create(int device, size_t hbm_budget, aq_ctx** out) {
    return guarded([&] {
        AQ_CHECK(out, AQ_ERR_INVALID, "null output pointer");
        int ndev = 0;
        AQ_HIP(hipGetDeviceCount(&ndev));
        AQ_CHECK(device >= 0 && device < ndev, AQ_ERR_INVALID, "no such HIP device");
        AQ_HIP(hipSetDevice(device));
        noalloc_scratch_init(device);
        {
            // keep memory released by hipFreeAsync in the device pool instead of
            // returning it to the driver at every synchronisation
            hipMemPool_t pool;
            if (hipDeviceGetDefaultMemPool(&pool, device) == hipSuccess) {
                uint64_t thr = UINT64_MAX;
                (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &thr);
            }
            (void)hipGetLastError();
        }
        auto* ctx = new aq_ctx();
        ctx->device = device;
        try {
            AQ_HIP(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
            size_t free_b = 0, total_b = 0;
            AQ_HIP(hipMemGetInfo(&free_b, &total_b));
            ctx->budget = hbm_budget ? hbm_budget : static_cast<size_t>(0.9 * double(free_b));
            ctx->d_partials = static_cast<double*>(dmalloc(ctx, sizeof(double) * (8 * 1024 + 64)));
            AQ_CHECK(ctx->d_partials, AQ_ERR_OOM, "device allocation of reduction slots failed");
            AQ_HIP(hipHostMalloc(&ctx->h_result, sizeof(double) * 64, hipHostMallocDefault));
        } catch (...) {
            if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
            delete ctx;
            throw;
        }
        *out = ctx;
    });
}

int aq_ctx_destroy(aq_ctx* ctx) {
    return guarded([&] {
        if (!ctx) return;
        (void)hipSetDevice(ctx->device);
        (void)hipStreamSynchronize(ctx->stream);
        resolve_timing(ctx);
        pool_release_all(ctx);
        for (auto& kv : ctx->cache) dfree(ctx, kv.second.ptr);
        for (auto e : ctx->event_pool) (void)hipEventDestroy(e);
        if (ctx->group_ready) (void)hipEventDestroy(ctx->group_ready);
        dfree(ctx, ctx->d_partials);
        (void)hipStreamSynchronize(ctx->stream);
        (void)hipHostFree(ctx->h_result);
        (void)hipStreamDestroy(ctx->stream);
        delete ctx;
    });
}

int aq_ctx_sync(aq_ctx* ctx) {
    return guarded([&] {
        AQ_CHECK(ctx, AQ_ERR_INVALID, "null context");
        AQ_HIP(hipStreamSynchronize(ctx->stream));
        resolve_timing(ctx);
    });
}

int aq_ctx_set_precision(aq_ctx* ctx, int precision, double min_reduced_flops) {
    return guarded([&] {
        AQ_CHECK(ctx, AQ_ERR_INVALID, "null context");
        AQ_CHECK(precision == AQ_PREC_FP64 || precision == AQ_PREC_FP32 || precision == AQ_PREC_BF16X3,
                 AQ_ERR_INVALID, "unknown precision");
        ctx->precision = precision;
        ctx->min_reduced_flops = min_reduced_flops;
    });
}

int aq_ctx_set_autotune(aq_ctx* ctx, int enable) {
    return guarded([&] {
        AQ_CHECK(ctx, AQ_ERR_INVALID, "null context");
        ctx->autotune = enable != 0;
    });
}

int aq_autotune_table(aq_tune_entry* out, int cap, int* n) {
    return guarded([&] {
        AQ_CHECK(n && cap >= 0, AQ_ERR_INVALID, "null count");
        *n = autotune_entries(out, cap);
    });
}

int aq_ctx_set_gemm_variant(aq_ctx* ctx, int precision, int variant) {
    return guarded([&] {
        AQ_CHECK(ctx, AQ_ERR_INVALID, "null context");
        AQ_CHECK(precision >= 0 && precision <= 2, AQ_ERR_INVALID, "unknown precision");
        AQ_CHECK(variant >= -1 && variant <= (precision == AQ_PREC_BF16X3 ? 4 : 2), AQ_ERR_INVALID,
                 "unknown GEMM variant");
        // AQ_PREC_* and the internal format ids coincide: FP64 = 0, FP32 = 1, BF16x3 = 2
        ctx->gemm_variant[precision] = variant;
    });
}

int aq_ctx_get_precision(const aq_ctx* ctx, int* precision) {
    return guarded([&] {
        AQ_CHECK(ctx && precision, AQ_ERR_INVALID, "null argument");
        *precision = ctx->precision;
    });
}

int aq_ctx_memory(const aq_ctx* ctx, size_t* in_use, size_t* budget) {
    return guarded([&] {
        AQ_CHECK(ctx, AQ_ERR_INVALID, "null context");
        if (in_use) *in_use = ctx_in_use(ctx);
        if (budget) *budget = ctx->budget;
    });
}

int aq_ctx_peak_memory(aq_ctx* ctx, size_t* peak, int reset) {
    return guarded([&] {
        AQ_CHECK(ctx && peak, AQ_ERR_INVALID, "null argument");
        *peak = ctx->peak_bytes;
        if (reset) ctx->peak_bytes = ctx_in_use(ctx);
    });
}

int aq_ctx_trim(aq_ctx* ctx) {
    return guarded([&] {
        AQ_CHECK(ctx, AQ_ERR_INVALID, "null context");
        AQ_HIP(hipStreamSynchronize(ctx->stream));
        pool_release_all(ctx);
        for (auto& kv : ctx->cache) dfree(ctx, kv.second.ptr);
        ctx->cache.clear();
        ctx->cache_lru.clear();
        ctx->cache_bytes = 0;
        AQ_HIP(hipStreamSynchronize(ctx->stream));
        trim_device_pool(ctx);
    });
}

int aq_ctx_enable_timing(aq_ctx* ctx, int enable) {
    return guarded([&] {
        AQ_CHECK(ctx, AQ_ERR_INVALID, "null context");
        ctx->timing = enable != 0;
    });
}

int aq_ctx_timing(aq_ctx* ctx, aq_timing_record* rec, int cap, int* n) {
    return guarded([&] {
        AQ_CHECK(ctx && n, AQ_ERR_INVALID, "null argument");
        AQ_HIP(hipStreamSynchronize(ctx->stream));
        resolve_timing(ctx);
        int i = 0;
        for (auto& kv : ctx->timing_classes) {
            if (rec && i < cap) {
                std::memset(rec[i].name, 0, sizeof(rec[i].name));
                std::strncpy(rec[i].name, kv.first.c_str(), sizeof(rec[i].name) - 1);
                rec[i].launches = kv.second.launches;
                rec[i].total_ms = kv.second.total_ms;
                rec[i].flops = kv.second.flops;
                rec[i].bytes = kv.second.bytes;
            }
            ++i;
        }
        *n = i;
    });
}

int aq_ctx_timing_reset(aq_ctx* ctx) {
    return guarded([&] {
        AQ_CHECK(ctx, AQ_ERR_INVALID, "null context");
        AQ_HIP(hipStreamSynchronize(ctx->stream));
        resolve_timing(ctx);
        ctx->timing_classes.clear();
    });
}

}  // extern "C"
