// Internal runtime of the CCSD engine: context, tensors, scratch pool,
// operand cache and kernel launch interfaces.  Not part of the C ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <list>
#include <map>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/aqamd.h"

namespace aqamd {

// ---------------------------------------------------------------- errors

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

void set_last_error(const std::string& msg);

#define AQ_HIP(expr)                                                                   \
    do {                                                                               \
        hipError_t _e = (expr);                                                        \
        if (_e != hipSuccess)                                                          \
            throw ::aqamd::Error(AQ_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(_e)); \
    } while (0)

#define AQ_CHECK(cond, code, msg)                                                      \
    do {                                                                               \
        if (!(cond)) throw ::aqamd::Error((code), (msg));                              \
    } while (0)

// Run `body` translating exceptions into a status code + last-error message.
template <class F>
int guarded(F&& body) {
    try {
        body();
        return AQ_OK;
    } catch (const Error& e) {
        set_last_error(e.what());
        return e.code;
    } catch (const std::bad_alloc&) {
        set_last_error("host allocation failed");
        return AQ_ERR_OOM;
    } catch (const std::exception& e) {
        set_last_error(e.what());
        return AQ_ERR_INVALID;
    }
}

// ------------------------------------------------------------ data formats

enum Fmt : int { FMT_F64 = 0, FMT_F32 = 1, FMT_BF16X2 = 2 };  // BF16X2 = hi plane + lo plane

inline size_t fmt_bytes(Fmt f) { return f == FMT_F64 ? 8 : f == FMT_F32 ? 4 : 4; }

// ------------------------------------------------------------ timing

struct TimingClass {
    int64_t launches = 0;
    double total_ms = 0.0, flops = 0.0, bytes = 0.0;
};

struct PendingTiming {
    std::string name;
    hipEvent_t start, stop;
    double flops, bytes;
};

// ------------------------------------------------------------ context

struct CacheEntry {
    void* ptr = nullptr;
    size_t bytes = 0;
    uint64_t version = 0;
    bool valid = false;
    int pins = 0;             // > 0 while an operation is using the copy
    std::list<std::string>::iterator lru;
};

// Recorded while a context's stream is being captured into a HIP graph: the
// cached operand copies the captured work reads without having rewritten them
// earlier in the same capture, with the source tensor's version at capture
// time.  A replay is valid only while those versions are unchanged (the graph
// reads such a copy as it stands).  Copies the graph rewrites itself are
// recomputed from the current data by every replay.
struct CaptureLog {
    std::vector<std::pair<const aq_tensor*, uint64_t>> fresh_reads;
    std::vector<const void*> rewritten;
};

}  // namespace aqamd

// Tensor ids are unique in the process (not per context): a context's operand
// cache keys copies by (id, version), and a context may read tensors that
// another context created (a served job's borrowed inputs).
uint64_t new_tensor_id();

struct aq_tensor {
    aq_ctx* ctx = nullptr;
    int ndim = 0;
    std::vector<int64_t> len;
    int64_t size = 1;
    double* data = nullptr;
    uint64_t id = 0;
    uint64_t version = 0;   // bumped by every write; keys the operand cache
};

namespace aqamd {
struct GemmBatcher;
}

struct aq_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    size_t budget = 0;
    size_t tensor_bytes = 0;
    size_t peak_bytes = 0;           // high-water mark of tensors + pool + caches
    int precision = AQ_PREC_FP64;
    double min_reduced_flops = 0.0;
    int gemm_variant[3] = {-1, -1, -1};   // per format; -1 = library default
    bool autotune = false;                // pick GEMM variants per shape bucket by timing

    // scratch pool: exact-size free lists
    std::map<size_t, std::vector<void*>> pool_free;
    size_t pool_bytes = 0;       // bytes held by the pool (free + in use)

    // operand cache: key -> entry (permuted / converted copies of tensors)
    std::unordered_map<std::string, aqamd::CacheEntry> cache;
    std::list<std::string> cache_lru;   // front = most recent
    size_t cache_bytes = 0;

    // reductions
    double* d_partials = nullptr;   // device: 8 slots x 1024 partials, then 64 result slots
    double* h_result = nullptr;     // pinned host result slots (64)

    // HIP-graph capture: mem_gen counts frees of device memory (a captured
    // graph holds raw pointers and is valid only while nothing it uses was
    // freed); `capture` is non-null while the stream is being captured, and
    // then nothing may allocate
    uint64_t mem_gen = 0;
    aqamd::CaptureLog* capture = nullptr;

    // cross-context GEMM grouping (a server's slot contexts): while `batcher`
    // is set, eager GEMMs of the grouped block shapes join other contexts'
    // GEMMs in one launch on the batcher's stream; group_ready marks this
    // stream's position at a submission
    aqamd::GemmBatcher* batcher = nullptr;
    hipEvent_t group_ready = nullptr;

    // timing
    bool timing = false;
    std::vector<aqamd::PendingTiming> pending;
    std::vector<hipEvent_t> event_pool;
    std::map<std::string, aqamd::TimingClass> timing_classes;
};

namespace aqamd {

// ---- memory (context.cpp)
void* dmalloc(aq_ctx* ctx, size_t bytes);          // stream-ordered raw allocation (null on failure)
void dfree(aq_ctx* ctx, void* p);                    // stream-ordered free
void note_cache_alloc(aq_ctx* ctx);                  // peak tracking after a cache allocation
void* dev_alloc(aq_ctx* ctx, size_t bytes);        // pooled scratch
void dev_free(aq_ctx* ctx, void* p, size_t bytes);   // back to the pool
double* tensor_alloc(aq_ctx* ctx, size_t bytes);     // persistent (counted against budget)
void tensor_free(aq_ctx* ctx, double* p, size_t bytes);
size_t ctx_in_use(const aq_ctx* ctx);
void evict_cache(aq_ctx* ctx, size_t need);
void invalidate_cache(aq_ctx* ctx);                  // every cached copy must be rebuilt before use
void noalloc_scratch_init(int device);               // noalloc.cpp: resolve the fixed device array once per device

struct Scratch {  // RAII pooled device buffer
    aq_ctx* ctx;
    void* p;
    size_t n;
    Scratch(aq_ctx* c, size_t bytes) : ctx(c), p(bytes ? dev_alloc(c, bytes) : nullptr), n(bytes) {}
    ~Scratch() { if (p) dev_free(ctx, p, n); }
    Scratch(const Scratch&) = delete;
    Scratch& operator=(const Scratch&) = delete;
    template <class T> T* as() const { return static_cast<T*>(p); }
};

// ---- timing (context.cpp)
struct TimedLaunch {
    aq_ctx* ctx;
    PendingTiming rec;
    bool on;
    TimedLaunch(aq_ctx* c, const char* name, double flops, double bytes);
    ~TimedLaunch();
};
void resolve_timing(aq_ctx* ctx);   // after a synchronize

// ---- strided views (ops.cpp)
constexpr int MAXDIM = 8;

struct StridedPlan {
    // output iteration space (unique labels of the destination)
    int nd = 0;
    int64_t len[MAXDIM];
    int64_t sdst[MAXDIM];   // destination stride per output dim
    int64_t ssrc[MAXDIM];   // source stride per output dim (0 = broadcast)
    // reduced (summed) source dims
    int nr = 0;
    int64_t rlen[MAXDIM];
    int64_t rsrc[MAXDIM];
    int64_t nout = 1;
    int64_t nred = 1;
};

// ---- kernels (elementwise.hip)
void launch_fill(aq_ctx* ctx, double* x, int64_t n, double alpha, double beta);
void launch_zero_bytes(aq_ctx* ctx, void* p, size_t bytes);   // bytes % 4 == 0; a kernel, not a memset
void launch_strided_sum(aq_ctx* ctx, const StridedPlan& p, double alpha, const double* src,
                        double beta, double* dst);
void launch_strided_convert(aq_ctx* ctx, const StridedPlan& p, const double* src, void* dst, Fmt fmt,
                            int64_t plane_stride);
void launch_div(aq_ctx* ctx, const double* A, const double* B, double* C, int64_t n, double alpha, double beta);
void launch_weight(aq_ctx* ctx, double* x, int ndim, const int64_t* len, const double* d_dev,
                   const int64_t* d_off, double shift, bool timed = true);
// deterministic reductions into ctx->h_result[slot]; caller syncs
void launch_dot(aq_ctx* ctx, const double* a, const double* b, int64_t n, int slot);
void launch_norm(aq_ctx* ctx, const double* x, int64_t n, int p, int slot);
double sync_result(aq_ctx* ctx, int slot);
void launch_scatter_keys(aq_ctx* ctx, double* x, int64_t n, const int64_t* keys, const double* vals,
                         double alpha, double beta);
void launch_gather_keys(aq_ctx* ctx, const double* x, int64_t n, const int64_t* keys, double* vals);
// Placement of a GEMM's [M x N] output in C.  Unmapped: C[m + n*ldc].  Mapped:
// C[rows(m) + cols(n)] with m = d0 + ml[0]*(d1 + ml[1]*d2) (digits of the GEMM
// row labels, first fastest) and rows(m) = sum dk*ms[k]; the same for n.  Lets a
// contraction write its result straight into a permuted C (no temporary and no
// permuting pass).  M and N below 2^31.
struct CMap {
    int mapped = 0;
    int nm = 0, nn = 0;
    int64_t ml[3] = {1, 1, 1}, ms[3] = {0, 0, 0};
    int64_t nl[3] = {1, 1, 1}, ns[3] = {0, 0, 0};
};
// offset of a GEMM row (or column) index x < 2^31 through nd mapped digits
__device__ __forceinline__ int64_t cmap_digits(int nd, const int64_t* len, const int64_t* st, int64_t x) {
    uint32_t r = uint32_t(x);
    int64_t off = 0;
    for (int k = 0; k + 1 < nd; ++k) {
        const uint32_t l = uint32_t(len[k]);
        const uint32_t q = r / l;
        off += int64_t(r - q * l) * st[k];
        r = q;
    }
    return nd > 0 ? off + int64_t(r) * st[nd - 1] : 0;
}
void launch_reduce_splitk(aq_ctx* ctx, const double* P, int nsplit, int64_t m, int64_t n,
                          int64_t batch, double beta, double* C, int64_t ldc, int64_t strideC, const CMap& map);

// fused CCSD update (elementwise.hip)
struct UpdateArgs {
    int o, v;
    double* Z1; double* Z2; double* T1; double* T2; double* Tau;
    const double* VABIJSA; const double* fAI;
    const double* dA; const double* dI;
};
void launch_ccsd_update(aq_ctx* ctx, const UpdateArgs& a, int slot_e1, int slot_e2, int slot_c1,
                        int slot_c2);

// packed particle-particle ladder (elementwise.hip): V+- [P x P], [Q x Q] and
// T+- [P x O+], [Q x O-] with P = v(v+1)/2, Q = v(v-1)/2, O+ = o(o+1)/2,
// O- = o(o-1)/2; Z[v,v,o,o] += alpha (S+ + S-)
void launch_ladder_pack_v(aq_ctx* ctx, const double* V, int v, double* Vp, double* Vm);
void launch_ladder_pack_t(aq_ctx* ctx, const double* T, int v, int o, double* Tp, double* Tm);
void launch_ladder_unpack(aq_ctx* ctx, const double* Sp, const double* Sm, int v, int o, double alpha, double* Z);
// Tau, TauSA, T2SA from T1 and T2 in one pass (rhfccsd.cxx:231-238)
void launch_tau_sa(aq_ctx* ctx, const double* T1, const double* T2, int v, int o, double* Tau, double* TauSA,
                   double* T2SA);
// max |X(p,q,r,s) - X(q,p,s,r)| of X[n1,n1,n2,n2] (synchronising)
double pair_asymmetry(aq_ctx* ctx, const double* X, int n1, int n2, int slot);

// AO integrals (elementwise.hip): (pq|rs) [n,n,n,n] from the compact
// canonical shell-quartet storage (fe::CompactEri), all pointers on the device
void launch_eri_expand(aq_ctx* ctx, const double* compact, const int* shell_of, const int* fn_of, const int* nf,
                       const int64_t* pair_size, const int64_t* prefix, const int64_t* rowstart, int64_t n,
                       int64_t total, double* out);

// DIIS helpers (elementwise.hip)
void launch_multi_dot(aq_ctx* ctx, const double* x, const double* const* ys, int k, int64_t n,
                      int first_slot);
void launch_lincomb(aq_ctx* ctx, double* out, const double* const* xs, const double* c, int k,
                    int64_t n);

// ---- GEMM (gemm.hip)
struct GemmDesc {
    int64_t M, N, K, batch;
    Fmt fmt;
    const void* A; int64_t lda; bool transA; int64_t strideA; int64_t planeA;  // plane = lo offset for BF16X2
    const void* B; int64_t ldb; bool transB; int64_t strideB; int64_t planeB;
    double alpha, beta;
    double* C; int64_t ldc; int64_t strideC;
    const char* tag;   // timing class name
    CMap cmap;         // output placement (unmapped: C[m + n*ldc] + b*strideC)
};
int autotune_entries(aq_tune_entry* out, int cap);   // gemm.hip
void gemm(aq_ctx* ctx, const GemmDesc& d);

// ---- cross-context GEMM grouping (gemm.hip).  A submission waits on the host
// until every context inside a contraction stage has a GEMM of the same key
// pending, or `window_us` has passed, then one grouped kernel runs them all.
GemmBatcher* batcher_create(int device, int window_us);
void batcher_destroy(GemmBatcher* b);
void batcher_stage(GemmBatcher* b, int delta);   // a context enters (+1) or leaves (-1) a stage
void batcher_stats(GemmBatcher* b, int64_t* launches, int64_t* problems, int* max_group, int64_t* solo);

// Marks the calling context as inside a contraction stage for its batcher.
struct StageMark {
    GemmBatcher* b;
    explicit StageMark(GemmBatcher* b_) : b(b_) { if (b) batcher_stage(b, +1); }
    ~StageMark() { if (b) batcher_stage(b, -1); }
    StageMark(const StageMark&) = delete;
    StageMark& operator=(const StageMark&) = delete;
};

}  // namespace aqamd
