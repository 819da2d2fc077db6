// Allocation-free entry point: the orbital-energy-denominator weight with its
// denominator table staged in a fixed device array instead of pooled device
// memory.  For callers that cannot allocate device memory (code under HIP
// graph capture, tight serving loops, harnesses that forbid hipMalloc).
//
// `scratch` is one file-scope device array per GPU, held until the process
// ends.  Host code never passes the array's name to a kernel or a copy (on the
// host it is only a stand-in for the symbol): aq_ctx_create resolves its
// device address once per device with hipGetSymbolAddress, with that device
// current, and every use goes through the address kept here.
#include <mutex>

#include "internal.hpp"

namespace aqamd {

namespace {

constexpr int kMaxDevices = 64;

// the denominator table of aq_weight_noalloc: sum of the tensor's lengths
// (one vector per dimension) <= AQ_NOALLOC_WEIGHT_MAX doubles
static __device__ __align__(256) double scratch[AQ_NOALLOC_WEIGHT_MAX];

struct DeviceScratch {
    std::mutex lock;          // serialises the calls that use this device's array
    std::once_flag resolved;
    double* ptr = nullptr;    // device address of `scratch` on this device
    hipError_t status = hipSuccess;
};

DeviceScratch g_scratch[kMaxDevices];

}  // namespace

// Called by aq_ctx_create with `device` current (the library's set-up call).
void noalloc_scratch_init(int device) {
    AQ_CHECK(device >= 0 && device < kMaxDevices, AQ_ERR_UNSUPPORTED, "device index beyond the scratch table");
    DeviceScratch& s = g_scratch[device];
    std::call_once(s.resolved, [&] {
        void* p = nullptr;
        s.status = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (s.status == hipSuccess) s.ptr = static_cast<double*>(p);
    });
    if (s.status != hipSuccess)
        throw Error(AQ_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(s.status));
}

}  // namespace aqamd

using namespace aqamd;

int aq_weight_noalloc(aq_ctx* ctx, aq_tensor* X, int ndim, const double* const* d, double shift) {
    return guarded([&] {
        AQ_CHECK(ctx && X, AQ_ERR_INVALID, "null argument");
        if (X->ndim == 0) return;   // ctf_tensor.cxx:484
        AQ_CHECK(ndim == X->ndim && d, AQ_ERR_INVALID, "weight needs one denominator vector per dimension");
        AQ_CHECK(X->ndim <= MAXDIM, AQ_ERR_UNSUPPORTED, "weight supports up to 8 dimensions");
        int64_t total = 0;
        int64_t off[MAXDIM];
        for (int k = 0; k < ndim; ++k) {
            AQ_CHECK(d[k] || X->len[k] == 0, AQ_ERR_INVALID, "null denominator vector");
            off[k] = total;
            total += X->len[k];
        }
        AQ_CHECK(total <= AQ_NOALLOC_WEIGHT_MAX, AQ_ERR_INVALID,
                 "aq_weight_noalloc: the tensor's lengths sum to " + std::to_string(total) +
                     " denominators, more than the fixed device table holds (" +
                     std::to_string(AQ_NOALLOC_WEIGHT_MAX) + "); use aq_weight");
        if (X->size == 0) return;
        AQ_CHECK(ctx->device >= 0 && ctx->device < 64, AQ_ERR_UNSUPPORTED, "device index beyond the scratch table");
        DeviceScratch& s = g_scratch[ctx->device];
        AQ_CHECK(s.ptr, AQ_ERR_HIP, "device scratch of aq_weight_noalloc not resolved for this device");
        std::lock_guard<std::mutex> g(s.lock);
        for (int k = 0; k < ndim; ++k)
            if (X->len[k] > 0)
                AQ_HIP(hipMemcpyAsync(s.ptr + off[k], d[k], size_t(X->len[k]) * sizeof(double),
                                      hipMemcpyHostToDevice, ctx->stream));
        launch_weight(ctx, X->data, ndim, X->len.data(), s.ptr, off, shift, /*timed=*/false);
        // the caller's host vectors and this device's table are free again
        // only once the copies and the kernel are done
        AQ_HIP(hipStreamSynchronize(ctx->stream));
        X->version += 1;
    });
}
