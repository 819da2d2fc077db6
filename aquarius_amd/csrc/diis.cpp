// C ABI of the device DIIS (DIIS<T>, src/convergence/diis.hpp:36-227): the
// extrapolator the CCSD drivers use, available on its own so that a caller
// can drive it with its own amplitude/residual tensors.
#include "diis.hpp"

using namespace aqamd;

struct aq_diis {
    aq_ctx* ctx = nullptr;
    int nx = 0, ndx = 0;
    std::vector<int64_t> sizes;   // element counts of x then dx, fixed by the first call
    std::unique_ptr<DIIS> impl;
};

int aq_diis_create(aq_ctx* ctx, int order, int start, double damping, int nx, int ndx, aq_diis** out) {
    return guarded([&] {
        AQ_CHECK(ctx && out, AQ_ERR_INVALID, "null argument");
        AQ_CHECK(order >= 0, AQ_ERR_INVALID, "DIIS order must be >= 0");
        AQ_CHECK(nx >= 1 && ndx >= 1, AQ_ERR_INVALID, "DIIS needs at least one x and one dx tensor");
        auto* d = new aq_diis();
        d->ctx = ctx;
        d->nx = nx;
        d->ndx = ndx;
        d->impl.reset(new DIIS(ctx, order, start, damping, nx, ndx));
        *out = d;
    });
}

int aq_diis_destroy(aq_diis* d) {
    return guarded([&] { delete d; });
}

int aq_diis_extrapolate(aq_diis* d, aq_tensor* const* x, aq_tensor* const* dx) {
    return guarded([&] {
        AQ_CHECK(d && x && dx, AQ_ERR_INVALID, "null argument");
        std::vector<int64_t> sizes;
        for (int i = 0; i < d->nx; ++i) {
            AQ_CHECK(x[i] && x[i]->ctx == d->ctx, AQ_ERR_INVALID, "x tensor missing or of another context");
            sizes.push_back(x[i]->size);
        }
        for (int i = 0; i < d->ndx; ++i) {
            AQ_CHECK(dx[i] && dx[i]->ctx == d->ctx, AQ_ERR_INVALID, "dx tensor missing or of another context");
            sizes.push_back(dx[i]->size);
        }
        if (d->sizes.empty()) d->sizes = sizes;
        AQ_CHECK(sizes == d->sizes, AQ_ERR_INVALID, "DIIS tensors changed size between calls");
        std::vector<Tensor> xs, dxs;
        for (int i = 0; i < d->nx; ++i) xs.push_back(Tensor::borrow(d->ctx, x[i]));
        for (int i = 0; i < d->ndx; ++i) dxs.push_back(Tensor::borrow(d->ctx, dx[i]));
        std::vector<Tensor*> px, pdx;
        for (auto& t : xs) px.push_back(&t);
        for (auto& t : dxs) pdx.push_back(&t);
        d->impl->extrapolate(px, pdx);
    });
}

int aq_diis_coefficients(const aq_diis* d, double* c, int cap, int* n, int* used_eps_retry) {
    return guarded([&] {
        AQ_CHECK(d, AQ_ERR_INVALID, "null argument");
        const auto& cf = d->impl->coefficients();
        if (n) *n = int(cf.size());
        for (int i = 0; c && i < cap && i < int(cf.size()); ++i) c[i] = cf[i];
        if (used_eps_retry) *used_eps_retry = d->impl->used_eps_retry ? 1 : 0;
    });
}
