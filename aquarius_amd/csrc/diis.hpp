// DIIS extrapolation (src/convergence/diis.hpp:16-227) over device-resident
// history, shared by the closed-shell and the spin-orbital CCSD drivers.
// Not part of the C ABI.
#pragma once

#include "internal.hpp"

#include "../../include/aqamd.hpp"

#include <algorithm>
#include <cmath>
#include <limits>
#include <memory>
#include <vector>

using aqamd::tensor::Tensor;

namespace aqamd {

inline void bump_t(const Tensor& t) { t.handle()->version += 1; }
inline double* dptr(const Tensor& t) { return t.handle()->data; }
inline int64_t tsize(const Tensor& t) { return t.handle()->size; }

// Solve the (n x n) symmetric system A x = b in place (row-major A), Gaussian
// elimination with partial pivoting.  Returns 0, or k > 0 when the k-th pivot
// is exactly zero (the dsysv 'info > 0' condition, diis.hpp:171-177).
inline int solve_small(int n, std::vector<double> A, std::vector<double>& b) {
    std::vector<int> perm(n);
    for (int i = 0; i < n; ++i) perm[i] = i;
    for (int k = 0; k < n; ++k) {
        int p = k;
        for (int i = k + 1; i < n; ++i)
            if (std::fabs(A[i * n + k]) > std::fabs(A[p * n + k])) p = i;
        if (A[p * n + k] == 0.0) return k + 1;
        if (p != k) {
            for (int j = 0; j < n; ++j) std::swap(A[k * n + j], A[p * n + j]);
            std::swap(b[k], b[p]);
        }
        for (int i = k + 1; i < n; ++i) {
            const double f = A[i * n + k] / A[k * n + k];
            if (f == 0.0) continue;
            for (int j = k; j < n; ++j) A[i * n + j] -= f * A[k * n + j];
            b[i] -= f * b[k];
        }
    }
    for (int k = n - 1; k >= 0; --k) {
        double s = b[k];
        for (int j = k + 1; j < n; ++j) s -= A[k * n + j] * b[j];
        b[k] = s / A[k * n + k];
    }
    return 0;
}

// Pulay DIIS over device-resident history (diis.hpp:36-227).
class DIIS {
  public:
    DIIS(aq_ctx* ctx, int order, int start, double damping, int nx, int ndx)
        : ctx_(ctx), nextrap_(order), start_(start), damping_(damping), nx_(nx), ndx_(ndx),
          e_((order + 1) * (order + 1), 0.0), c_(order + 1, 0.0), old_x_(order), old_dx_(order) {}

    bool used_eps_retry = false;

    void extrapolate(const std::vector<Tensor*>& x, const std::vector<Tensor*>& dx) {
        const int n = nextrap_, n1 = n + 1;
        if (n <= 1) return;
        // slot k holds iteration n-k (diis.hpp:82-85)
        std::rotate(old_x_.begin(), old_x_.end() - 1, old_x_.end());
        std::rotate(old_dx_.begin(), old_dx_.end() - 1, old_dx_.end());
        {
            std::vector<double> r(e_.size());
            for (int i = 0; i < n1; ++i)
                for (int j = 0; j < n1; ++j) r[i * n1 + j] = e_[((i + n) % n1) * n1 + (j + n) % n1];
            e_.swap(r);
            std::rotate(c_.begin(), c_.end() - 1, c_.end());
        }
        store(old_x_[0], x);    // :92-105
        store(old_dx_[0], dx);  // :107-120

        // e[i][0] = <dx, old_dx[i]> for every stored slot (:122-135)
        int nreal = 1;
        while (nreal < n && !old_dx_[nreal].empty()) ++nreal;
        std::vector<double> p(nreal, 0.0);
        // up to kDotSlots history slots per launch (the context's reduction
        // slots); each inner product is reduced on its own, so the grouping
        // does not change any value
        constexpr int kDotSlots = 8;
        for (int j = 0; j < ndx_; ++j) {
            for (int first = 0; first < nreal; first += kDotSlots) {
                const int kk = std::min(kDotSlots, nreal - first);
                std::vector<const double*> ys;
                for (int i = 0; i < kk; ++i) ys.push_back(dptr(old_dx_[first + i][j]));
                launch_multi_dot(ctx_, dptr(*dx[j]), ys.data(), kk, tsize(*dx[j]), 0);
                AQ_HIP(hipStreamSynchronize(ctx_->stream));
                for (int i = 0; i < kk; ++i) p[first + i] += ctx_->h_result[i];
            }
        }
        resolve_timing(ctx_);
        E(0, 0) = p[0];
        for (int i = 1; i < nreal; ++i) {
            E(i, 0) = p[i];
            E(0, i) = p[i];
        }
        for (int i = 0; i < nreal; ++i) {   // :141-149
            E(i, nreal) = -1.0;
            E(nreal, i) = -1.0;
            c_[i] = 0.0;
        }
        E(nreal, nreal) = 0.0;
        c_[nreal] = -1.0;

        if (nreal == 1) return;   // :151

        if (--start_ > 1) {       // :153-164
            if (damping_ > 0.0)
                for (int i = 0; i < nx_; ++i) {
                    double* a = dptr(old_x_[0][i]);
                    const double* b = dptr(old_x_[1][i]);
                    const double* xs[2] = {a, b};
                    const double cf[2] = {damping_ - 1.0, damping_};
                    launch_lincomb(ctx_, a, xs, cf, 2, tsize(old_x_[0][i]));
                    bump_t(old_x_[0][i]);
                }
            return;
        }

        const int m = nreal + 1;   // :166-188
        auto sub = [&]() {
            std::vector<double> A(m * m);
            for (int i = 0; i < m; ++i)
                for (int j = 0; j < m; ++j) A[i * m + j] = E(i, j);
            return A;
        };
        std::vector<double> b(c_.begin(), c_.begin() + m);
        std::vector<double> A = sub();
        int info = solve_small(m, A, b);
        used_eps_retry = false;
        if (info > 0) {
            const double eps = E(nreal - 1, nreal - 1) * std::numeric_limits<double>::epsilon();
            A = sub();
            for (int i = 0; i < nreal; ++i) A[i * m + i] += eps;
            b.assign(c_.begin(), c_.begin() + m);
            info = solve_small(m, A, b);
            used_eps_retry = true;
        }
        AQ_CHECK(info == 0, AQ_ERR_NUMERIC, "DIIS: Info in hesv: " + std::to_string(info));
        for (int i = 0; i < m; ++i) c_[i] = b[i];

        // x = sum_i c_i old_x[i], dx = sum_i c_i old_dx[i] (:205-226)
        for (int j = 0; j < ndx_; ++j) {
            std::vector<const double*> xs;
            for (int i = 0; i < nreal; ++i) xs.push_back(dptr(old_dx_[i][j]));
            launch_lincomb(ctx_, dptr(*dx[j]), xs.data(), c_.data(), nreal, tsize(*dx[j]));
            bump_t(*dx[j]);
        }
        for (int j = 0; j < nx_; ++j) {
            std::vector<const double*> xs;
            for (int i = 0; i < nreal; ++i) xs.push_back(dptr(old_x_[i][j]));
            launch_lincomb(ctx_, dptr(*x[j]), xs.data(), c_.data(), nreal, tsize(*x[j]));
            bump_t(*x[j]);
        }
    }

    const std::vector<double>& coefficients() const { return c_; }

  private:
    double& E(int i, int j) { return e_[i * (nextrap_ + 1) + j]; }

    void store(std::vector<Tensor>& slot, const std::vector<Tensor*>& src) {
        if (slot.empty()) {
            for (auto* t : src) slot.emplace_back(*t);   // deep copy
        } else {
            for (size_t i = 0; i < src.size(); ++i) {
                AQ_HIP(hipMemcpyAsync(dptr(slot[i]), dptr(*src[i]), size_t(tsize(*src[i])) * 8,
                                      hipMemcpyDeviceToDevice, ctx_->stream));
                bump_t(slot[i]);
            }
        }
    }

    aq_ctx* ctx_;
    int nextrap_, start_;
    double damping_;
    int nx_, ndx_;
    std::vector<double> e_, c_;
    std::vector<std::vector<Tensor>> old_x_, old_dx_;
};

}  // namespace aqamd
