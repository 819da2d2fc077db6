// Host helpers of the integral front end: status checks, tensor ownership,
// small dense algebra.  Not part of the C ABI.
#pragma once

#include "internal.hpp"

#include <algorithm>
#include <cmath>
#include <initializer_list>
#include <vector>

namespace aqamd {
namespace fe {

inline void ck(int rc) {
    if (rc != AQ_OK) throw Error(rc, aq_last_error());
}

// RAII for tensors created here.
struct Owned {
    aq_tensor* t = nullptr;
    Owned() = default;
    Owned(aq_ctx* ctx, std::initializer_list<int64_t> len) {
        std::vector<int64_t> l(len);
        ck(aq_tensor_create(ctx, int(l.size()), l.data(), 0, &t));
    }
    ~Owned() {
        if (t) aq_tensor_destroy(t);
    }
    Owned(const Owned&) = delete;
    Owned& operator=(const Owned&) = delete;
    aq_tensor* release() {
        aq_tensor* r = t;
        t = nullptr;
        return r;
    }
};

// C = A^T B (n x n row-major operands).
inline void gemm_tn(int n, const double* A, const double* B, double* C) {
    std::fill(C, C + size_t(n) * size_t(n), 0.0);
    for (int k = 0; k < n; ++k)
        for (int i = 0; i < n; ++i) {
            const double a = A[size_t(k) * size_t(n) + size_t(i)];
            if (a == 0.0) continue;
            const double* b = B + size_t(k) * size_t(n);
            double* c = C + size_t(i) * size_t(n);
            for (int j = 0; j < n; ++j) c[j] += a * b[j];
        }
}

// C = A B (n x n row-major).
inline void gemm_nn(int n, const double* A, const double* B, double* C) {
    std::fill(C, C + size_t(n) * size_t(n), 0.0);
    for (int i = 0; i < n; ++i)
        for (int k = 0; k < n; ++k) {
            const double a = A[size_t(i) * size_t(n) + size_t(k)];
            if (a == 0.0) continue;
            const double* b = B + size_t(k) * size_t(n);
            double* c = C + size_t(i) * size_t(n);
            for (int j = 0; j < n; ++j) c[j] += a * b[j];
        }
}

// Dense solve with partial pivoting; returns false on an exactly zero pivot.
inline bool solve(int n, std::vector<double> A, std::vector<double>& b) {
    for (int k = 0; k < n; ++k) {
        int p = k;
        for (int i = k + 1; i < n; ++i)
            if (std::fabs(A[size_t(i) * size_t(n) + size_t(k)]) > std::fabs(A[size_t(p) * size_t(n) + size_t(k)])) p = i;
        if (A[size_t(p) * size_t(n) + size_t(k)] == 0.0) return false;
        if (p != k) {
            for (int j = 0; j < n; ++j) std::swap(A[size_t(k) * size_t(n) + size_t(j)], A[size_t(p) * size_t(n) + size_t(j)]);
            std::swap(b[size_t(k)], b[size_t(p)]);
        }
        for (int i = k + 1; i < n; ++i) {
            const double f = A[size_t(i) * size_t(n) + size_t(k)] / A[size_t(k) * size_t(n) + size_t(k)];
            for (int j = k; j < n; ++j) A[size_t(i) * size_t(n) + size_t(j)] -= f * A[size_t(k) * size_t(n) + size_t(j)];
            b[size_t(i)] -= f * b[size_t(k)];
        }
    }
    for (int i = n - 1; i >= 0; --i) {
        double s = b[size_t(i)];
        for (int j = i + 1; j < n; ++j) s -= A[size_t(i) * size_t(n) + size_t(j)] * b[size_t(j)];
        b[size_t(i)] = s / A[size_t(i) * size_t(n) + size_t(i)];
    }
    return true;
}

}  // namespace fe
}  // namespace aqamd
