// aqamd.hpp -- header-only C++ host mirror of Aquarius's tensor expression
// surface, implemented on the aqamd.h C ABI.
//
// The operators map one to one onto the reference's front end
// (src/tensor/indexable_tensor.hpp:288-346, 514-518; src/tensor/tensor.hpp:
// 151-161, 903-906):
//   C["abij"]  = k*A["..."]            -> aq_sum(k, A, "...", 0, C, "abij")
//   C["abij"] += k*A["..."]            -> aq_sum(k, A, "...", 1, C, "abij")
//   C["abij"] -= k*A["..."]            -> aq_sum(-k, A, "...", 1, C, "abij")
//   C["abij"] (=|+=|-=) k*A[..]*B[..]  -> aq_contract(+-k, A, .., B, .., 0|1, C, "abij")
//   scalar(A["ai"]*B["ai"])            -> aq_dot
//   T += Z, T = Z                      -> whole-tensor aq_sum
// Failures surface as exceptions named after the reference's TensorError
// family (tensor.hpp:834-898).
#pragma once

#include <cstdint>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "aqamd.h"

namespace aqamd {
namespace tensor {

class TensorError : public std::runtime_error {
  public:
    int code;
    TensorError(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};
class InvalidNdimError : public TensorError {
  public:
    explicit InvalidNdimError(const std::string& m) : TensorError(AQ_ERR_INVALID, m) {}
};
class OutOfMemoryError : public TensorError {
  public:
    explicit OutOfMemoryError(const std::string& m) : TensorError(AQ_ERR_OOM, m) {}
};

inline void check(int rc) {
    if (rc == AQ_OK) return;
    std::string m = aq_last_error();
    if (rc == AQ_ERR_OOM) throw OutOfMemoryError(m);
    if (m.find("invalid ndim") != std::string::npos) throw InvalidNdimError(m);
    throw TensorError(rc, m);
}

class Tensor;

struct IndexedTensor;
struct IndexedTensorMult;

// A tensor with an index string and a scalar factor (IndexedTensor,
// indexable_tensor.hpp:242-420).
struct IndexedTensor {
    Tensor* tensor_;
    std::string idx_;
    double factor_;

    IndexedTensor(Tensor& t, std::string idx, double f = 1.0) : tensor_(&t), idx_(std::move(idx)), factor_(f) {}

    IndexedTensor operator-() const { return IndexedTensor(*tensor_, idx_, -factor_); }
    IndexedTensor operator*(double f) const { return IndexedTensor(*tensor_, idx_, factor_ * f); }
    friend IndexedTensor operator*(double f, const IndexedTensor& t) { return t * f; }

    inline IndexedTensorMult operator*(const IndexedTensor& other) const;

    inline IndexedTensor& operator=(const IndexedTensor& other);
    inline IndexedTensor& operator+=(const IndexedTensor& other);
    inline IndexedTensor& operator-=(const IndexedTensor& other);
    inline IndexedTensor& operator=(const IndexedTensorMult& other);
    inline IndexedTensor& operator+=(const IndexedTensorMult& other);
    inline IndexedTensor& operator-=(const IndexedTensorMult& other);
    inline IndexedTensor& operator*=(double f);
};

// Product of two indexed tensors with a combined factor (IndexedTensorMult,
// indexable_tensor.hpp:448-512).
struct IndexedTensorMult {
    IndexedTensor A_, B_;
    double factor_;
    IndexedTensorMult(const IndexedTensor& A, const IndexedTensor& B)
        : A_(A), B_(B), factor_(A.factor_ * B.factor_) {}
    IndexedTensorMult operator-() const {
        IndexedTensorMult r(*this);
        r.factor_ = -r.factor_;
        return r;
    }
    IndexedTensorMult operator*(double f) const {
        IndexedTensorMult r(*this);
        r.factor_ *= f;
        return r;
    }
    friend IndexedTensorMult operator*(double f, const IndexedTensorMult& m) { return m * f; }
};

// Dense column-major FP64 tensor on the device (the CTFTensor role,
// ctf_tensor.hpp:16-285).  Owns its aq_tensor handle.
class Tensor {
  public:
    Tensor() = default;
    Tensor(aq_ctx* ctx, const std::vector<int64_t>& len, bool zero = true) : ctx_(ctx) {
        check(aq_tensor_create(ctx, int(len.size()), len.data(), zero ? 1 : 0, &h_));
    }
    // deep copy (ctf_tensor.cxx:40-78)
    Tensor(const Tensor& o) : ctx_(o.ctx_) {
        if (o.h_) check(aq_tensor_copy(ctx_, o.h_, &h_));
    }
    Tensor(Tensor&& o) noexcept : ctx_(o.ctx_), h_(o.h_), borrowed_(o.borrowed_) { o.h_ = nullptr; }
    // non-owning view of an existing handle
    static Tensor borrow(aq_ctx* ctx, aq_tensor* h) {
        Tensor t;
        t.ctx_ = ctx;
        t.h_ = h;
        t.borrowed_ = true;
        return t;
    }
    ~Tensor() {
        if (h_ && !borrowed_) aq_tensor_destroy(h_);
    }
    Tensor& operator=(Tensor&& o) noexcept {
        if (this != &o) {
            if (h_ && !borrowed_) aq_tensor_destroy(h_);
            ctx_ = o.ctx_;
            h_ = o.h_;
            borrowed_ = o.borrowed_;
            o.h_ = nullptr;
        }
        return *this;
    }

    aq_tensor* handle() const { return h_; }
    aq_ctx* context() const { return ctx_; }
    int ndim() const {
        int n = 0;
        check(aq_tensor_ndim(h_, &n));
        return n;
    }
    std::vector<int64_t> lengths() const {
        std::vector<int64_t> l(ndim());
        check(aq_tensor_lengths(h_, l.data()));
        return l;
    }
    // implicit index string "ABC..." (indexable_tensor.hpp:58-63)
    std::string implicit() const {
        std::string s;
        for (int k = 0; k < ndim(); ++k) s.push_back(char('A' + k));
        return s;
    }

    IndexedTensor operator[](const std::string& idx) { return IndexedTensor(*this, idx); }

    // CTFTensor methods
    void mult(double alpha, const Tensor& A, const std::string& ia, const Tensor& B, const std::string& ib,
              double beta, const std::string& ic) {
        check(aq_contract(ctx_, alpha, A.h_, ia.c_str(), B.h_, ib.c_str(), beta, h_, ic.c_str()));
    }
    void sum(double alpha, const Tensor& A, const std::string& ia, double beta, const std::string& ib) {
        check(aq_sum(ctx_, alpha, A.h_, ia.c_str(), beta, h_, ib.c_str()));
    }
    void scale(double alpha, const std::string& idx) { check(aq_scale(ctx_, h_, alpha, idx.c_str())); }
    void sum(double alpha, double beta) { check(aq_fill(ctx_, h_, alpha, beta)); }
    double dot(const Tensor& A, const std::string& ia, const std::string& ib) const {
        double r = 0.0;
        check(aq_dot(ctx_, A.h_, ia.c_str(), h_, ib.c_str(), &r));
        return r;
    }
    void weight(const std::vector<const std::vector<double>*>& d, double shift = 0.0) {
        std::vector<const double*> p;
        for (auto* v : d) p.push_back(v->data());
        check(aq_weight(ctx_, h_, int(p.size()), p.data(), shift));
    }
    double norm(int p) const {
        double r = 0.0;
        check(aq_norm(ctx_, h_, p, &r));
        return r;
    }
    std::vector<double> getAllData() const {
        int64_t n = 0;
        check(aq_tensor_size(h_, &n));
        std::vector<double> v(static_cast<size_t>(n));
        check(aq_read_all(ctx_, h_, v.data()));
        return v;
    }
    void setAllData(const std::vector<double>& v) { check(aq_write_all(ctx_, h_, v.data())); }

    // whole-tensor arithmetic (tensor.hpp:151-161)
    Tensor& operator=(const Tensor& o) {
        if (this != &o) sum(1.0, o, o.implicit(), 0.0, implicit());
        return *this;
    }
    Tensor& operator+=(const Tensor& o) {
        sum(1.0, o, o.implicit(), 1.0, implicit());
        return *this;
    }
    Tensor& operator-=(const Tensor& o) {
        sum(-1.0, o, o.implicit(), 1.0, implicit());
        return *this;
    }
    Tensor& operator*=(double f) {
        scale(f, implicit());
        return *this;
    }

  private:
    aq_ctx* ctx_ = nullptr;
    aq_tensor* h_ = nullptr;
    bool borrowed_ = false;
};

inline IndexedTensorMult IndexedTensor::operator*(const IndexedTensor& other) const {
    return IndexedTensorMult(*this, other);
}
inline IndexedTensor& IndexedTensor::operator=(const IndexedTensor& o) {
    tensor_->sum(o.factor_, *o.tensor_, o.idx_, 0.0, idx_);
    return *this;
}
inline IndexedTensor& IndexedTensor::operator+=(const IndexedTensor& o) {
    tensor_->sum(o.factor_, *o.tensor_, o.idx_, factor_, idx_);
    return *this;
}
inline IndexedTensor& IndexedTensor::operator-=(const IndexedTensor& o) {
    tensor_->sum(-o.factor_, *o.tensor_, o.idx_, factor_, idx_);
    return *this;
}
inline IndexedTensor& IndexedTensor::operator=(const IndexedTensorMult& m) {
    tensor_->mult(m.factor_, *m.A_.tensor_, m.A_.idx_, *m.B_.tensor_, m.B_.idx_, 0.0, idx_);
    return *this;
}
inline IndexedTensor& IndexedTensor::operator+=(const IndexedTensorMult& m) {
    tensor_->mult(m.factor_, *m.A_.tensor_, m.A_.idx_, *m.B_.tensor_, m.B_.idx_, factor_, idx_);
    return *this;
}
inline IndexedTensor& IndexedTensor::operator-=(const IndexedTensorMult& m) {
    tensor_->mult(-m.factor_, *m.A_.tensor_, m.A_.idx_, *m.B_.tensor_, m.B_.idx_, factor_, idx_);
    return *this;
}
inline IndexedTensor& IndexedTensor::operator*=(double f) {
    tensor_->scale(f, idx_);
    return *this;
}

// scalar(A[..]*B[..]) (indexable_tensor.hpp:514-518)
inline double scalar(const IndexedTensorMult& m) {
    return m.factor_ * m.B_.tensor_->dot(*m.A_.tensor_, m.A_.idx_, m.B_.idx_);
}

}  // namespace tensor
}  // namespace aqamd
