// Spin-orbital CCSD on an unrestricted reference, issued through the tensor
// operator surface statement by statement as the reference's CCSD<U> task
// issues them (src/cc/ccsd.cxx:30-193; defaults as its task spec: convergence
// 1e-9, 50 iterations, DIIS order 5, start 1).
//
// The reference's spin-orbital tensors are antisymmetric, and a product
// written into an antisymmetric output is antisymmetrised over every index
// pair whose two indices come from different operands (the implicit P(pq) of
// its spin-orbital contractions, src/tensor/spinorbital_tensor.cxx:455-790).
// Here the tensors are dense over all spin orbitals (occupied alpha, occupied
// beta, virtual alpha, virtual beta), and that antisymmetriser is explicit:
// the product goes to a scratch tensor and is added to the output under each
// signed index permutation (`as_mult`).  The weight, energy, convergence and
// DIIS are the closed-shell driver's, FP64.
#include "diis.hpp"

#include <chrono>

using aqamd::tensor::Tensor;

namespace {

struct Pair {
    int p, q;
};

}  // namespace

struct aq_soccsd {
    aq_ctx* ctx = nullptr;
    aq_rhfccsd_config cfg{};
    bool ccd = false;
    int64_t o = 0, v = 0;
    Tensor fAI, fME, fAE, fMI, VABIJ, VMNEF, VAMEF, VABEJ, VABEF, VMNIJ, VMNEJ, VAMIJ, VAMEI;
    Tensor T1, T2, Z1, Z2, Tau, FME, FMI, FAE, WMNIJ, WMNEJ, WAMIJ, WAMEI;
    Tensor Xvvoo, Xoooo, Xvooo;
    std::vector<double> dA, dI;
    std::unique_ptr<aqamd::DIIS> diis;
    double mp2 = 0.0, energy = 0.0, conv = std::numeric_limits<double>::max();

    static void ck(int rc) { aqamd::tensor::check(rc); }

    // C[ic] += alpha * AS(A[ia] B[ib]): P(pq) for every antisymmetric pair of
    // C (positions in `pairs`) whose labels come from different operands.
    void as_mult(double alpha, const Tensor& A, const char* ia, const Tensor& B, const char* ib, Tensor& C,
                 const char* ic, std::initializer_list<Pair> pairs, Tensor* scratch) {
        const std::string sa(ia), sc(ic);
        std::vector<Pair> need;
        for (const Pair& pr : pairs) {
            const bool pa = sa.find(sc[size_t(pr.p)]) != std::string::npos;
            const bool qa = sa.find(sc[size_t(pr.q)]) != std::string::npos;
            if (pa != qa) need.push_back(pr);
        }
        if (need.empty()) {
            ck(aq_contract(ctx, alpha, A.handle(), ia, B.handle(), ib, 1.0, C.handle(), ic));
            return;
        }
        AQ_CHECK(scratch, AQ_ERR_INVALID, std::string("antisymmetrised product without scratch: ") + ic);
        ck(aq_contract(ctx, alpha, A.handle(), ia, B.handle(), ib, 0.0, scratch->handle(), ic));
        const int nperm = 1 << need.size();
        for (int mask = 0; mask < nperm; ++mask) {
            std::string lab = sc;
            double sign = 1.0;
            for (size_t k = 0; k < need.size(); ++k)
                if (mask & (1 << k)) {
                    std::swap(lab[size_t(need[k].p)], lab[size_t(need[k].q)]);
                    sign = -sign;
                }
            ck(aq_sum(ctx, sign, scratch->handle(), lab.c_str(), 1.0, C.handle(), ic));
        }
    }
    void plain(double alpha, const Tensor& A, const char* ia, const Tensor& B, const char* ib, Tensor& C,
               const char* ic) {
        ck(aq_contract(ctx, alpha, A.handle(), ia, B.handle(), ib, 1.0, C.handle(), ic));
    }
    void copy(const Tensor& A, const char* ia, Tensor& C, const char* ic) {
        ck(aq_sum(ctx, 1.0, A.handle(), ia, 0.0, C.handle(), ic));
    }

    // Tau["abij"] = T2["abij"] + 0.5 * T1["ai"] * T1["bj"]   (P(ab) P(ij) implicit)
    void tau() {
        copy(T2, "abij", Tau, "abij");
        as_mult(0.5, T1, "ai", T1, "bj", Tau, "abij", {{0, 1}, {2, 3}}, &Xvvoo);
    }

    int next_precision() const {
        if (cfg.precision == AQ_PREC_FP64) return AQ_PREC_FP64;
        if (cfg.refine_below > 0.0 && conv < cfg.refine_below) return AQ_PREC_FP64;
        return cfg.precision;
    }

    // ccsd.cxx:45-66
    void setup() {
        auto mk = [&](std::initializer_list<int64_t> l) { return Tensor(ctx, std::vector<int64_t>(l), true); };
        T1 = mk({v, o});
        T2 = mk({v, v, o, o});
        Z1 = mk({v, o});
        Z2 = mk({v, v, o, o});
        Tau = mk({v, v, o, o});
        FME = mk({o, v});
        FMI = mk({o, o});
        FAE = mk({v, v});
        WMNIJ = mk({o, o, o, o});
        WMNEJ = mk({o, o, v, o});
        WAMIJ = mk({v, o, o, o});
        WAMEI = mk({v, o, v, o});
        Xvvoo = mk({v, v, o, o});
        Xoooo = mk({o, o, o, o});
        Xvooo = mk({v, o, o, o});
        // Denominator: dA = -f_aa, dI = +f_ii
        {
            std::vector<int64_t> ka(static_cast<size_t>(v)), ki(static_cast<size_t>(o));
            for (int64_t a = 0; a < v; ++a) ka[size_t(a)] = a + a * v;
            for (int64_t i = 0; i < o; ++i) ki[size_t(i)] = i + i * o;
            dA.assign(size_t(v), 0.0);
            dI.assign(size_t(o), 0.0);
            ck(aq_read(ctx, fAE.handle(), v, ka.data(), dA.data()));
            ck(aq_read(ctx, fMI.handle(), o, ki.data(), dI.data()));
            for (auto& x : dA) x = -x;
        }
        copy(fAI, "ai", T1, "ai");                          // :57
        if (ccd) ck(aq_fill(ctx, T1.handle(), 0.0, 0.0));
        copy(VABIJ, "abij", T2, "abij");                    // :58
        weight1(T1);
        weight2(T2);                                        // :60
        tau();                                              // :62-63
        double e1 = 0.0, e2 = 0.0;
        ck(aq_dot(ctx, fAI.handle(), "ai", T1.handle(), "ai", &e1));
        ck(aq_dot(ctx, VABIJ.handle(), "abij", Tau.handle(), "abij", &e2));
        mp2 = e1 + 0.25 * e2;                               // :65
        diis.reset(new aqamd::DIIS(ctx, cfg.diis_order, cfg.diis_start, cfg.diis_damping, 2, 2));
    }
    void weight1(Tensor& X) {
        const double* d[2] = {dA.data(), dI.data()};
        ck(aq_weight(ctx, X.handle(), 2, d, 0.0));
    }
    void weight2(Tensor& X) {
        const double* d[4] = {dA.data(), dA.data(), dI.data(), dI.data()};
        ck(aq_weight(ctx, X.handle(), 4, d, 0.0));
    }

    // ccsd.cxx:92-187
    void iterate() {
        ck(aq_ctx_set_precision(ctx, next_precision(), cfg.min_reduced_flops));
        tau();                                                                        // :114-115

        copy(fME, "me", FME, "me");                                                   // :121
        plain(1.0, VMNEF, "mnef", T1, "fn", FME, "me");                               // :122
        copy(fMI, "mi", FMI, "mi");                                                   // :124
        plain(0.5, VMNEF, "mnef", T2, "efin", FMI, "mi");                             // :125
        plain(1.0, FME, "me", T1, "ei", FMI, "mi");                                   // :126
        plain(1.0, VMNEJ, "nmfi", T1, "fn", FMI, "mi");                               // :127
        copy(fAE, "ae", FAE, "ae");                                                   // :129
        plain(-0.5, VMNEF, "mnef", T2, "afmn", FAE, "ae");                            // :130
        plain(-1.0, FME, "me", T1, "am", FAE, "ae");                                  // :131
        plain(1.0, VAMEF, "amef", T1, "fm", FAE, "ae");                               // :132

        copy(VMNIJ, "mnij", WMNIJ, "mnij");                                           // :134
        as_mult(0.5, VMNEF, "mnef", Tau, "efij", WMNIJ, "mnij", {{0, 1}, {2, 3}}, &Xoooo);   // :135
        as_mult(1.0, VMNEJ, "mnej", T1, "ei", WMNIJ, "mnij", {{0, 1}, {2, 3}}, &Xoooo);      // :136
        copy(VMNEJ, "mnej", WMNEJ, "mnej");                                           // :138
        as_mult(1.0, VMNEF, "mnef", T1, "fj", WMNEJ, "mnej", {{0, 1}}, nullptr);      // :139
        copy(VAMIJ, "amij", WAMIJ, "amij");                                           // :141
        as_mult(0.5, VAMEF, "amef", Tau, "efij", WAMIJ, "amij", {{2, 3}}, nullptr);   // :142
        as_mult(1.0, VAMEI, "amej", T1, "ei", WAMIJ, "amij", {{2, 3}}, &Xvooo);       // :143
        copy(VAMEI, "amei", WAMEI, "amei");                                           // :145
        plain(0.5, VMNEF, "mnef", T2, "afni", WAMEI, "amei");                         // :146
        plain(1.0, VAMEF, "amef", T1, "fi", WAMEI, "amei");                           // :147
        plain(-1.0, WMNEJ, "nmei", T1, "an", WAMEI, "amei");                          // :148

        copy(fAI, "ai", Z1, "ai");                                                    // :157
        plain(1.0, fAE, "ae", T1, "ei", Z1, "ai");                                    // :158
        plain(-1.0, FMI, "mi", T1, "am", Z1, "ai");                                   // :159
        plain(-1.0, VAMEI, "amei", T1, "em", Z1, "ai");                               // :160
        plain(1.0, FME, "me", T2, "aeim", Z1, "ai");                                  // :161
        plain(0.5, VAMEF, "amef", Tau, "efim", Z1, "ai");                             // :162
        plain(-0.5, WMNEJ, "mnei", T2, "eamn", Z1, "ai");                             // :163

        copy(VABIJ, "abij", Z2, "abij");                                              // :165
        as_mult(1.0, VABEJ, "abej", T1, "ei", Z2, "abij", {{0, 1}, {2, 3}}, &Xvvoo);  // :166
        as_mult(-1.0, WAMIJ, "amij", T1, "bm", Z2, "abij", {{0, 1}, {2, 3}}, &Xvvoo); // :167
        as_mult(1.0, FAE, "ae", T2, "ebij", Z2, "abij", {{0, 1}, {2, 3}}, &Xvvoo);    // :168
        as_mult(-1.0, FMI, "mi", T2, "abmj", Z2, "abij", {{0, 1}, {2, 3}}, &Xvvoo);   // :169
        as_mult(0.5, VABEF, "abef", Tau, "efij", Z2, "abij", {{0, 1}, {2, 3}}, &Xvvoo);   // :170
        as_mult(0.5, WMNIJ, "mnij", Tau, "abmn", Z2, "abij", {{0, 1}, {2, 3}}, &Xvvoo);   // :171
        as_mult(1.0, WAMEI, "amei", T2, "ebjm", Z2, "abij", {{0, 1}, {2, 3}}, &Xvvoo);    // :172

        ck(aq_ctx_set_precision(ctx, AQ_PREC_FP64, 0.0));
        weight1(Z1);                                                                  // :178
        weight2(Z2);
        if (ccd) ck(aq_fill(ctx, Z1.handle(), 0.0, 0.0));
        ck(aq_sum(ctx, 1.0, Z1.handle(), "ai", 1.0, T1.handle(), "ai"));              // :179
        ck(aq_sum(ctx, 1.0, Z2.handle(), "abij", 1.0, T2.handle(), "abij"));
        tau();                                                                        // :181-182
        double e1 = 0.0, e2 = 0.0, n1 = 0.0, n2 = 0.0;
        ck(aq_dot(ctx, fAI.handle(), "ai", T1.handle(), "ai", &e1));
        ck(aq_dot(ctx, VABIJ.handle(), "abij", Tau.handle(), "abij", &e2));
        energy = e1 + 0.25 * e2;                                                      // :183
        ck(aq_norm(ctx, Z1.handle(), 0, &n1));
        ck(aq_norm(ctx, Z2.handle(), 0, &n2));
        conv = std::max(n1, n2);                                                      // :184
        diis->extrapolate({&T1, &T2}, {&Z1, &Z2});                                    // :186
    }
};

using namespace aqamd;

extern "C" {

int aq_soccsd_create(aq_ctx* ctx, const aq_tensor* const* in, const aq_rhfccsd_config* cfg, int ccd,
                     aq_soccsd** out) {
    return guarded([&] {
        AQ_CHECK(ctx && in && out, AQ_ERR_INVALID, "null argument");
        for (int k = 0; k < 13; ++k) AQ_CHECK(in[k], AQ_ERR_INVALID, "null input tensor");
        AQ_CHECK(in[0]->ndim == 2, AQ_ERR_INVALID, "fAI must be [v,o]");
        const int64_t v = in[0]->len[0], o = in[0]->len[1];
        const std::vector<std::vector<int64_t>> shapes = {
            {v, o}, {o, v}, {v, v}, {o, o}, {v, v, o, o}, {o, o, v, v}, {v, o, v, v},
            {v, v, v, o}, {v, v, v, v}, {o, o, o, o}, {o, o, v, o}, {v, o, o, o}, {v, o, v, o}};
        const char* names[13] = {"fAI", "fME", "fAE", "fMI", "VABIJ", "VMNEF", "VAMEF",
                                 "VABEJ", "VABEF", "VMNIJ", "VMNEJ", "VAMIJ", "VAMEI"};
        for (int k = 0; k < 13; ++k)
            AQ_CHECK(in[k]->len == shapes[size_t(k)], AQ_ERR_INVALID, std::string("invalid lengths of ") + names[k]);
        auto* cc = new aq_soccsd();
        try {
            cc->ctx = ctx;
            if (cfg)
                cc->cfg = *cfg;
            else
                aq_rhfccsd_default_config(&cc->cfg);
            cc->ccd = ccd != 0;
            cc->o = o;
            cc->v = v;
            Tensor* dst[13] = {&cc->fAI, &cc->fME, &cc->fAE, &cc->fMI, &cc->VABIJ, &cc->VMNEF, &cc->VAMEF,
                               &cc->VABEJ, &cc->VABEF, &cc->VMNIJ, &cc->VMNEJ, &cc->VAMIJ, &cc->VAMEI};
            for (int k = 0; k < 13; ++k) *dst[k] = Tensor::borrow(ctx, const_cast<aq_tensor*>(in[k]));
            cc->setup();
        } catch (...) {
            delete cc;
            throw;
        }
        *out = cc;
    });
}

int aq_soccsd_destroy(aq_soccsd* cc) {
    return guarded([&] { delete cc; });
}

int aq_soccsd_mp2(const aq_soccsd* cc, double* e) {
    return guarded([&] {
        AQ_CHECK(cc && e, AQ_ERR_INVALID, "null argument");
        *e = cc->mp2;
    });
}

int aq_soccsd_iterate(aq_soccsd* cc, double* energy, double* conv) {
    return guarded([&] {
        AQ_CHECK(cc, AQ_ERR_INVALID, "null solver");
        cc->iterate();
        if (energy) *energy = cc->energy;
        if (conv) *conv = cc->conv;
    });
}

int aq_soccsd_run(aq_soccsd* cc, aq_iter_callback cb, void* user, double* energy, double* conv, int* iterations,
                  int* converged) {
    return guarded([&] {
        AQ_CHECK(cc, AQ_ERR_INVALID, "null solver");
        cc->conv = std::numeric_limits<double>::max();   // iterative.hpp:130
        int it = 1;
        for (; it <= cc->cfg.max_iterations && !(cc->conv < cc->cfg.convergence); ++it) {
            const int prec = cc->next_precision();
            auto t0 = std::chrono::steady_clock::now();
            cc->iterate();
            const double dt = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
            if (cb) cb(it, cc->energy, cc->conv, dt, prec, user);
        }
        if (energy) *energy = cc->energy;
        if (conv) *conv = cc->conv;
        if (iterations) *iterations = it - 1;
        if (converged) *converged = cc->conv < cc->cfg.convergence ? 1 : 0;
    });
}

int aq_soccsd_tensor(const aq_soccsd* cc, int which, const aq_tensor** out) {
    return guarded([&] {
        AQ_CHECK(cc && out, AQ_ERR_INVALID, "null argument");
        const Tensor* t = nullptr;
        switch (which) {
            case 0: t = &cc->T1; break;
            case 1: t = &cc->T2; break;
            case 2: t = &cc->Z1; break;
            case 3: t = &cc->Z2; break;
            case 4: t = &cc->Tau; break;
            default: throw Error(AQ_ERR_INVALID, "unknown tensor selector");
        }
        *out = t->handle();
    });
}

}  // extern "C"
