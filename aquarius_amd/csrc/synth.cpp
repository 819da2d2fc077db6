// Seeded synthetic closed-shell integrals (the workload construction of the
// synthetic CCSD configurations; no reference counterpart):
//   eps_occ ~ U[-1.5,-0.3] Eh and eps_vir ~ U[0.2,5.0] Eh, each sorted;
//   B^Q_pq = B^Q_qp ~ sigma N(0,1) over the o+v orbitals (occupied first),
//   sigma defaulting to 0.05*sqrt(150/(v*sqrt(nQ))) (size-normalised coupling);
//   (pq|rs) = sum_Q B^Q_pq B^Q_rs  (8-fold symmetric, positive semidefinite).
// The uniform stream is counter based (splitmix64 finaliser), normals by
// Box-Muller, so any slice can be generated independently and the host and
// any restatement draw identical uniforms.
//
// The seven Dirac blocks <pq|rs> = (pr|qs) are built on the device with the
// engine's own contractions, in the layout of rhfaomoints.cxx:58-65.
#include "internal.hpp"

#include "../../include/aqamd.hpp"

#include <algorithm>
#include <cmath>
#include <thread>

namespace aqamd {

static inline uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

static inline double u01(uint64_t key, uint64_t c) {
    const uint64_t x = mix64(key + (c + 1) * 0x9E3779B97F4A7C15ull);
    return double(x >> 11) * (1.0 / 9007199254740992.0);
}

static inline uint64_t stream_key(uint64_t seed, uint64_t stream) { return mix64(seed ^ (stream * 0xD1B54A32D192ED03ull)); }

static inline double normal(uint64_t key, uint64_t k) {
    const uint64_t j = k >> 1;
    const double u1 = u01(key, 2 * j), u2 = u01(key, 2 * j + 1);
    const double r = std::sqrt(-2.0 * std::log1p(-u1));
    const double th = 2.0 * M_PI * u2;
    return (k & 1) ? r * std::sin(th) : r * std::cos(th);
}

}  // namespace aqamd

using namespace aqamd;
using aqamd::tensor::Tensor;

extern "C" {

double aq_synth_default_sigma(int v, int nQ) {
    if (v <= 0 || nQ <= 0) return 0.05;
    return 0.05 * std::sqrt(150.0 / (double(v) * std::sqrt(double(nQ))));
}

int aq_synth_host(uint64_t seed, int o, int v, int nQ, double sigma, double* eps, double* B) {
    return guarded([&] {
        AQ_CHECK(o > 0 && v > 0 && nQ > 0, AQ_ERR_INVALID, "o, v, nQ must be positive");
        const double sg = sigma > 0.0 ? sigma : aq_synth_default_sigma(v, nQ);
        const int n = o + v;
        if (eps) {
            const uint64_t ko = stream_key(seed, 1), kv = stream_key(seed, 2);
            for (int i = 0; i < o; ++i) eps[i] = -1.5 + 1.2 * u01(ko, uint64_t(i));
            for (int a = 0; a < v; ++a) eps[o + a] = 0.2 + 4.8 * u01(kv, uint64_t(a));
            std::sort(eps, eps + o);
            std::sort(eps + o, eps + n);
        }
        if (B) {
            const uint64_t kb = stream_key(seed, 3);
            const uint64_t npair = uint64_t(n) * (n + 1) / 2;
            auto work = [&](int q0, int q1) {
                for (int Q = q0; Q < q1; ++Q) {
                    double* BQ = B + size_t(Q) * n * n;
                    for (int p = 0; p < n; ++p)
                        for (int q = p; q < n; ++q) {
                            const uint64_t t = uint64_t(p) * n - uint64_t(p) * (p - 1) / 2 + uint64_t(q - p);
                            const double z = sg * normal(kb, uint64_t(Q) * npair + t);
                            BQ[p + size_t(q) * n] = z;
                            BQ[q + size_t(p) * n] = z;
                        }
                }
            };
            const int nt = std::max(1, std::min<int>(16, int(std::thread::hardware_concurrency())));
            std::vector<std::thread> th;
            const int per = (nQ + nt - 1) / nt;
            for (int t = 0; t < nt; ++t) {
                const int q0 = t * per, q1 = std::min(nQ, q0 + per);
                if (q0 < q1) th.emplace_back(work, q0, q1);
            }
            for (auto& x : th) x.join();
        }
    });
}

int aq_synth_build(aq_ctx* ctx, int o, int v, int nQ, const double* eps, const double* B, aq_tensor** out) {
    return guarded([&] {
        AQ_CHECK(ctx && eps && B && out, AQ_ERR_INVALID, "null argument");
        AQ_CHECK(o > 0 && v > 0 && nQ > 0, AQ_ERR_INVALID, "o, v, nQ must be positive");
        const int n = o + v;
        for (int k = 0; k < 10; ++k) out[k] = nullptr;
        auto mk = [&](std::vector<int64_t> l) {
            aq_tensor* t = nullptr;
            tensor::check(aq_tensor_create(ctx, int(l.size()), l.data(), 1, &t));
            return t;
        };
        std::vector<aq_tensor*> made;
        try {
            // Fock: fAI = 0, fAB = diag(eps_vir), fIJ = diag(eps_occ)
            aq_tensor* fAI = mk({v, o});
            made.push_back(fAI);
            aq_tensor* fAB = mk({v, v});
            made.push_back(fAB);
            aq_tensor* fIJ = mk({o, o});
            made.push_back(fIJ);
            {
                std::vector<int64_t> ka(v), ki(o);
                std::vector<double> va(v), vi(o);
                for (int a = 0; a < v; ++a) { ka[a] = int64_t(a) * (v + 1); va[a] = eps[o + a]; }
                for (int i = 0; i < o; ++i) { ki[i] = int64_t(i) * (o + 1); vi[i] = eps[i]; }
                tensor::check(aq_write(ctx, fAB, v, ka.data(), va.data(), 1.0, 0.0));
                tensor::check(aq_write(ctx, fIJ, o, ki.data(), vi.data(), 1.0, 0.0));
            }
            // three-index factor blocks, Q outermost: Bvv[a,c,Q], Bvo[a,i,Q], Boo[i,j,Q]
            Tensor Bvv(ctx, {v, v, nQ}, false), Bvo(ctx, {v, o, nQ}, false), Boo(ctx, {o, o, nQ}, false);
            {
                std::vector<double> hvv(size_t(v) * v * nQ), hvo(size_t(v) * o * nQ), hoo(size_t(o) * o * nQ);
                for (int Q = 0; Q < nQ; ++Q) {
                    const double* BQ = B + size_t(Q) * n * n;   // symmetric: B[p + q*n]
                    for (int c = 0; c < v; ++c)
                        for (int a = 0; a < v; ++a) hvv[a + size_t(v) * (c + size_t(v) * Q)] = BQ[(o + a) + size_t(o + c) * n];
                    for (int i = 0; i < o; ++i)
                        for (int a = 0; a < v; ++a) hvo[a + size_t(v) * (i + size_t(o) * Q)] = BQ[(o + a) + size_t(i) * n];
                    for (int j = 0; j < o; ++j)
                        for (int i = 0; i < o; ++i) hoo[i + size_t(o) * (j + size_t(o) * Q)] = BQ[i + size_t(j) * n];
                }
                Bvv.setAllData(hvv);
                Bvo.setAllData(hvo);
                Boo.setAllData(hoo);
            }
            const int prev = ctx->precision;
            ctx->precision = AQ_PREC_FP64;
            aq_tensor* V[7];
            V[0] = mk({v, v, v, v}); made.push_back(V[0]);
            V[1] = mk({v, v, v, o}); made.push_back(V[1]);
            V[2] = mk({v, v, o, o}); made.push_back(V[2]);
            V[3] = mk({v, o, v, o}); made.push_back(V[3]);
            V[4] = mk({v, o, o, v}); made.push_back(V[4]);
            V[5] = mk({v, o, o, o}); made.push_back(V[5]);
            V[6] = mk({o, o, o, o}); made.push_back(V[6]);
            auto con = [&](aq_tensor* C, const char* ic, const Tensor& X, const char* ix, const Tensor& Y, const char* iy) {
                tensor::check(aq_contract(ctx, 1.0, X.handle(), ix, Y.handle(), iy, 0.0, C, ic));
            };
            con(V[0], "abcd", Bvv, "acQ", Bvv, "bdQ");   // (ac|bd)
            con(V[1], "abci", Bvv, "acQ", Bvo, "biQ");   // (ac|bi)
            con(V[2], "abij", Bvo, "aiQ", Bvo, "bjQ");   // (ai|bj)
            con(V[3], "aibj", Bvv, "abQ", Boo, "ijQ");   // (ab|ij)
            con(V[4], "aijb", Bvo, "ajQ", Bvo, "biQ");   // (aj|ib)
            con(V[5], "aijk", Bvo, "ajQ", Boo, "ikQ");   // (aj|ik)
            con(V[6], "ijkl", Boo, "ikQ", Boo, "jlQ");   // (ik|jl)
            ctx->precision = prev;
            AQ_HIP(hipStreamSynchronize(ctx->stream));
            aq_ctx_trim(ctx);   // drop the factor-sized scratch
            out[0] = fAI;
            out[1] = fAB;
            out[2] = fIJ;
            for (int k = 0; k < 7; ++k) out[3 + k] = V[k];
        } catch (...) {
            for (auto* t : made) aq_tensor_destroy(t);
            for (int k = 0; k < 10; ++k) out[k] = nullptr;
            throw;
        }
    });
}

}  // extern "C"
