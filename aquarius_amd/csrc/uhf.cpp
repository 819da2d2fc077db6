// Open-shell front end: unrestricted SCF and the spin-orbital integrals of
// the spin-orbital CCSD path.
//
// Reference: the localaoscf task (AOUHF<LocalUHF>, src/scf/uhf.cxx:17-409,
// options src/scf/aouhf.cxx:266-282) and the spin-orbital AO->MO task
// (src/operator/aomoints.cxx) that feeds CCSD<U> (src/cc/ccsd.cxx:92-107).
// The Coulomb and exchange builds run on the device through the contraction
// operator over the resident AO tensor; the n x n algebra stays on the host.
// Spin orbitals are dense, ordered occupied alpha, occupied beta, virtual
// alpha, virtual beta; the MO tensor (PQ|RS) comes from two half
// transformations per spin, so integrals between orbitals of different spin
// within one charge distribution are zero by construction, and each
// antisymmetrised block <PQ||RS> = (PR|QS) - (PS|QR) is cut out of it with
// slice + permuted sums.
#include "fe_util.hpp"
#include "frontend.hpp"

#include <cstring>

using namespace aqamd;
using aqamd::fe::ck;
using aqamd::fe::gemm_nn;
using aqamd::fe::gemm_tn;
using aqamd::fe::Owned;
using aqamd::fe::solve;

extern "C" {

int aq_scf_uhf(aq_ctx* ctx, const aq_molecule* mol, const aq_tensor* eri, int multiplicity,
               const aq_scf_config* cfg_in, aq_scf_result* res, double* eps_out, double* C_out, double* F_out) {
    return guarded([&] {
        AQ_CHECK(ctx && mol && eri && res, AQ_ERR_INVALID, "null argument");
        aq_scf_config cfg;
        if (cfg_in)
            cfg = *cfg_in;
        else
            aq_scf_default_config(&cfg);
        const int n = mol->d.nao;
        const int ne = mol->d.nelectrons();
        AQ_CHECK(multiplicity >= 1 && (ne + multiplicity - 1) % 2 == 0 && ne >= multiplicity - 1, AQ_ERR_INVALID,
                 "multiplicity does not fit the electron count");
        const int na = (ne + multiplicity - 1) / 2, nb = (ne - multiplicity + 1) / 2;
        AQ_CHECK(na <= n && na >= 1, AQ_ERR_INVALID, "more occupied orbitals than basis functions");
        AQ_CHECK(eri->ndim == 4 && eri->len[0] == n && eri->len[1] == n && eri->len[2] == n && eri->len[3] == n,
                 AQ_ERR_INVALID, "AO tensor must be [nao,nao,nao,nao]");
        const size_t N = size_t(n), NN = N * N;
        std::vector<double> S(NN), T(NN), V(NN), H(NN);
        fe::ints_1e(mol->d, S.data(), T.data(), V.data());
        for (size_t i = 0; i < NN; ++i) H[i] = T[i] + V[i];
        const double enuc = mol->d.nuclear_repulsion();

        std::vector<double> A = S, s, U, X(NN), tmp(NN), tmp2(NN);
        fe::eigh(n, A, s, U);
        AQ_CHECK(s[0] > 1e-10, AQ_ERR_NUMERIC, "overlap matrix is singular to working precision");
        for (size_t i = 0; i < N; ++i)
            for (size_t j = 0; j < N; ++j) {
                double x = 0.0;
                for (size_t k = 0; k < N; ++k) x += U[i * N + k] * U[j * N + k] / std::sqrt(s[k]);
                X[i * N + j] = x;
            }
        std::vector<double> Ca(NN), Cb(NN), ea(N), eb(N);
        auto diagonalise = [&](const std::vector<double>& F, std::vector<double>& C, std::vector<double>& e) {
            gemm_tn(n, X.data(), F.data(), tmp.data());
            gemm_nn(n, tmp.data(), X.data(), tmp2.data());
            std::vector<double> Cp;
            fe::eigh(n, tmp2, e, Cp);
            gemm_nn(n, X.data(), Cp.data(), C.data());
        };
        diagonalise(H, Ca, ea);
        Cb = Ca;
        eb = ea;
        auto density = [&](const std::vector<double>& C, int nocc, std::vector<double>& D) {
            for (size_t i = 0; i < N; ++i)
                for (size_t j = 0; j < N; ++j) {
                    double d = 0.0;
                    for (int k = 0; k < nocc; ++k) d += C[i * N + size_t(k)] * C[j * N + size_t(k)];
                    D[i * N + j] = d;
                }
        };
        auto commutator = [&](const std::vector<double>& F, const std::vector<double>& D, std::vector<double>& err) {
            gemm_nn(n, F.data(), D.data(), tmp.data());
            gemm_nn(n, tmp.data(), S.data(), tmp2.data());
            for (size_t i = 0; i < N; ++i)
                for (size_t j = 0; j < N; ++j) err[i * N + j] = tmp2[i * N + j] - tmp2[j * N + i];
            gemm_tn(n, X.data(), err.data(), tmp.data());
            gemm_nn(n, tmp.data(), X.data(), err.data());
        };

        Owned Dt(ctx, {n, n}), Dad(ctx, {n, n}), Dbd(ctx, {n, n}), Jd(ctx, {n, n}), Kad(ctx, {n, n}), Kbd(ctx, {n, n});
        std::vector<double> Da(NN), Db(NN), Dsum(NN), J(NN), Ka(NN), Kb(NN), Fa(NN), Fb(NN), erra(NN), errb(NN);
        std::vector<std::vector<double>> hist_a, hist_b, hist_e;
        double E = 0.0, E_old = 0.0, conv = 0.0;
        int it = 0;
        bool done = false;
        for (it = 1; it <= cfg.max_iterations; ++it) {
            density(Ca, na, Da);
            density(Cb, nb, Db);
            for (size_t i = 0; i < NN; ++i) Dsum[i] = Da[i] + Db[i];
            ck(aq_write_all(ctx, Dt.t, Dsum.data()));
            ck(aq_write_all(ctx, Dad.t, Da.data()));
            ck(aq_write_all(ctx, Dbd.t, Db.data()));
            ck(aq_contract(ctx, 1.0, eri, "mnls", Dt.t, "ls", 0.0, Jd.t, "mn"));
            ck(aq_contract(ctx, 1.0, eri, "mlns", Dad.t, "ls", 0.0, Kad.t, "mn"));
            ck(aq_contract(ctx, 1.0, eri, "mlns", Dbd.t, "ls", 0.0, Kbd.t, "mn"));
            ck(aq_read_all(ctx, Jd.t, J.data()));
            ck(aq_read_all(ctx, Kad.t, Ka.data()));
            ck(aq_read_all(ctx, Kbd.t, Kb.data()));
            E = enuc;
            for (size_t i = 0; i < NN; ++i) {
                Fa[i] = H[i] + J[i] - Ka[i];
                Fb[i] = H[i] + J[i] - Kb[i];
                E += 0.5 * (Da[i] * (H[i] + Fa[i]) + Db[i] * (H[i] + Fb[i]));   // uhf.cxx:300-310
            }
            commutator(Fa, Da, erra);
            commutator(Fb, Db, errb);
            conv = 0.0;
            for (size_t i = 0; i < NN; ++i) conv = std::max(conv, std::max(std::fabs(erra[i]), std::fabs(errb[i])));
            if (it > 1 && std::fabs(E - E_old) < cfg.convergence && conv < 100.0 * cfg.convergence) {
                done = true;
                break;
            }
            E_old = E;
            hist_a.push_back(Fa);
            hist_b.push_back(Fb);
            std::vector<double> e2(erra);
            e2.insert(e2.end(), errb.begin(), errb.end());
            hist_e.push_back(std::move(e2));
            if (int(hist_a.size()) > std::max(cfg.diis_order, 1)) {
                hist_a.erase(hist_a.begin());
                hist_b.erase(hist_b.begin());
                hist_e.erase(hist_e.begin());
            }
            const int m = int(hist_a.size());
            std::vector<double> Fxa = Fa, Fxb = Fb;
            if (m >= 2) {
                std::vector<double> B(size_t(m + 1) * size_t(m + 1), -1.0), rhs(size_t(m + 1), 0.0);
                B[size_t(m) * size_t(m + 1) + size_t(m)] = 0.0;
                for (int i = 0; i < m; ++i)
                    for (int j = 0; j < m; ++j) {
                        double b = 0.0;
                        for (size_t k = 0; k < 2 * NN; ++k) b += hist_e[size_t(i)][k] * hist_e[size_t(j)][k];
                        B[size_t(i) * size_t(m + 1) + size_t(j)] = b;
                    }
                rhs[size_t(m)] = -1.0;
                if (solve(m + 1, B, rhs)) {
                    std::fill(Fxa.begin(), Fxa.end(), 0.0);
                    std::fill(Fxb.begin(), Fxb.end(), 0.0);
                    for (int i = 0; i < m; ++i)
                        for (size_t k = 0; k < NN; ++k) {
                            Fxa[k] += rhs[size_t(i)] * hist_a[size_t(i)][k];
                            Fxb[k] += rhs[size_t(i)] * hist_b[size_t(i)][k];
                        }
                }
            }
            diagonalise(Fxa, Ca, ea);
            diagonalise(Fxb, Cb, eb);
        }
        diagonalise(Fa, Ca, ea);
        diagonalise(Fb, Cb, eb);
        res->energy = E;
        res->conv = conv;
        res->iterations = done ? it : cfg.max_iterations;
        res->converged = done ? 1 : 0;
        res->nocc = na;
        res->nocc_beta = nb;
        if (eps_out) {
            std::memcpy(eps_out, ea.data(), N * sizeof(double));
            std::memcpy(eps_out + N, eb.data(), N * sizeof(double));
        }
        if (C_out)
            for (size_t i = 0; i < N; ++i)
                for (size_t p = 0; p < N; ++p) {
                    C_out[i + p * N] = Ca[i * N + p];
                    C_out[NN + i + p * N] = Cb[i * N + p];
                }
        if (F_out) {
            std::memcpy(F_out, Fa.data(), NN * sizeof(double));
            std::memcpy(F_out + NN, Fb.data(), NN * sizeof(double));
        }
    });
}

int aq_aomo_uhf(aq_ctx* ctx, const aq_tensor* eri, int nalpha, int nbeta, const double* Ccm, const double* F,
                aq_tensor** out) {
    return guarded([&] {
        AQ_CHECK(ctx && eri && Ccm && F && out, AQ_ERR_INVALID, "null argument");
        AQ_CHECK(eri->ndim == 4, AQ_ERR_INVALID, "AO tensor must be 4-dimensional");
        const int64_t n = eri->len[0];
        AQ_CHECK(eri->len[1] == n && eri->len[2] == n && eri->len[3] == n, AQ_ERR_INVALID, "AO tensor must be [n,n,n,n]");
        AQ_CHECK(nalpha >= 1 && nbeta >= 0 && nalpha < n && nbeta < n, AQ_ERR_INVALID, "bad occupation");
        const size_t N = size_t(n), NN = N * N;
        const int64_t ns = 2 * n, o = nalpha + nbeta, v = ns - o;
        // spin-orbital columns: occ alpha, occ beta, vir alpha, vir beta
        std::vector<int> spin(static_cast<size_t>(ns)), orb(static_cast<size_t>(ns));
        {
            int64_t P = 0;
            for (int p = 0; p < nalpha; ++p) spin[size_t(P)] = 0, orb[size_t(P++)] = p;
            for (int p = 0; p < nbeta; ++p) spin[size_t(P)] = 1, orb[size_t(P++)] = p;
            for (int64_t p = nalpha; p < n; ++p) spin[size_t(P)] = 0, orb[size_t(P++)] = int(p);
            for (int64_t p = nbeta; p < n; ++p) spin[size_t(P)] = 1, orb[size_t(P++)] = int(p);
        }
        // C_sigma[mu, P] = C^sigma[mu, orb(P)] for spin(P) == sigma, else 0 (column-major [n, ns])
        std::vector<double> Cs[2] = {std::vector<double>(N * size_t(ns), 0.0), std::vector<double>(N * size_t(ns), 0.0)};
        for (int64_t P = 0; P < ns; ++P) {
            const int sg = spin[size_t(P)];
            const double* src = Ccm + size_t(sg) * NN + size_t(orb[size_t(P)]) * N;
            std::memcpy(Cs[sg].data() + size_t(P) * N, src, N * sizeof(double));
        }
        // spin-orbital Fock f[P,Q] = d(sP,sQ) C_P^T F_s C_Q (row-major ns x ns)
        std::vector<double> f(size_t(ns) * size_t(ns), 0.0);
        for (int64_t P = 0; P < ns; ++P)
            for (int64_t Q = 0; Q < ns; ++Q) {
                if (spin[size_t(P)] != spin[size_t(Q)]) continue;
                const int sg = spin[size_t(P)];
                const double* cp = Ccm + size_t(sg) * NN + size_t(orb[size_t(P)]) * N;
                const double* cq = Ccm + size_t(sg) * NN + size_t(orb[size_t(Q)]) * N;
                const double* Fs = F + size_t(sg) * NN;
                double acc = 0.0;
                for (size_t m = 0; m < N; ++m) {
                    double r = 0.0;
                    for (size_t l = 0; l < N; ++l) r += Fs[m * N + l] * cq[l];
                    acc += cp[m] * r;
                }
                f[size_t(P) * size_t(ns) + size_t(Q)] = acc;
            }
        // (PQ|RS) = sum_{sigma,tau} (C^sigma_P C^sigma_Q | C^tau_R C^tau_S)
        Owned Ca(ctx, {n, ns}), Cb(ctx, {n, ns});
        ck(aq_write_all(ctx, Ca.t, Cs[0].data()));
        ck(aq_write_all(ctx, Cb.t, Cs[1].data()));
        Owned mo(ctx, {ns, ns, ns, ns});
        {
            Owned half(ctx, {ns, ns, n, n});
            {
                Owned x(ctx, {ns, n, n, n});
                ck(aq_contract(ctx, 1.0, Ca.t, "mP", eri, "mnlk", 0.0, x.t, "Pnlk"));
                ck(aq_contract(ctx, 1.0, Ca.t, "nQ", x.t, "Pnlk", 0.0, half.t, "PQlk"));
                ck(aq_contract(ctx, 1.0, Cb.t, "mP", eri, "mnlk", 0.0, x.t, "Pnlk"));
                ck(aq_contract(ctx, 1.0, Cb.t, "nQ", x.t, "Pnlk", 1.0, half.t, "PQlk"));
            }
            Owned y(ctx, {ns, ns, ns, n});
            ck(aq_contract(ctx, 1.0, Ca.t, "lR", half.t, "PQlk", 0.0, y.t, "PQRk"));
            ck(aq_contract(ctx, 1.0, Ca.t, "kS", y.t, "PQRk", 0.0, mo.t, "PQRS"));
            ck(aq_contract(ctx, 1.0, Cb.t, "lR", half.t, "PQlk", 0.0, y.t, "PQRk"));
            ck(aq_contract(ctx, 1.0, Cb.t, "kS", y.t, "PQRk", 1.0, mo.t, "PQRS"));
        }
        ck(aq_ctx_trim(ctx));
        // Fock blocks fAI [v,o], fME [o,v], fAE [v,v], fMI [o,o]
        auto fblock = [&](int64_t r0, int64_t c0, int64_t nr, int64_t nc, Owned& dst) {
            std::vector<double> h(size_t(nr * nc));
            for (int64_t c = 0; c < nc; ++c)
                for (int64_t r = 0; r < nr; ++r) h[size_t(r + c * nr)] = f[size_t(r0 + r) * size_t(ns) + size_t(c0 + c)];
            ck(aq_tensor_create(ctx, 2, std::vector<int64_t>{nr, nc}.data(), 0, &dst.t));
            ck(aq_write_all(ctx, dst.t, h.data()));
        };
        Owned fAI, fME, fAE, fMI;
        fblock(o, 0, v, o, fAI);
        fblock(0, o, o, v, fME);
        fblock(o, o, v, v, fAE);
        fblock(0, 0, o, o, fMI);
        // <pq||rs> = (pr|qs) - (ps|qr); ranges: 0 = occupied, 1 = virtual
        auto block = [&](int r0, int r1, int r2, int r3, Owned& dst) {
            const int64_t len[2] = {o, v}, off[2] = {0, o};
            const int64_t lp = len[r0], lq = len[r1], lr = len[r2], ls = len[r3];
            ck(aq_tensor_create(ctx, 4, std::vector<int64_t>{lp, lq, lr, ls}.data(), 0, &dst.t));
            const int64_t startB[4] = {0, 0, 0, 0};
            {
                Owned box(ctx, {lp, lr, lq, ls});
                const int64_t startA[4] = {off[r0], off[r2], off[r1], off[r3]};
                const int64_t ext[4] = {lp, lr, lq, ls};
                ck(aq_slice(ctx, 1.0, mo.t, startA, 0.0, box.t, startB, ext));
                ck(aq_sum(ctx, 1.0, box.t, "prqs", 0.0, dst.t, "pqrs"));
            }
            {
                Owned box(ctx, {lp, ls, lq, lr});
                const int64_t startA[4] = {off[r0], off[r3], off[r1], off[r2]};
                const int64_t ext[4] = {lp, ls, lq, lr};
                ck(aq_slice(ctx, 1.0, mo.t, startA, 0.0, box.t, startB, ext));
                ck(aq_sum(ctx, -1.0, box.t, "psqr", 1.0, dst.t, "pqrs"));
            }
        };
        const int O_ = 0, V_ = 1;
        Owned VABIJ, VMNEF, VAMEF, VABEJ, VABEF, VMNIJ, VMNEJ, VAMIJ, VAMEI;
        block(V_, V_, O_, O_, VABIJ);
        block(O_, O_, V_, V_, VMNEF);
        block(V_, O_, V_, V_, VAMEF);
        block(V_, V_, V_, O_, VABEJ);
        block(V_, V_, V_, V_, VABEF);
        block(O_, O_, O_, O_, VMNIJ);
        block(O_, O_, V_, O_, VMNEJ);
        block(V_, O_, O_, O_, VAMIJ);
        block(V_, O_, V_, O_, VAMEI);
        ck(aq_ctx_sync(ctx));
        Owned* all[13] = {&fAI, &fME, &fAE, &fMI, &VABIJ, &VMNEF, &VAMEF, &VABEJ, &VABEF, &VMNIJ, &VMNEJ, &VAMIJ, &VAMEI};
        for (int k = 0; k < 13; ++k) out[k] = all[k]->release();
    });
}

}  // extern "C"
