// Closed-shell SCF and the AO->MO step, plus the C ABI of the integral front
// end.
//
// Reference: the localaoscf task (AOUHF<LocalUHF>, src/scf/aouhf.cxx and
// src/scf/uhf.cxx:17-409; options aouhf.cxx:266-282) and the RHF AO->MO task
// (src/operator/rhfaomoints.cxx:38-180).  A closed-shell UHF with equal alpha
// and beta orbitals is RHF, which is what is solved here.  The Coulomb and
// exchange builds J = (mn|ls) D_ls and K = (ml|ns) D_ls run on the device
// through the contraction operator over the resident AO tensor; the n x n
// algebra (orthogonalisation, DIIS, diagonalisation) stays on the host.  The
// four quarter transformations of the AO->MO step are device contractions,
// and the seven Dirac blocks <pq|rs> = (pr|qs) are cut from the MO tensor
// with slice + permuted sum.
#include "frontend.hpp"
#include "internal.hpp"

#include <algorithm>
#include <limits>
#include <cmath>
#include <cstring>
#include <numeric>

using namespace aqamd;

#include "fe_util.hpp"

using aqamd::fe::ck;
using aqamd::fe::Owned;
using aqamd::fe::gemm_nn;
using aqamd::fe::gemm_tn;
using aqamd::fe::solve;

namespace aqamd {
namespace fe {

// Symmetric eigensolver: Householder reduction to tridiagonal form, then the
// implicit QL iteration with Wilkinson shifts (Bowdler, Martin, Reinsch,
// Wilkinson, Numer. Math. 11, 293 (1968)).  The orthogonal factor is kept as
// Q^T (row k = k-th column of Q), so both the reflections and the plane
// rotations update contiguous rows.  A is n x n symmetric row-major (destroyed);
// w receives the eigenvalues ascending, U the eigenvectors as the columns of a
// row-major matrix.
void eigh(int n, std::vector<double>& A, std::vector<double>& w, std::vector<double>& U) {
    const size_t N = size_t(n);
    std::vector<double> QT(N * N, 0.0);
    for (size_t i = 0; i < N; ++i) QT[i * N + i] = 1.0;
    auto a = [&](size_t i, size_t j) -> double& { return A[i * N + j]; };
    std::vector<double> v(N), pv(N), u(N);

    // A <- H A H with H = I - beta v v^T acting on rows/columns k+1..n-1
    for (size_t k = 0; k + 2 < N; ++k) {
        const size_t m = N - k - 1;   // length of the reflected part
        double xn = 0.0;
        for (size_t t = 0; t < m; ++t) xn += a(k + 1 + t, k) * a(k + 1 + t, k);
        xn = std::sqrt(xn);
        if (xn == 0.0) continue;
        const double alpha = a(k + 1, k) > 0.0 ? -xn : xn;
        for (size_t t = 0; t < m; ++t) v[t] = a(k + 1 + t, k);
        v[0] -= alpha;
        double vv = 0.0;
        for (size_t t = 0; t < m; ++t) vv += v[t] * v[t];
        if (vv == 0.0) continue;
        const double beta = 2.0 / vv;
        // p = beta A_sub v, K = beta (v.p) / 2, q = p - K v, A_sub -= v q^T + q v^T
        double vp = 0.0;
        for (size_t r = 0; r < m; ++r) {
            const double* row = &A[(k + 1 + r) * N + k + 1];
            double acc = 0.0;
            for (size_t t = 0; t < m; ++t) acc += row[t] * v[t];
            pv[r] = beta * acc;
            vp += v[r] * pv[r];
        }
        const double K = 0.5 * beta * vp;
        for (size_t r = 0; r < m; ++r) pv[r] -= K * v[r];
        for (size_t r = 0; r < m; ++r) {
            double* row = &A[(k + 1 + r) * N + k + 1];
            const double vr = v[r], qr = pv[r];
            for (size_t t = 0; t < m; ++t) row[t] -= vr * pv[t] + qr * v[t];
        }
        a(k + 1, k) = a(k, k + 1) = alpha;
        for (size_t t = 1; t < m; ++t) a(k + 1 + t, k) = a(k, k + 1 + t) = 0.0;
        // Q <- Q H, i.e. rows k+1.. of Q^T: QT_sub -= beta v (v^T QT_sub)
        std::fill(u.begin(), u.end(), 0.0);
        for (size_t t = 0; t < m; ++t) {
            const double* row = &QT[(k + 1 + t) * N];
            const double vt = v[t];
            for (size_t c = 0; c < N; ++c) u[c] += vt * row[c];
        }
        for (size_t t = 0; t < m; ++t) {
            double* row = &QT[(k + 1 + t) * N];
            const double f = beta * v[t];
            for (size_t c = 0; c < N; ++c) row[c] -= f * u[c];
        }
    }

    // implicit QL on the tridiagonal (d, e): e[i] couples i and i+1
    std::vector<double> d(N), e(N, 0.0);
    for (size_t i = 0; i < N; ++i) d[i] = a(i, i);
    for (size_t i = 0; i + 1 < N; ++i) e[i] = a(i + 1, i);
    const double eps = std::numeric_limits<double>::epsilon();
    auto rotate = [&](size_t i, double c, double sn) {   // rows i, i+1 of Q^T
        double* r0 = &QT[i * N];
        double* r1 = &QT[(i + 1) * N];
        for (size_t c2 = 0; c2 < N; ++c2) {
            const double f = r1[c2];
            r1[c2] = sn * r0[c2] + c * f;
            r0[c2] = c * r0[c2] - sn * f;
        }
    };
    for (size_t l = 0; l < N; ++l) {
        int iter = 0;
        for (;;) {
            size_t m = l;
            for (; m + 1 < N; ++m) {
                const double dd = std::fabs(d[m]) + std::fabs(d[m + 1]);
                if (std::fabs(e[m]) <= eps * dd) break;
            }
            if (m == l) break;
            AQ_CHECK(++iter <= 60, AQ_ERR_INVALID, "eigensolver did not converge");
            double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
            double r = std::hypot(g, 1.0);
            g = d[m] - d[l] + e[l] / (g + (g >= 0.0 ? r : -r));
            double sn = 1.0, c = 1.0, p = 0.0;
            bool early = false;
            for (size_t i = m; i-- > l;) {
                const double f = sn * e[i], b = c * e[i];
                r = std::hypot(f, g);
                e[i + 1] = r;
                if (r == 0.0) {   // split: the rest decouples
                    d[i + 1] -= p;
                    e[m] = 0.0;
                    early = true;
                    break;
                }
                sn = f / r;
                c = g / r;
                g = d[i + 1] - p;
                r = (d[i] - g) * sn + 2.0 * c * b;
                p = sn * r;
                d[i + 1] = g + p;
                g = c * r - b;
                rotate(i, c, sn);
            }
            if (early) continue;
            d[l] -= p;
            e[l] = g;
            e[m] = 0.0;
        }
    }

    std::vector<size_t> order(N);
    std::iota(order.begin(), order.end(), size_t(0));
    std::stable_sort(order.begin(), order.end(), [&](size_t x, size_t y) { return d[x] < d[y]; });
    w.resize(N);
    U.assign(N * N, 0.0);
    for (size_t k = 0; k < N; ++k) {
        w[k] = d[order[k]];
        const double* col = &QT[order[k] * N];
        for (size_t i = 0; i < N; ++i) U[i * N + k] = col[i];
    }
}

}  // namespace fe
}  // namespace aqamd

namespace {

// The ten CCSD inputs from a full MO tensor mo[p,q,r,s] = (pq|rs) and the MO
// Fock matrix f (row-major n x n): fAI, fAB, fIJ and the seven Dirac blocks
// <pq|rs> = (pr|qs), each cut from mo as the box [p-range, r-range,
// q-range, s-range] and permuted into [p, q, r, s].
void rhf_blocks(aq_ctx* ctx, const aq_tensor* mo, int nocc, const std::vector<double>& f, aq_tensor** out) {
    const int64_t n = mo->len[0];
    const int64_t o = nocc, v = n - nocc;
    const size_t N = size_t(n);
    Owned fAI(ctx, {v, o}), fAB(ctx, {v, v}), fIJ(ctx, {o, o});
    {
        std::vector<double> h(size_t(v * o));
        for (int64_t i = 0; i < o; ++i)
            for (int64_t a = 0; a < v; ++a) h[size_t(a + i * v)] = f[size_t(o + a) * N + size_t(i)];
        ck(aq_write_all(ctx, fAI.t, h.data()));
        h.assign(size_t(v * v), 0.0);
        for (int64_t b = 0; b < v; ++b)
            for (int64_t a = 0; a < v; ++a) h[size_t(a + b * v)] = f[size_t(o + a) * N + size_t(o + b)];
        ck(aq_write_all(ctx, fAB.t, h.data()));
        h.assign(size_t(o * o), 0.0);
        for (int64_t j = 0; j < o; ++j)
            for (int64_t i = 0; i < o; ++i) h[size_t(i + j * o)] = f[size_t(i) * N + size_t(j)];
        ck(aq_write_all(ctx, fIJ.t, h.data()));
    }
    auto block = [&](int r0, int r1, int r2, int r3, Owned& dst) {
        const int64_t lp = r0 == 0 ? o : v, lq = r1 == 0 ? o : v, lr = r2 == 0 ? o : v, ls = r3 == 0 ? o : v;
        const int64_t off[2] = {0, o};
        Owned box(ctx, {lp, lr, lq, ls});
        const int64_t startA[4] = {off[r0 != 0], off[r2 != 0], off[r1 != 0], off[r3 != 0]};
        const int64_t startB[4] = {0, 0, 0, 0};
        const int64_t ext[4] = {lp, lr, lq, ls};
        ck(aq_slice(ctx, 1.0, mo, startA, 0.0, box.t, startB, ext));
        ck(aq_tensor_create(ctx, 4, std::vector<int64_t>{lp, lq, lr, ls}.data(), 0, &dst.t));
        ck(aq_sum(ctx, 1.0, box.t, "prqs", 0.0, dst.t, "pqrs"));
    };
    const int V_ = 1, O_ = 0;
    Owned VABCD, VABCI, VABIJ, VAIBJ, VAIJB, VAIJK, VIJKL;
    block(V_, V_, V_, V_, VABCD);
    block(V_, V_, V_, O_, VABCI);
    block(V_, V_, O_, O_, VABIJ);
    block(V_, O_, V_, O_, VAIBJ);
    block(V_, O_, O_, V_, VAIJB);
    block(V_, O_, O_, O_, VAIJK);
    block(O_, O_, O_, O_, VIJKL);
    ck(aq_ctx_sync(ctx));
    Owned* all[10] = {&fAI, &fAB, &fIJ, &VABCD, &VABCI, &VABIJ, &VAIBJ, &VAIJB, &VAIJK, &VIJKL};
    for (int k = 0; k < 10; ++k) out[k] = all[k]->release();
}

}  // namespace

extern "C" {

int aq_molecule_create(int natoms, const int* Z, const double* xyz, const char* genbas, int spherical, int charge,
                       aq_molecule** out) {
    return guarded([&] {
        AQ_CHECK(out && Z && xyz && genbas && natoms >= 1, AQ_ERR_INVALID, "null argument or no atoms");
        auto basis = fe::parse_genbas(genbas);
        auto* m = new aq_molecule();
        try {
            m->d.Z.assign(Z, Z + natoms);
            m->d.xyz.assign(xyz, xyz + 3 * size_t(natoms));
            m->d.charge = charge;
            m->d.spherical = spherical != 0;
            fe::build_shells(m->d, basis);
        } catch (...) {
            delete m;
            throw;
        }
        *out = m;
    });
}

int aq_molecule_destroy(aq_molecule* mol) {
    return guarded([&] { delete mol; });
}

int aq_molecule_info(const aq_molecule* mol, int* nao, int* nelectrons, int* nshells, double* nuclear_repulsion) {
    return guarded([&] {
        AQ_CHECK(mol, AQ_ERR_INVALID, "null molecule");
        if (nao) *nao = mol->d.nao;
        if (nelectrons) *nelectrons = mol->d.nelectrons();
        if (nshells) *nshells = int(mol->d.shells.size());
        if (nuclear_repulsion) *nuclear_repulsion = mol->d.nuclear_repulsion();
    });
}

int aq_ints_1e(const aq_molecule* mol, double* S, double* T, double* V) {
    return guarded([&] {
        AQ_CHECK(mol && S && T && V, AQ_ERR_INVALID, "null argument");
        fe::ints_1e(mol->d, S, T, V);
    });
}

int aq_ints_2e(const aq_molecule* mol, int nthreads, double* eri) {
    return guarded([&] {
        AQ_CHECK(mol && eri, AQ_ERR_INVALID, "null argument");
        fe::ints_2e(mol->d, nthreads, eri);
    });
}

// The host evaluates each canonical shell quartet once into a compact pinned
// buffer (about n^4/8 values); one copy moves it to the device, where a
// gather kernel expands the 8-fold permutational symmetry into (mn|ls).
int aq_ints_2e_device(aq_ctx* ctx, const aq_molecule* mol, int nthreads, aq_tensor** out) {
    return guarded([&] {
        AQ_CHECK(ctx && mol && out, AQ_ERR_INVALID, "null argument");
        const int64_t n = mol->d.nao;
        const fe::CompactEri lay = fe::compact_layout(mol->d);
        struct Pinned {
            double* p = nullptr;
            ~Pinned() {
                if (p) (void)hipHostFree(p);
            }
        } host;
        AQ_HIP(hipHostMalloc(reinterpret_cast<void**>(&host.p), size_t(std::max<int64_t>(lay.total, 1)) * 8, 0));
        fe::ints_2e_compact(mol->d, nthreads, lay, host.p);
        Owned e(ctx, {n, n, n, n});
        {
            Scratch dc(ctx, size_t(std::max<int64_t>(lay.total, 1)) * 8);
            Scratch dt(ctx, (lay.shell_of.size() + lay.fn_of.size() + lay.nf.size()) * sizeof(int) +
                                (lay.pair_size.size() * 3) * sizeof(int64_t) + 64);
            char* base = dt.as<char>();
            int64_t* d_ps = reinterpret_cast<int64_t*>(base);
            int64_t* d_pre = d_ps + lay.pair_size.size();
            int64_t* d_row = d_pre + lay.prefix.size();
            int* d_sh = reinterpret_cast<int*>(d_row + lay.rowstart.size());
            int* d_fn = d_sh + lay.shell_of.size();
            int* d_nf = d_fn + lay.fn_of.size();
            auto up = [&](void* dst, const void* src, size_t bytes) {
                if (bytes) AQ_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
            };
            up(dc.p, host.p, size_t(lay.total) * 8);
            up(d_ps, lay.pair_size.data(), lay.pair_size.size() * sizeof(int64_t));
            up(d_pre, lay.prefix.data(), lay.prefix.size() * sizeof(int64_t));
            up(d_row, lay.rowstart.data(), lay.rowstart.size() * sizeof(int64_t));
            up(d_sh, lay.shell_of.data(), lay.shell_of.size() * sizeof(int));
            up(d_fn, lay.fn_of.data(), lay.fn_of.size() * sizeof(int));
            up(d_nf, lay.nf.data(), lay.nf.size() * sizeof(int));
            launch_eri_expand(ctx, dc.as<double>(), d_sh, d_fn, d_nf, d_ps, d_pre, d_row, n, lay.total, e.t->data);
            AQ_HIP(hipStreamSynchronize(ctx->stream));   // the host tables and pinned buffer die here
        }
        e.t->version += 1;
        *out = e.release();
    });
}

int aq_scf_default_config(aq_scf_config* cfg) {
    return guarded([&] {
        AQ_CHECK(cfg, AQ_ERR_INVALID, "null config");
        cfg->convergence = 1e-12;
        cfg->max_iterations = 150;
        cfg->diis_order = 6;
    });
}

int aq_scf_rhf(aq_ctx* ctx, const aq_molecule* mol, const aq_tensor* eri, const aq_scf_config* cfg_in,
               aq_scf_result* res, double* eps_out, double* C_out, double* F_out) {
    return guarded([&] {
        AQ_CHECK(ctx && mol && eri && res, AQ_ERR_INVALID, "null argument");
        aq_scf_config cfg;
        if (cfg_in)
            cfg = *cfg_in;
        else
            aq_scf_default_config(&cfg);
        const int n = mol->d.nao;
        const int ne = mol->d.nelectrons();
        AQ_CHECK(ne > 0 && ne % 2 == 0, AQ_ERR_UNSUPPORTED, "closed-shell SCF needs an even electron count");
        const int nocc = ne / 2;
        AQ_CHECK(nocc <= n, AQ_ERR_INVALID, "more occupied orbitals than basis functions");
        AQ_CHECK(eri->ndim == 4 && eri->len[0] == n && eri->len[1] == n && eri->len[2] == n && eri->len[3] == n,
                 AQ_ERR_INVALID, "AO tensor must be [nao,nao,nao,nao]");
        const size_t N = size_t(n), NN = N * N;
        std::vector<double> S(NN), T(NN), V(NN), H(NN);
        fe::ints_1e(mol->d, S.data(), T.data(), V.data());
        for (size_t i = 0; i < NN; ++i) H[i] = T[i] + V[i];
        const double enuc = mol->d.nuclear_repulsion();

        // X = S^{-1/2}
        std::vector<double> A = S, s, U, X(NN), tmp(NN), tmp2(NN);
        fe::eigh(n, A, s, U);
        AQ_CHECK(s[0] > 1e-10, AQ_ERR_NUMERIC, "overlap matrix is singular to working precision");
        for (size_t i = 0; i < N; ++i)
            for (size_t j = 0; j < N; ++j) {
                double x = 0.0;
                for (size_t k = 0; k < N; ++k) x += U[i * N + k] * U[j * N + k] / std::sqrt(s[k]);
                X[i * N + j] = x;
            }
        // C = X C' with F' = X^T F X = C' eps C'^T
        std::vector<double> C(NN), eps(N);
        auto diagonalise = [&](const std::vector<double>& F) {
            gemm_tn(n, X.data(), F.data(), tmp.data());   // X^T F
            gemm_nn(n, tmp.data(), X.data(), tmp2.data()); // X^T F X
            std::vector<double> Cp;
            fe::eigh(n, tmp2, eps, Cp);
            gemm_nn(n, X.data(), Cp.data(), C.data());
        };
        diagonalise(H);

        Owned Dd(ctx, {n, n}), Jd(ctx, {n, n}), Kd(ctx, {n, n});
        std::vector<double> D(NN), J(NN), K(NN), F(NN), err(NN);
        std::vector<std::vector<double>> fh, eh;
        double E = 0.0, E_old = 0.0, conv = 0.0;
        int it = 0;
        bool done = false;
        for (it = 1; it <= cfg.max_iterations; ++it) {
            for (size_t i = 0; i < N; ++i)
                for (size_t j = 0; j < N; ++j) {
                    double d = 0.0;
                    for (int k = 0; k < nocc; ++k) d += C[i * N + size_t(k)] * C[j * N + size_t(k)];
                    D[i * N + j] = d;
                }
            // symmetric matrices: row- and column-major agree
            ck(aq_write_all(ctx, Dd.t, D.data()));
            ck(aq_contract(ctx, 1.0, eri, "mnls", Dd.t, "ls", 0.0, Jd.t, "mn"));
            ck(aq_contract(ctx, 1.0, eri, "mlns", Dd.t, "ls", 0.0, Kd.t, "mn"));
            ck(aq_read_all(ctx, Jd.t, J.data()));
            ck(aq_read_all(ctx, Kd.t, K.data()));
            E = enuc;
            for (size_t i = 0; i < NN; ++i) {
                F[i] = H[i] + 2.0 * J[i] - K[i];
                E += D[i] * (H[i] + F[i]);
            }
            // orthonormal-basis commutator X^T (F D S - S D F) X (uhf.cxx:338-400)
            gemm_nn(n, F.data(), D.data(), tmp.data());
            gemm_nn(n, tmp.data(), S.data(), tmp2.data());
            for (size_t i = 0; i < N; ++i)
                for (size_t j = 0; j < N; ++j) err[i * N + j] = tmp2[i * N + j] - tmp2[j * N + i];
            gemm_tn(n, X.data(), err.data(), tmp.data());
            gemm_nn(n, tmp.data(), X.data(), err.data());
            conv = 0.0;
            for (double x : err) conv = std::max(conv, std::fabs(x));
            if (it > 1 && std::fabs(E - E_old) < cfg.convergence && conv < 100.0 * cfg.convergence) {
                done = true;
                break;
            }
            E_old = E;
            fh.push_back(F);
            eh.push_back(err);
            if (int(fh.size()) > std::max(cfg.diis_order, 1)) {
                fh.erase(fh.begin());
                eh.erase(eh.begin());
            }
            const int m = int(fh.size());
            std::vector<double> Fx = F;
            if (m >= 2) {
                std::vector<double> B(size_t(m + 1) * size_t(m + 1), -1.0), rhs(size_t(m + 1), 0.0);
                B[size_t(m) * size_t(m + 1) + size_t(m)] = 0.0;
                for (int i = 0; i < m; ++i)
                    for (int j = 0; j < m; ++j) {
                        double b = 0.0;
                        for (size_t k = 0; k < NN; ++k) b += eh[size_t(i)][k] * eh[size_t(j)][k];
                        B[size_t(i) * size_t(m + 1) + size_t(j)] = b;
                    }
                rhs[size_t(m)] = -1.0;
                if (solve(m + 1, B, rhs)) {
                    std::fill(Fx.begin(), Fx.end(), 0.0);
                    for (int i = 0; i < m; ++i)
                        for (size_t k = 0; k < NN; ++k) Fx[k] += rhs[size_t(i)] * fh[size_t(i)][k];
                }
            }
            diagonalise(Fx);
        }
        // canonical orbitals of the final Fock matrix
        diagonalise(F);
        res->energy = E;
        res->conv = conv;
        res->iterations = done ? it : cfg.max_iterations;
        res->converged = done ? 1 : 0;
        res->nocc = nocc;
        res->nocc_beta = nocc;
        if (eps_out) std::memcpy(eps_out, eps.data(), N * sizeof(double));
        // column-major [nao, nmo]: C_out[m + p n] = C[m][p]
        if (C_out)
            for (size_t i = 0; i < N; ++i)
                for (size_t p = 0; p < N; ++p) C_out[i + p * N] = C[i * N + p];
        if (F_out) std::memcpy(F_out, F.data(), NN * sizeof(double));
    });
}

int aq_aomo_rhf(aq_ctx* ctx, const aq_tensor* eri, int nocc, const double* Ccm, const double* F, aq_tensor** out) {
    return guarded([&] {
        AQ_CHECK(ctx && eri && Ccm && F && out, AQ_ERR_INVALID, "null argument");
        AQ_CHECK(eri->ndim == 4, AQ_ERR_INVALID, "AO tensor must be 4-dimensional");
        const int64_t n = eri->len[0];
        AQ_CHECK(eri->len[1] == n && eri->len[2] == n && eri->len[3] == n, AQ_ERR_INVALID, "AO tensor must be [n,n,n,n]");
        AQ_CHECK(nocc >= 1 && nocc < n, AQ_ERR_INVALID, "need 1 <= nocc < nao");
        const size_t N = size_t(n);
        // Fock matrix in the MO basis, f = C^T F C
        std::vector<double> CF(N * N, 0.0), f(N * N, 0.0);
        for (size_t p = 0; p < N; ++p)
            for (size_t l = 0; l < N; ++l) {
                double s = 0.0;
                for (size_t m = 0; m < N; ++m) s += Ccm[m + p * N] * F[m * N + l];
                CF[p * N + l] = s;
            }
        for (size_t p = 0; p < N; ++p)
            for (size_t q = 0; q < N; ++q) {
                double s = 0.0;
                for (size_t l = 0; l < N; ++l) s += CF[p * N + l] * Ccm[l + q * N];
                f[p * N + q] = s;
            }
        // four quarter transformations
        Owned Cd(ctx, {n, n});
        ck(aq_write_all(ctx, Cd.t, Ccm));
        Owned mo(ctx, {n, n, n, n});
        {
            Owned x1(ctx, {n, n, n, n});
            ck(aq_contract(ctx, 1.0, Cd.t, "mp", eri, "mnls", 0.0, x1.t, "pnls"));
            ck(aq_contract(ctx, 1.0, Cd.t, "nq", x1.t, "pnls", 0.0, mo.t, "pqls"));
            ck(aq_contract(ctx, 1.0, Cd.t, "lr", mo.t, "pqls", 0.0, x1.t, "pqrs"));
            ck(aq_contract(ctx, 1.0, Cd.t, "st", x1.t, "pqrs", 0.0, mo.t, "pqrt"));
        }
        ck(aq_ctx_trim(ctx));
        rhf_blocks(ctx, mo.t, nocc, f, out);
    });
}

int aq_mo_rhf_blocks(aq_ctx* ctx, const aq_tensor* mo_eri, int nocc, const double* h, double* escf_elec,
                     double* f_out, aq_tensor** out) {
    return guarded([&] {
        AQ_CHECK(ctx && mo_eri && h && out, AQ_ERR_INVALID, "null argument");
        AQ_CHECK(mo_eri->ndim == 4, AQ_ERR_INVALID, "MO tensor must be 4-dimensional");
        const int64_t n = mo_eri->len[0];
        AQ_CHECK(mo_eri->len[1] == n && mo_eri->len[2] == n && mo_eri->len[3] == n, AQ_ERR_INVALID,
                 "MO tensor must be [n,n,n,n]");
        AQ_CHECK(nocc >= 1 && nocc < n, AQ_ERR_INVALID, "need 1 <= nocc < norb");
        const size_t N = size_t(n);
        // closed-shell Fock of the orbitals as given: f = h + 2 J[D] - K[D], D = occupied projector
        std::vector<double> D(N * N, 0.0), J(N * N), K(N * N), f(N * N);
        for (int k = 0; k < nocc; ++k) D[size_t(k) * N + size_t(k)] = 1.0;
        Owned Dd(ctx, {n, n}), Jd(ctx, {n, n}), Kd(ctx, {n, n});
        ck(aq_write_all(ctx, Dd.t, D.data()));
        ck(aq_contract(ctx, 1.0, mo_eri, "pqkl", Dd.t, "kl", 0.0, Jd.t, "pq"));
        ck(aq_contract(ctx, 1.0, mo_eri, "pklq", Dd.t, "kl", 0.0, Kd.t, "pq"));
        ck(aq_read_all(ctx, Jd.t, J.data()));
        ck(aq_read_all(ctx, Kd.t, K.data()));
        double e = 0.0;
        // h, J and K in column-major key order (p + q n); f row-major for the block cutter
        for (size_t p = 0; p < N; ++p)
            for (size_t q = 0; q < N; ++q) f[p * N + q] = h[p + q * N] + 2.0 * J[p + q * N] - K[p + q * N];
        for (int i = 0; i < nocc; ++i) e += h[size_t(i) * (N + 1)] + f[size_t(i) * (N + 1)];
        if (escf_elec) *escf_elec = e;
        if (f_out)
            for (size_t p = 0; p < N; ++p)
                for (size_t q = 0; q < N; ++q) f_out[p + q * N] = f[p * N + q];
        rhf_blocks(ctx, mo_eri, nocc, f, out);
    });
}

}  // extern "C"
