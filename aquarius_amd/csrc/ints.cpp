// One- and two-electron Gaussian integrals on the host (McMurchie-Davidson).
//
// Reference: the 1eints and 2eints tasks (src/integrals/1eints.cxx,
// src/integrals/2eints.cxx, os.cxx) compute the same overlap, kinetic,
// nuclear-attraction and (mn|ls) Coulomb integrals with Obara-Saika
// recursions; only the values are shared.  Here every shell pair is expanded
// in Hermite Gaussians (Helgaker, Joergensen, Olsen, "Molecular
// Electronic-Structure Theory", sec. 9.5-9.9):
//   E^{ij}_t  (9.5.6-9.5.7), R^n_{tuv} (9.9.18-9.9.20),
//   (ab|cd) = 2 pi^{5/2} / (p q sqrt(p+q)) sum_{tuv} E^{ab}_{tuv}
//             sum_{t'u'v'} (-1)^{t'+u'+v'} E^{cd}_{t'u'v'} R_{t+t',u+u',v+v'}.
// Shell quartets with a Cauchy-Schwarz bound below 1e-15 (the reference's
// calc_cutoff, 2eints.cxx:524-531) are written as zeros.  The contractions of
// one general contraction (same centre, L and exponents) share their
// primitive quartets.  Work is spread over threads by canonical bra pair of
// shell families; distinct canonical quartets own disjoint output elements,
// so the threads never write the same address.
#include "frontend.hpp"
#include "internal.hpp"

#include <algorithm>
#include <atomic>
#include <exception>
#include <cmath>
#include <cstring>
#include <mutex>
#include <thread>

namespace aqamd {
namespace fe {

// ------------------------------------------------------------ Boys function
//
// F_m(T) = int_0^1 t^{2m} exp(-T t^2) dt.  Table of F_m(k dT) for T < 40 from
// the positive series e^{-T} sum_k (2T)^k / ((2m+1)(2m+3)...(2m+2k+1)); a point
// is F_{mmax} by a 7-term Taylor expansion around the nearest grid point
// (dF_m/dT = -F_{m+1}, |remainder| < 1.3e-15 F_m), then downward recursion
// F_m = (2T F_{m+1} + e^{-T}) / (2m+1).  Above T = 40: F_0 = sqrt(pi/T)/2
// (erf(sqrt(40)) = 1 to 1e-18) and the upward recursion, stable there.
namespace {

constexpr double kDT = 0.05;
constexpr double kTmax = 40.0;
constexpr int kNT = int(kTmax / kDT) + 2;
constexpr int kTaylor = 7;
constexpr int kTabM = kMaxM + kTaylor + 1;

const double* boys_table() {
    static std::vector<double> tab;
    static std::once_flag once;
    std::call_once(once, [] {
        tab.assign(size_t(kNT) * kTabM, 0.0);
        for (int k = 0; k < kNT; ++k) {
            const double T = k * kDT;
            const double e = std::exp(-T);
            for (int m = 0; m < kTabM; ++m) {
                double term = 1.0 / (2 * m + 1), sum = term;
                for (int j = 1; j < 1000; ++j) {
                    term *= 2.0 * T / (2 * m + 2 * j + 1);
                    sum += term;
                    if (term < 1e-18 * sum) break;
                }
                tab[size_t(k) * kTabM + size_t(m)] = e * sum;
            }
        }
    });
    return tab.data();
}

}  // namespace

void boys(int mmax, double T, double* F) {
    if (T < kTmax) {
        const double* tab = boys_table();
        const int k = int(T / kDT + 0.5);
        const double d = -(T - k * kDT);
        const double* row = tab + size_t(k) * kTabM + size_t(mmax);
        double f = 0.0, pw = 1.0;
        for (int j = 0; j < kTaylor; ++j) {
            f += row[j] * pw;
            pw *= d / (j + 1);
        }
        F[mmax] = f;
        const double e = std::exp(-T);
        for (int m = mmax - 1; m >= 0; --m) F[m] = (2.0 * T * F[m + 1] + e) / (2 * m + 1);
    } else {
        const double e = std::exp(-T), inv = 0.5 / T;
        F[0] = 0.5 * std::sqrt(M_PI / T);
        for (int m = 0; m < mmax; ++m) F[m + 1] = ((2 * m + 1) * F[m] - e) * inv;
    }
}

// ------------------------------------------------------- Hermite machinery
namespace {

// E[(i*(lb+1) + j)*(la+lb+1) + t], i <= la, j <= lb, t <= i+j, with E^{00}_0 = 1.
void hermite_e(int la, int lb, double XPA, double XPB, double inv2p, double* E) {
    const int nt = la + lb + 1;
    auto at = [&](int i, int j, int t) -> double& { return E[(size_t(i) * size_t(lb + 1) + size_t(j)) * size_t(nt) + size_t(t)]; };
    std::fill(E, E + size_t(la + 1) * size_t(lb + 1) * size_t(nt), 0.0);
    at(0, 0, 0) = 1.0;
    for (int i = 0; i <= la; ++i)
        for (int j = 0; j <= lb; ++j) {
            if (i == 0 && j == 0) continue;
            const bool up_i = i > 0;
            const int si = up_i ? i - 1 : i, sj = up_i ? j : j - 1;
            const double X = up_i ? XPA : XPB;
            const int smax = si + sj;
            for (int t = 0; t <= i + j; ++t) {
                double v = t <= smax ? X * at(si, sj, t) : 0.0;
                if (t > 0 && t - 1 <= smax) v += inv2p * at(si, sj, t - 1);
                if (t + 1 <= smax) v += (t + 1) * at(si, sj, t + 1);
                at(i, j, t) = v;
            }
        }
}

struct Herm {
    int L = -1;
    std::vector<std::array<int, 3>> tuv;      // t+u+v <= L
};

const Herm& hermite_list(int L) {
    static Herm tab[4 * kMaxL + 1];
    static std::once_flag once;
    std::call_once(once, [] {
        for (int l = 0; l <= 4 * kMaxL; ++l) {
            tab[l].L = l;
            for (int t = 0; t <= l; ++t)
                for (int u = 0; u <= l - t; ++u)
                    for (int v = 0; v <= l - t - u; ++v) tab[l].tuv.push_back({t, u, v});
        }
    });
    return tab[L];
}

// R_{tuv} for t+u+v <= L into R[(t*(L+1)+u)*(L+1)+v]; work: (L+1)^4 doubles.
void hermite_r(int L, double alpha, double X, double Y, double Z, double* R, double* work) {
    double F[kMaxM + 1];
    boys(L, alpha * (X * X + Y * Y + Z * Z), F);
    const size_t D = size_t(L + 1);
    auto W = [&](int n, int t, int u, int v) -> double& {
        return work[((size_t(n) * D + size_t(t)) * D + size_t(u)) * D + size_t(v)];
    };
    double f = 1.0;
    for (int n = 0; n <= L; ++n) {
        W(n, 0, 0, 0) = f * F[n];
        f *= -2.0 * alpha;
    }
    for (int s = 1; s <= L; ++s)
        for (int t = 0; t <= s; ++t)
            for (int u = 0; u <= s - t; ++u) {
                const int v = s - t - u;
                for (int n = 0; n <= L - s; ++n) {
                    double r;
                    if (t > 0) {
                        r = X * W(n + 1, t - 1, u, v);
                        if (t > 1) r += (t - 1) * W(n + 1, t - 2, u, v);
                    } else if (u > 0) {
                        r = Y * W(n + 1, t, u - 1, v);
                        if (u > 1) r += (u - 1) * W(n + 1, t, u - 2, v);
                    } else {
                        r = Z * W(n + 1, t, u, v - 1);
                        if (v > 1) r += (v - 1) * W(n + 1, t, u, v - 2);
                    }
                    W(n, t, u, v) = r;
                }
            }
    for (int t = 0; t <= L; ++t)
        for (int u = 0; u <= L - t; ++u)
            for (int v = 0; v <= L - t - u; ++v) R[(size_t(t) * D + size_t(u)) * D + size_t(v)] = W(0, t, u, v);
}

// out[f1 * ldo + f2] += sum T1[f1,c1] T2[f2,c2] cart[c1 * n2 + c2]  (nullptr T: identity)
void to_ao_2(const Shell& s1, const Shell& s2, const double* cart, double* out, int ldo) {
    const int n1 = ncart(s1.L), n2 = ncart(s2.L);
    std::vector<double> half(size_t(s1.nf) * size_t(n2), 0.0);
    for (int f = 0; f < s1.nf; ++f)
        for (int c = 0; c < n1; ++c) {
            const double w = s1.T ? s1.T[size_t(f) * size_t(n1) + size_t(c)] : (f == c ? 1.0 : 0.0);
            if (w == 0.0) continue;
            for (int c2 = 0; c2 < n2; ++c2) half[size_t(f) * size_t(n2) + size_t(c2)] += w * cart[size_t(c) * size_t(n2) + size_t(c2)];
        }
    for (int f = 0; f < s1.nf; ++f)
        for (int g = 0; g < s2.nf; ++g) {
            double acc = 0.0;
            for (int c2 = 0; c2 < n2; ++c2) {
                const double w = s2.T ? s2.T[size_t(g) * size_t(n2) + size_t(c2)] : (g == c2 ? 1.0 : 0.0);
                acc += w * half[size_t(f) * size_t(n2) + size_t(c2)];
            }
            out[size_t(f) * size_t(ldo) + size_t(g)] = acc;
        }
}

}  // namespace

// ----------------------------------------------------------- one-electron

void ints_1e(const MoleculeData& m, double* S, double* T, double* V) {
    const int N = m.nao;
    const int natom = int(m.Z.size());
    std::vector<double> Ex, Ey, Ez, R, work, cs, ct, cv, blk;
    for (size_t ia = 0; ia < m.shells.size(); ++ia)
        for (size_t ib = 0; ib <= ia; ++ib) {
            const Shell& a = m.shells[ia];
            const Shell& b = m.shells[ib];
            const int la = a.L, lb = b.L, na = ncart(la), nb = ncart(lb);
            const auto& Pa = cart_powers(la);
            const auto& Pb = cart_powers(lb);
            const int L = la + lb;
            const size_t ne = size_t(la + 1) * size_t(lb + 3) * size_t(la + lb + 3);
            Ex.resize(ne);
            Ey.resize(ne);
            Ez.resize(ne);
            R.resize(size_t(L + 1) * size_t(L + 1) * size_t(L + 1));
            work.resize(size_t(L + 1) * size_t(L + 1) * size_t(L + 1) * size_t(L + 1));
            cs.assign(size_t(na) * size_t(nb), 0.0);
            ct.assign(size_t(na) * size_t(nb), 0.0);
            cv.assign(size_t(na) * size_t(nb), 0.0);
            const double AB2 = (a.A[0] - b.A[0]) * (a.A[0] - b.A[0]) + (a.A[1] - b.A[1]) * (a.A[1] - b.A[1]) +
                               (a.A[2] - b.A[2]) * (a.A[2] - b.A[2]);
            const int nt = la + lb + 3;
            auto e = [&](const std::vector<double>& E, int i, int j, int t) {
                return (j < 0 || t > i + j) ? 0.0 : E[(size_t(i) * size_t(lb + 3) + size_t(j)) * size_t(nt) + size_t(t)];
            };
            for (size_t i = 0; i < a.a.size(); ++i)
                for (size_t j = 0; j < b.a.size(); ++j) {
                    const double ai = a.a[i], bj = b.a[j], p = ai + bj;
                    const double K = std::exp(-ai * bj / p * AB2) * a.c[i] * b.c[j];
                    double P[3];
                    for (int d = 0; d < 3; ++d) P[d] = (ai * a.A[d] + bj * b.A[d]) / p;
                    hermite_e(la, lb + 2, P[0] - a.A[0], P[0] - b.A[0], 0.5 / p, Ex.data());
                    hermite_e(la, lb + 2, P[1] - a.A[1], P[1] - b.A[1], 0.5 / p, Ey.data());
                    hermite_e(la, lb + 2, P[2] - a.A[2], P[2] - b.A[2], 0.5 / p, Ez.data());
                    const double sq = std::sqrt(M_PI / p);
                    const std::vector<double>* Es[3] = {&Ex, &Ey, &Ez};
                    for (int ca = 0; ca < na; ++ca)
                        for (int cb = 0; cb < nb; ++cb) {
                            double s1[3], t1[3];
                            for (int d = 0; d < 3; ++d) {
                                const int pi = Pa[size_t(ca)][size_t(d)], pj = Pb[size_t(cb)][size_t(d)];
                                const std::vector<double>& E = *Es[d];
                                s1[d] = e(E, pi, pj, 0) * sq;
                                t1[d] = -0.5 * (pj * (pj - 1) * e(E, pi, pj - 2, 0) * sq -
                                                2.0 * bj * (2 * pj + 1) * s1[d] + 4.0 * bj * bj * e(E, pi, pj + 2, 0) * sq);
                            }
                            cs[size_t(ca) * size_t(nb) + size_t(cb)] += K * s1[0] * s1[1] * s1[2];
                            ct[size_t(ca) * size_t(nb) + size_t(cb)] +=
                                K * (t1[0] * s1[1] * s1[2] + s1[0] * t1[1] * s1[2] + s1[0] * s1[1] * t1[2]);
                        }
                    for (int at = 0; at < natom; ++at) {
                        const double* C = &m.xyz[3 * size_t(at)];
                        hermite_r(L, p, P[0] - C[0], P[1] - C[1], P[2] - C[2], R.data(), work.data());
                        const double f = -m.Z[size_t(at)] * 2.0 * M_PI / p * K;
                        const size_t D = size_t(L + 1);
                        for (int ca = 0; ca < na; ++ca)
                            for (int cb = 0; cb < nb; ++cb) {
                                const auto& pa = Pa[size_t(ca)];
                                const auto& pb = Pb[size_t(cb)];
                                double acc = 0.0;
                                for (int t = 0; t <= pa[0] + pb[0]; ++t)
                                    for (int u = 0; u <= pa[1] + pb[1]; ++u)
                                        for (int v = 0; v <= pa[2] + pb[2]; ++v)
                                            acc += e(Ex, pa[0], pb[0], t) * e(Ey, pa[1], pb[1], u) *
                                                   e(Ez, pa[2], pb[2], v) * R[(size_t(t) * D + size_t(u)) * D + size_t(v)];
                                cv[size_t(ca) * size_t(nb) + size_t(cb)] += f * acc;
                            }
                    }
                }
            blk.resize(size_t(a.nf) * size_t(b.nf));
            double* dst[3] = {S, T, V};
            const std::vector<double>* src[3] = {&cs, &ct, &cv};
            for (int k = 0; k < 3; ++k) {
                to_ao_2(a, b, src[k]->data(), blk.data(), b.nf);
                for (int f = 0; f < a.nf; ++f)
                    for (int g = 0; g < b.nf; ++g) {
                        const double x = blk[size_t(f) * size_t(b.nf) + size_t(g)];
                        dst[k][size_t(a.ao0 + f) * size_t(N) + size_t(b.ao0 + g)] = x;
                        dst[k][size_t(b.ao0 + g) * size_t(N) + size_t(a.ao0 + f)] = x;
                    }
            }
        }
}

// ----------------------------------------------------------- two-electron
namespace {

// One primitive product of a family pair: exponent sum and centre; its
// Hermite expansion, without contraction coefficients, at E[...] of the pair.
struct PrimPair {
    double p, P[3];
    size_t E;                       // offset of E[h * nab + ab] in the family pair's table
};

// Shells of one centre with the same L and the same exponents -- the
// contractions of one general contraction, e.g. cc-pVDZ's 1s and 2s on O --
// at consecutive shell indices [first, first + n).  Their primitive
// integrals are computed once and weighted per contraction.
struct Family {
    int first = 0, n = 0;
};

std::vector<Family> families(const MoleculeData& m) {
    std::vector<Family> f;
    for (int s = 0; s < int(m.shells.size()); ++s) {
        const Shell& x = m.shells[size_t(s)];
        if (!f.empty()) {
            const Shell& y = m.shells[size_t(f.back().first)];
            if (y.atom == x.atom && y.L == x.L && y.a == x.a) {
                f.back().n += 1;
                continue;
            }
        }
        f.push_back(Family{s, 1});
    }
    return f;
}

// The shell pairs (a, b), a >= b, of two families F1 >= F2.  Member m's
// integrals are the raw primitive ones weighted by coef[m * nprim + k] =
// c_a[i] c_b[j] for primitive product k = (i, j).
struct FamPair {
    int L = 0, nab = 0, nh = 0;
    std::vector<PrimPair> prims;
    std::vector<double> E;                  // per primitive product: nh x nab
    std::vector<std::pair<int, int>> members;
    std::vector<int64_t> pair_index;        // a(a+1)/2 + b per member
    std::vector<double> coef;               // [member][primitive product]
    std::vector<double> Q;                  // per member: sqrt(max |(ab|ab)| * nab)
};

FamPair make_fampair(const MoleculeData& m, const Family& F1, const Family& F2) {
    const Shell& a = m.shells[size_t(F1.first)];
    const Shell& b = m.shells[size_t(F2.first)];
    FamPair pd;
    for (int x = F1.first; x < F1.first + F1.n; ++x)
        for (int y = F2.first; y < F2.first + F2.n; ++y)
            if (x >= y) {
                pd.members.emplace_back(x, y);
                pd.pair_index.push_back(int64_t(x) * (x + 1) / 2 + y);
            }
    pd.L = a.L + b.L;
    const int na = ncart(a.L), nb = ncart(b.L);
    pd.nab = na * nb;
    const Herm& H = hermite_list(pd.L);
    pd.nh = int(H.tuv.size());
    const auto& Pa = cart_powers(a.L);
    const auto& Pb = cart_powers(b.L);
    const int nt = pd.L + 1;
    std::vector<double> Ex(size_t(a.L + 1) * size_t(b.L + 1) * size_t(nt)), Ey(Ex.size()), Ez(Ex.size());
    const double AB2 = (a.A[0] - b.A[0]) * (a.A[0] - b.A[0]) + (a.A[1] - b.A[1]) * (a.A[1] - b.A[1]) +
                       (a.A[2] - b.A[2]) * (a.A[2] - b.A[2]);
    std::vector<std::pair<size_t, size_t>> ij;
    for (size_t i = 0; i < a.a.size(); ++i)
        for (size_t j = 0; j < b.a.size(); ++j) {
            const double ai = a.a[i], bj = b.a[j], p = ai + bj;
            const double K = std::exp(-ai * bj / p * AB2);
            if (K < 1e-300) continue;
            PrimPair pp;
            pp.p = p;
            for (int d = 0; d < 3; ++d) pp.P[d] = (ai * a.A[d] + bj * b.A[d]) / p;
            hermite_e(a.L, b.L, pp.P[0] - a.A[0], pp.P[0] - b.A[0], 0.5 / p, Ex.data());
            hermite_e(a.L, b.L, pp.P[1] - a.A[1], pp.P[1] - b.A[1], 0.5 / p, Ey.data());
            hermite_e(a.L, b.L, pp.P[2] - a.A[2], pp.P[2] - b.A[2], 0.5 / p, Ez.data());
            pp.E = pd.E.size();
            pd.E.resize(pd.E.size() + size_t(pd.nh) * size_t(pd.nab), 0.0);
            double* Ep = pd.E.data() + pp.E;
            auto e = [&](const std::vector<double>& E, int x, int y, int t) {
                return t > x + y ? 0.0 : E[(size_t(x) * size_t(b.L + 1) + size_t(y)) * size_t(nt) + size_t(t)];
            };
            for (int h = 0; h < pd.nh; ++h) {
                const auto& tuv = H.tuv[size_t(h)];
                for (int ca = 0; ca < na; ++ca)
                    for (int cb = 0; cb < nb; ++cb) {
                        const auto& pa = Pa[size_t(ca)];
                        const auto& pb = Pb[size_t(cb)];
                        Ep[size_t(h) * size_t(pd.nab) + size_t(ca * nb + cb)] =
                            K * e(Ex, pa[0], pb[0], tuv[0]) * e(Ey, pa[1], pb[1], tuv[1]) * e(Ez, pa[2], pb[2], tuv[2]);
                    }
            }
            pd.prims.push_back(pp);
            ij.emplace_back(i, j);
        }
    const size_t np = pd.prims.size();
    pd.coef.resize(pd.members.size() * np);
    for (size_t mi = 0; mi < pd.members.size(); ++mi) {
        const Shell& sa = m.shells[size_t(pd.members[mi].first)];
        const Shell& sb = m.shells[size_t(pd.members[mi].second)];
        for (size_t k = 0; k < np; ++k) pd.coef[mi * np + k] = sa.c[ij[k].first] * sb.c[ij[k].second];
    }
    return pd;
}

struct QuartetWork {
    std::vector<double> R, work, tmp, Y, W, Z, out, ao, half;
};

// Cartesian (ab|cd) of every bra member mb and ket member mk into
// w.out[(mb * nmk + mk) * nab * ncd + ab * ncd + cd].  The Hermite integrals
// and their product with the ket expansion are formed once per primitive
// quartet and weighted per member; a family of one folds its weight into the
// accumulation instead.
void quartet(const FamPair& bra, const FamPair& ket, QuartetWork& w) {
    const int Lt = bra.L + ket.L;
    const size_t D = size_t(Lt + 1);
    w.R.resize(D * D * D);
    w.work.resize(D * D * D * D);
    const Herm& Hb = hermite_list(bra.L);
    const Herm& Hk = hermite_list(ket.L);
    const int nhb = bra.nh, nhk = ket.nh, nab = bra.nab, ncd = ket.nab;
    const size_t nmb = bra.members.size(), nmk = ket.members.size();
    const size_t npb = bra.prims.size(), npk = ket.prims.size();
    const size_t blk = size_t(nab) * size_t(ncd), wsz = size_t(nhb) * size_t(ncd);
    w.tmp.resize(size_t(nhb) * size_t(nhk));
    w.Y.resize(wsz);
    w.W.resize(wsz * nmk);
    w.Z.resize(blk);
    w.out.assign(blk * nmb * nmk, 0.0);
    for (size_t k1 = 0; k1 < npb; ++k1) {
        const PrimPair& p1 = bra.prims[k1];
        std::fill(w.W.begin(), w.W.end(), 0.0);
        for (size_t k2 = 0; k2 < npk; ++k2) {
            const PrimPair& p2 = ket.prims[k2];
            const double p = p1.p, q = p2.p;
            const double alpha = p * q / (p + q);
            const double pref = 2.0 * std::pow(M_PI, 2.5) / (p * q * std::sqrt(p + q));
            hermite_r(Lt, alpha, p1.P[0] - p2.P[0], p1.P[1] - p2.P[1], p1.P[2] - p2.P[2], w.R.data(), w.work.data());
            for (int hb = 0; hb < nhb; ++hb) {
                const auto& b3 = Hb.tuv[size_t(hb)];
                for (int hk = 0; hk < nhk; ++hk) {
                    const auto& k3 = Hk.tuv[size_t(hk)];
                    const double sgn = ((k3[0] + k3[1] + k3[2]) & 1) ? -pref : pref;
                    w.tmp[size_t(hb) * size_t(nhk) + size_t(hk)] =
                        sgn * w.R[(size_t(b3[0] + k3[0]) * D + size_t(b3[1] + k3[1])) * D + size_t(b3[2] + k3[2])];
                }
            }
            const double* E2 = ket.E.data() + p2.E;
            // one ket member: W += (c * tmp) x E2; several: Y = tmp x E2, W_m += c_m Y
            const double c2 = nmk == 1 ? ket.coef[k2] : 1.0;
            double* acc = nmk == 1 ? w.W.data() : w.Y.data();
            if (nmk > 1) std::fill(w.Y.begin(), w.Y.end(), 0.0);
            for (int hb = 0; hb < nhb; ++hb) {
                double* Wr = acc + size_t(hb) * size_t(ncd);
                const double* tr = w.tmp.data() + size_t(hb) * size_t(nhk);
                for (int hk = 0; hk < nhk; ++hk) {
                    const double t = tr[hk] * c2;
                    const double* er = E2 + size_t(hk) * size_t(ncd);
                    for (int cd = 0; cd < ncd; ++cd) Wr[cd] += t * er[cd];
                }
            }
            if (nmk > 1)
                for (size_t mk = 0; mk < nmk; ++mk) {
                    const double c = ket.coef[mk * npk + k2];
                    double* Wm = w.W.data() + mk * wsz;
                    for (size_t x = 0; x < wsz; ++x) Wm[x] += c * w.Y[x];
                }
        }
        // out(mb, mk) += c_mb(k1) * E1^T W_mk
        const double* E1 = bra.E.data() + p1.E;
        for (size_t mk = 0; mk < nmk; ++mk) {
            const double* Wm = w.W.data() + mk * wsz;
            const double c1 = nmb == 1 ? bra.coef[k1] : 1.0;
            double* acc = nmb == 1 ? w.out.data() + mk * blk : w.Z.data();
            if (nmb > 1) std::fill(w.Z.begin(), w.Z.end(), 0.0);
            for (int hb = 0; hb < nhb; ++hb) {
                const double* er = E1 + size_t(hb) * size_t(nab);
                const double* Wr = Wm + size_t(hb) * size_t(ncd);
                for (int ab = 0; ab < nab; ++ab) {
                    if (er[ab] == 0.0) continue;
                    const double e = er[ab] * c1;
                    double* o = acc + size_t(ab) * size_t(ncd);
                    for (int cd = 0; cd < ncd; ++cd) o[cd] += e * Wr[cd];
                }
            }
            if (nmb > 1)
                for (size_t mb = 0; mb < nmb; ++mb) {
                    const double c = bra.coef[mb * npb + k1];
                    double* o = w.out.data() + (mb * nmk + mk) * blk;
                    for (size_t x = 0; x < blk; ++x) o[x] += c * w.Z[x];
                }
        }
    }
}

// Transform one index of a [n0][n1][n2][n3] block (index `k`) by shell s.
void transform_index(std::vector<double>& x, int dims[4], int k, const Shell& s, std::vector<double>& tmp) {
    if (!s.T) return;
    const int nc = dims[k], nf = s.nf;
    size_t outer = 1, inner = 1;
    for (int i = 0; i < k; ++i) outer *= size_t(dims[i]);
    for (int i = k + 1; i < 4; ++i) inner *= size_t(dims[i]);
    tmp.assign(outer * size_t(nf) * inner, 0.0);
    for (size_t o = 0; o < outer; ++o)
        for (int f = 0; f < nf; ++f)
            for (int c = 0; c < nc; ++c) {
                const double w = s.T[size_t(f) * size_t(nc) + size_t(c)];
                if (w == 0.0) continue;
                const double* src = x.data() + (o * size_t(nc) + size_t(c)) * inner;
                double* dst = tmp.data() + (o * size_t(nf) + size_t(f)) * inner;
                for (size_t i = 0; i < inner; ++i) dst[i] += w * src[i];
            }
    x.swap(tmp);
    dims[k] = nf;
}

// shell pair index a(a+1)/2 + b (a >= b) -> (a, b)
void unpair(int64_t idx, int& a, int& b) {
    int64_t x = int64_t((std::sqrt(8.0 * double(idx) + 1.0) - 1.0) / 2.0);
    while (x * (x + 1) / 2 > idx) --x;
    while ((x + 1) * (x + 2) / 2 <= idx) ++x;
    a = int(x);
    b = int(idx - x * (x + 1) / 2);
}

}  // namespace

// Every canonical shell quartet (bra pair i >= ket pair j, each pair a >= b,
// i and j the pair indices a(a+1)/2 + b) evaluated once, its pure-harmonic
// block [fa][fb][fc][fd] handed to sink(i, j, block) (block == nullptr:
// screened to zero).  The work runs over pairs of shell families (the
// contractions of a general contraction share their primitive integrals),
// on `nthreads` workers; distinct (i, j) are distinct calls.
template <class Sink>
static void for_each_quartet(const MoleculeData& m, int nthreads, Sink&& sink) {
    const std::vector<Family> fam = families(m);
    const int64_t nfam = int64_t(fam.size());
    const size_t nfp = size_t(nfam) * size_t(nfam + 1) / 2;
    std::vector<FamPair> fps(nfp);
    if (nthreads <= 0) nthreads = int(std::max(1u, std::thread::hardware_concurrency()));
    nthreads = std::min(nthreads, 64);

    // a failure in a worker (e.g. a host allocation) stops the others and is
    // rethrown on the calling thread once every worker has joined
    auto run = [&](auto&& body, size_t n) {
        std::atomic<size_t> next{0};
        std::exception_ptr err;
        std::mutex err_mu;
        std::vector<std::thread> pool;
        for (int t = 0; t < nthreads; ++t)
            pool.emplace_back([&] {
                try {
                    QuartetWork w;
                    for (size_t i; (i = next.fetch_add(1)) < n;) body(i, w);
                } catch (...) {
                    std::lock_guard<std::mutex> lk(err_mu);
                    if (!err) err = std::current_exception();
                    next.store(n);
                }
            });
        for (auto& th : pool) th.join();
        if (err) std::rethrow_exception(err);
    };

    // family pairs (F1 >= F2) and the Schwarz factor of each member shell pair
    run(
        [&](size_t idx, QuartetWork& w) {
            int f1, f2;
            unpair(int64_t(idx), f1, f2);
            FamPair& fp = fps[idx];
            fp = make_fampair(m, fam[size_t(f1)], fam[size_t(f2)]);
            quartet(fp, fp, w);
            const size_t nm = fp.members.size(), blk = size_t(fp.nab) * size_t(fp.nab);
            fp.Q.resize(nm);
            for (size_t mi = 0; mi < nm; ++mi) {
                const double* o = w.out.data() + (mi * nm + mi) * blk;
                double mx = 0.0;
                for (int i = 0; i < fp.nab; ++i) mx = std::max(mx, std::fabs(o[size_t(i) * size_t(fp.nab) + size_t(i)]));
                // the cartesian diagonal bounds every pure-harmonic combination up to
                // the (normalised) transform's row norm, which is <= sqrt(ncart)
                fp.Q[mi] = std::sqrt(mx * fp.nab);
            }
        },
        nfp);

    const double cutoff = 1e-15;
    // bra family pairs from the last (largest shell indices) down
    run(
        [&](size_t r, QuartetWork& w) {
            const size_t I = nfp - 1 - r;
            const FamPair& bra = fps[I];
            const double qb = *std::max_element(bra.Q.begin(), bra.Q.end());
            std::vector<double> tmp;
            for (size_t J = 0; J <= I; ++J) {
                const FamPair& ket = fps[J];
                const double qk = *std::max_element(ket.Q.begin(), ket.Q.end());
                const bool any = qb * qk >= cutoff;
                if (any) quartet(bra, ket, w);
                const size_t nmb = bra.members.size(), nmk = ket.members.size();
                const size_t blk = size_t(bra.nab) * size_t(ket.nab);
                for (size_t mb = 0; mb < nmb; ++mb)
                    for (size_t mk = 0; mk < nmk; ++mk) {
                        const int64_t i = bra.pair_index[mb], j = ket.pair_index[mk];
                        if (I == J && i < j) continue;   // the same quartet as (mk, mb)
                        const int64_t hi = std::max(i, j), lo = std::min(i, j);
                        if (!any || bra.Q[mb] * ket.Q[mk] < cutoff) {
                            sink(size_t(hi), size_t(lo), static_cast<const double*>(nullptr));
                            continue;
                        }
                        const Shell& sa = m.shells[size_t(bra.members[mb].first)];
                        const Shell& sb = m.shells[size_t(bra.members[mb].second)];
                        const Shell& sc = m.shells[size_t(ket.members[mk].first)];
                        const Shell& sd = m.shells[size_t(ket.members[mk].second)];
                        int dims[4] = {ncart(sa.L), ncart(sb.L), ncart(sc.L), ncart(sd.L)};
                        const double* o = w.out.data() + (mb * nmk + mk) * blk;
                        w.ao.assign(o, o + blk);
                        transform_index(w.ao, dims, 0, sa, tmp);
                        transform_index(w.ao, dims, 1, sb, tmp);
                        transform_index(w.ao, dims, 2, sc, tmp);
                        transform_index(w.ao, dims, 3, sd, tmp);
                        if (i >= j) {
                            sink(size_t(i), size_t(j), static_cast<const double*>(w.ao.data()));
                        } else {
                            // bra and ket exchanged: [fc][fd][fa][fb]
                            const size_t nb2 = size_t(dims[0]) * size_t(dims[1]), nk2 = size_t(dims[2]) * size_t(dims[3]);
                            w.half.resize(nb2 * nk2);
                            for (size_t x = 0; x < nb2; ++x)
                                for (size_t y = 0; y < nk2; ++y) w.half[y * nb2 + x] = w.ao[x * nk2 + y];
                            sink(size_t(j), size_t(i), static_cast<const double*>(w.half.data()));
                        }
                    }
            }
        },
        nfp);
}

void ints_2e(const MoleculeData& m, int nthreads, double* eri) {
    const size_t N = size_t(m.nao);
    for_each_quartet(m, nthreads, [&](size_t i, size_t j, const double* blk) {
        int ia, ib, ic, id;
        unpair(int64_t(i), ia, ib);
        unpair(int64_t(j), ic, id);
        const Shell& sa = m.shells[size_t(ia)];
        const Shell& sb = m.shells[size_t(ib)];
        const Shell& sc = m.shells[size_t(ic)];
        const Shell& sd = m.shells[size_t(id)];
        const int d0 = sa.nf, d1 = sb.nf, d2 = sc.nf, d3 = sd.nf;
        for (int f0 = 0; f0 < d0; ++f0)
            for (int f1 = 0; f1 < d1; ++f1)
                for (int f2 = 0; f2 < d2; ++f2)
                    for (int f3 = 0; f3 < d3; ++f3) {
                        const double x =
                            blk ? blk[((size_t(f0) * size_t(d1) + size_t(f1)) * size_t(d2) + size_t(f2)) * size_t(d3) +
                                      size_t(f3)]
                                : 0.0;
                        const size_t M = size_t(sa.ao0 + f0), Nn = size_t(sb.ao0 + f1), Lc = size_t(sc.ao0 + f2),
                                     S = size_t(sd.ao0 + f3);
                        auto put = [&](size_t p, size_t q, size_t r2, size_t s2) {
                            eri[p + N * (q + N * (r2 + N * s2))] = x;
                        };
                        put(M, Nn, Lc, S);
                        put(Nn, M, Lc, S);
                        put(M, Nn, S, Lc);
                        put(Nn, M, S, Lc);
                        put(Lc, S, M, Nn);
                        put(S, Lc, M, Nn);
                        put(Lc, S, Nn, M);
                        put(S, Lc, Nn, M);
                    }
    });
}

CompactEri compact_layout(const MoleculeData& m) {
    CompactEri c;
    const int ns = int(m.shells.size());
    const size_t npair = size_t(ns) * size_t(ns + 1) / 2;
    c.pair_size.resize(npair);
    c.prefix.resize(npair);
    c.rowstart.resize(npair);
    for (int a = 0, idx = 0; a < ns; ++a)
        for (int b = 0; b <= a; ++b, ++idx)
            c.pair_size[size_t(idx)] = int64_t(m.shells[size_t(a)].nf) * int64_t(m.shells[size_t(b)].nf);
    int64_t pre = 0, row = 0;
    for (size_t i = 0; i < npair; ++i) {
        c.prefix[i] = pre;                       // sum of pair sizes before i
        pre += c.pair_size[i];
        c.rowstart[i] = row;                     // blocks (i', j' <= i') for i' < i
        row += c.pair_size[i] * pre;             // pre now includes pair i
    }
    c.total = row;
    AQ_CHECK(npair == 0 || c.rowstart[npair - 1] + c.pair_size[npair - 1] * (c.prefix[npair - 1] + c.pair_size[npair - 1]) == c.total,
             AQ_ERR_INVALID, "inconsistent compact ERI layout");
    c.shell_of.resize(size_t(m.nao));
    c.fn_of.resize(size_t(m.nao));
    for (int s = 0; s < ns; ++s)
        for (int f = 0; f < m.shells[size_t(s)].nf; ++f) {
            c.shell_of[size_t(m.shells[size_t(s)].ao0 + f)] = s;
            c.fn_of[size_t(m.shells[size_t(s)].ao0 + f)] = f;
        }
    c.nf.resize(size_t(ns));
    for (int s = 0; s < ns; ++s) c.nf[size_t(s)] = m.shells[size_t(s)].nf;
    return c;
}

void ints_2e_compact(const MoleculeData& m, int nthreads, const CompactEri& c, double* out) {
    for_each_quartet(m, nthreads, [&](size_t i, size_t j, const double* blk) {
        const size_t n = size_t(c.pair_size[i]) * size_t(c.pair_size[j]);
        double* dst = out + c.rowstart[i] + c.pair_size[i] * c.prefix[j];
        if (blk) std::memcpy(dst, blk, n * sizeof(double));
        else std::memset(dst, 0, n * sizeof(double));
    });
}

}  // namespace fe
}  // namespace aqamd
