// GENBAS reader, shell construction and the cartesian -> pure-harmonic map.
//
// Reference: BasisSet::readBasisSet (src/input/basis.cxx:15-75) for the file
// format, Shell (src/integrals/shell.cxx:114-160) for the contraction
// normalisation and the use of pure harmonics (the reference default,
// src/input/molecule.cxx:1364-1367).  Energies do not depend on the order or
// the normalisation of the AO functions; both are this library's own.
#include "frontend.hpp"
#include "internal.hpp"

#include <algorithm>
#include <cctype>
#include <cmath>
#include <deque>
#include <mutex>
#include <sstream>

namespace aqamd {
namespace fe {

namespace {

const char* kSymbols[] = {"H",  "He", "Li", "Be", "B",  "C",  "N",  "O",  "F",  "Ne", "Na", "Mg",
                          "Al", "Si", "P",  "S",  "Cl", "Ar", "K",  "Ca", "Sc", "Ti", "V",  "Cr",
                          "Mn", "Fe", "Co", "Ni", "Cu", "Zn", "Ga", "Ge", "As", "Se", "Br", "Kr"};
constexpr int kNumSymbols = int(sizeof(kSymbols) / sizeof(kSymbols[0]));

std::string upper(std::string s) {
    for (auto& ch : s) ch = char(std::toupper(static_cast<unsigned char>(ch)));
    return s;
}

std::string trim(const std::string& s) {
    const size_t b = s.find_first_not_of(" \t\r");
    if (b == std::string::npos) return "";
    return s.substr(b, s.find_last_not_of(" \t\r") - b + 1);
}

double to_double(std::string s, int lineno) {
    for (auto& ch : s)
        if (ch == 'D' || ch == 'd') ch = 'E';
    size_t pos = 0;
    double v = 0.0;
    try {
        v = std::stod(s, &pos);
    } catch (...) {
        pos = 0;
    }
    AQ_CHECK(pos == s.size() && pos > 0, AQ_ERR_INVALID,
             "GENBAS line " + std::to_string(lineno) + ": bad number '" + s + "'");
    return v;
}

double dfact(int n) {
    double r = 1.0;
    for (int k = n; k > 1; k -= 2) r *= k;
    return r;
}

double binom(int n, int k) {
    if (k < 0 || k > n) return 0.0;
    double r = 1.0;
    for (int i = 1; i <= k; ++i) r = r * (n - k + i) / i;
    return r;
}

}  // namespace

const char* element_symbol(int Z) { return Z >= 1 && Z <= kNumSymbols ? kSymbols[Z - 1] : nullptr; }

std::map<std::string, std::vector<ShellBasis>> parse_genbas(const std::string& text) {
    std::vector<std::string> lines;
    {
        std::istringstream is(text);
        std::string l;
        while (std::getline(is, l)) lines.push_back(l);
    }
    std::map<std::string, std::vector<ShellBasis>> out;
    size_t i = 0;
    while (i < lines.size()) {
        const std::string& l = lines[i++];
        if (trim(l).empty() || l[0] == '!') continue;
        const size_t sep = l.find(':');
        AQ_CHECK(sep != std::string::npos && sep >= 1 && sep <= 2, AQ_ERR_INVALID,
                 "GENBAS line " + std::to_string(i) + ": element header 'X:NAME' expected");
        const std::string elem = upper(trim(l.substr(0, sep)));
        ++i;  // comment line
        std::deque<std::string> tok;
        auto take = [&](size_t n) {
            while (tok.size() < n) {
                AQ_CHECK(i < lines.size(), AQ_ERR_INVALID, "GENBAS: premature end of file in " + elem);
                std::istringstream ls(lines[i++]);
                std::string t;
                while (ls >> t) tok.push_back(t);
            }
            std::vector<double> v;
            for (size_t k = 0; k < n; ++k) {
                v.push_back(to_double(tok.front(), int(i)));
                tok.pop_front();
            }
            return v;
        };
        const int nshell = int(take(1)[0]);
        AQ_CHECK(nshell >= 1 && nshell <= 16, AQ_ERR_INVALID, "GENBAS: bad shell count for " + elem);
        const auto Ls = take(size_t(nshell));
        const auto nc = take(size_t(nshell));
        const auto np = take(size_t(nshell));
        std::vector<ShellBasis> shells(static_cast<size_t>(nshell));
        for (int s = 0; s < nshell; ++s) {
            ShellBasis& b = shells[size_t(s)];
            b.L = int(Ls[size_t(s)]);
            b.ncontr = int(nc[size_t(s)]);
            b.nprim = int(np[size_t(s)]);
            AQ_CHECK(b.L >= 0 && b.L <= kMaxL && b.ncontr >= 1 && b.nprim >= 1, AQ_ERR_UNSUPPORTED,
                     "GENBAS: " + elem + " shell L=" + std::to_string(b.L) + " not supported (L <= 4)");
            b.exps = take(size_t(b.nprim));
            b.coef = take(size_t(b.nprim) * size_t(b.ncontr));
        }
        AQ_CHECK(tok.empty(), AQ_ERR_INVALID, "GENBAS: trailing numbers in " + elem);
        out[elem] = std::move(shells);
    }
    return out;
}

const std::vector<std::array<int, 3>>& cart_powers(int L) {
    static std::vector<std::array<int, 3>> tab[kMaxL + 1];
    static std::once_flag once;
    std::call_once(once, [] {
        for (int l = 0; l <= kMaxL; ++l)
            for (int x = l; x >= 0; --x)
                for (int y = l - x; y >= 0; --y) tab[l].push_back({x, y, l - x - y});
    });
    return tab[L];
}

// Real solid harmonics (Helgaker, Joergensen, Olsen, "Molecular
// Electronic-Structure Theory", eqs. 6.4.47-6.4.48) over the monomials, then
// normalised with the overlap of axially normalised cartesian Gaussians of one
// exponent: <c|c'> = prod_d (p_d + p'_d - 1)!! / (2L - 1)!!.
const std::vector<double>& sph_transform(int L) {
    static std::vector<double> tab[kMaxL + 1];
    static std::once_flag once;
    std::call_once(once, [] {
        for (int l = 0; l <= kMaxL; ++l) {
            const auto& P = cart_powers(l);
            const int nc = ncart(l);
            auto col = [&](int x, int y, int z) {
                for (int k = 0; k < nc; ++k)
                    if (P[size_t(k)][0] == x && P[size_t(k)][1] == y && P[size_t(k)][2] == z) return k;
                return -1;
            };
            std::vector<double> M(size_t(2 * l + 1) * size_t(nc), 0.0);
            for (int m = -l; m <= l; ++m) {
                const int am = std::abs(m);
                const int twovm = m < 0 ? 1 : 0;   // 2 v_m
                double* row = M.data() + size_t(m + l) * size_t(nc);
                for (int t = 0; t <= (l - am) / 2; ++t)
                    for (int u = 0; u <= t; ++u)
                        // v = v_m, v_m + 1, ... ; 2v runs over the integers of the parity of 2 v_m
                        for (int twov = twovm; twov <= am; twov += 2) {
                            const int sgn = ((t + (twov - twovm) / 2) & 1) ? -1 : 1;
                            const double c = sgn * std::pow(0.25, t) * binom(l, t) * binom(l - t, am + t) *
                                             binom(t, u) * binom(am, twov);
                            const int py = 2 * u + twov;
                            const int px = 2 * t + am - py;
                            const int pz = l - 2 * t - am;
                            if (px < 0) continue;
                            const int k = col(px, py, pz);
                            if (k >= 0) row[k] += c;
                        }
                double nrm = 0.0;
                for (int a = 0; a < nc; ++a)
                    for (int b = 0; b < nc; ++b) {
                        const int sx = P[size_t(a)][0] + P[size_t(b)][0], sy = P[size_t(a)][1] + P[size_t(b)][1],
                                  sz = P[size_t(a)][2] + P[size_t(b)][2];
                        if ((sx | sy | sz) & 1) continue;
                        nrm += row[a] * row[b] * dfact(sx - 1) * dfact(sy - 1) * dfact(sz - 1) / dfact(2 * l - 1);
                    }
                for (int a = 0; a < nc; ++a) row[a] /= std::sqrt(nrm);
            }
            tab[l] = std::move(M);
        }
    });
    return tab[L];
}

double MoleculeData::nuclear_repulsion() const {
    double e = 0.0;
    for (size_t i = 0; i < Z.size(); ++i)
        for (size_t j = 0; j < i; ++j) {
            const double dx = xyz[3 * i] - xyz[3 * j], dy = xyz[3 * i + 1] - xyz[3 * j + 1],
                         dz = xyz[3 * i + 2] - xyz[3 * j + 2];
            e += Z[i] * Z[j] / std::sqrt(dx * dx + dy * dy + dz * dz);
        }
    return e;
}

int MoleculeData::nelectrons() const {
    int n = -charge;
    for (int z : Z) n += z;
    return n;
}

// Each contraction becomes one segmented shell over its nonzero primitives,
// normalised over normalised primitives (the normalisation of
// shell.cxx:114-134); the primitive normalisation (2a/pi)^(3/4) (4a)^(L/2) /
// sqrt((2L-1)!!) of the x^L component is folded into the coefficients.
void build_shells(MoleculeData& m, const std::map<std::string, std::vector<ShellBasis>>& basis) {
    m.shells.clear();
    m.nao = 0;
    m.lmax = 0;
    for (size_t at = 0; at < m.Z.size(); ++at) {
        const char* sym = element_symbol(m.Z[at]);
        AQ_CHECK(sym, AQ_ERR_INVALID, "unsupported atomic number " + std::to_string(m.Z[at]));
        auto it = basis.find(upper(sym));
        AQ_CHECK(it != basis.end(), AQ_ERR_INVALID, std::string("basis set has no entry for ") + sym);
        for (const ShellBasis& b : it->second) {
            for (int k = 0; k < b.ncontr; ++k) {
                Shell s;
                s.atom = int(at);
                s.L = b.L;
                for (int d = 0; d < 3; ++d) s.A[d] = m.xyz[3 * at + size_t(d)];
                for (int j = 0; j < b.nprim; ++j) {
                    const double c = b.coef[size_t(j) * size_t(b.ncontr) + size_t(k)];
                    if (c == 0.0) continue;
                    s.a.push_back(b.exps[size_t(j)]);
                    s.c.push_back(c);
                }
                AQ_CHECK(!s.a.empty(), AQ_ERR_INVALID, std::string("empty contraction in the basis of ") + sym);
                double nrm = 0.0;
                for (size_t i = 0; i < s.a.size(); ++i)
                    for (size_t j = 0; j < s.a.size(); ++j)
                        nrm += s.c[i] * s.c[j] *
                               std::pow(2.0 * std::sqrt(s.a[i] * s.a[j]) / (s.a[i] + s.a[j]), b.L + 1.5);
                const double inv = 1.0 / std::sqrt(nrm);
                for (size_t i = 0; i < s.a.size(); ++i)
                    s.c[i] *= inv * std::pow(2.0 * s.a[i] / M_PI, 0.75) * std::pow(4.0 * s.a[i], 0.5 * b.L) /
                              std::sqrt(dfact(2 * b.L - 1));
                s.nf = m.spherical ? 2 * b.L + 1 : ncart(b.L);
                s.T = m.spherical && b.L >= 1 ? sph_transform(b.L).data() : nullptr;
                s.ao0 = m.nao;
                m.nao += s.nf;
                m.lmax = std::max(m.lmax, b.L);
                m.shells.push_back(std::move(s));
            }
        }
    }
}

}  // namespace fe
}  // namespace aqamd
