// Integral front end: GENBAS basis sets, Gaussian shells, one- and
// two-electron integrals (host), closed-shell SCF and the AO->MO step that
// produces the CCSD engine's input blocks.  Not part of the C ABI.
#pragma once

#include <array>
#include <cstdint>
#include <map>
#include <string>
#include <vector>

namespace aqamd {
namespace fe {

constexpr int kMaxL = 4;                       // up to g functions
constexpr int kMaxM = 4 * kMaxL + 8;           // Boys orders kept in the table

// One angular-momentum block of a GENBAS entry (basis.cxx:15-75):
// coef[j * ncontr + k] is primitive j of contraction k (file order).
struct ShellBasis {
    int L = 0, nprim = 0, ncontr = 0;
    std::vector<double> exps, coef;
};

std::map<std::string, std::vector<ShellBasis>> parse_genbas(const std::string& text);
const char* element_symbol(int Z);

inline int ncart(int L) { return (L + 1) * (L + 2) / 2; }
// Cartesian components of a shell: x power descending, then y descending.
const std::vector<std::array<int, 3>>& cart_powers(int L);
// (2L+1) x ncart(L) row-major map from cartesian monomials to real solid
// harmonics m = -L..L, each normalised over axially normalised Gaussians.
const std::vector<double>& sph_transform(int L);

// One segmented contraction on one centre.
struct Shell {
    int atom = 0, L = 0;
    double A[3] = {0, 0, 0};
    std::vector<double> a;        // exponents
    std::vector<double> c;        // coefficients x primitive normalisation (x^L)
    int ao0 = 0, nf = 0;          // first AO, number of functions
    const double* T = nullptr;    // nf x ncart(L) to AO (nullptr: identity)
};

struct MoleculeData {
    std::vector<int> Z;
    std::vector<double> xyz;      // bohr, 3 per atom
    int charge = 0;
    bool spherical = true;
    std::vector<Shell> shells;
    int nao = 0, lmax = 0;
    double nuclear_repulsion() const;
    int nelectrons() const;
};

void build_shells(MoleculeData& m, const std::map<std::string, std::vector<ShellBasis>>& basis);

// F_m(T), m = 0..mmax, into F.
void boys(int mmax, double T, double* F);

// S, T, V: nao x nao (symmetric).  eri[m + n N + l N^2 + s N^3] = (mn|ls).
void ints_1e(const MoleculeData& m, double* S, double* T, double* V);
void ints_2e(const MoleculeData& m, int nthreads, double* eri);

// Compact storage of the canonical shell quartets: bra pair i = (a >= b)
// index a(a+1)/2 + b, ket pair j <= i; block (i, j) of pair_size[i] *
// pair_size[j] values [fa][fb][fc][fd] starts at rowstart[i] + pair_size[i] *
// prefix[j].  shell_of / fn_of map an AO to its shell and its function within.
struct CompactEri {
    std::vector<int64_t> pair_size, prefix, rowstart;
    std::vector<int> shell_of, fn_of, nf;
    int64_t total = 0;
};
CompactEri compact_layout(const MoleculeData& m);
void ints_2e_compact(const MoleculeData& m, int nthreads, const CompactEri& c, double* out);

// Symmetric eigenproblem A = U diag(w) U^T (cyclic Jacobi), n x n, A overwritten;
// eigenvalues ascending, U column-major (column k = eigenvector k).
void eigh(int n, std::vector<double>& A, std::vector<double>& w, std::vector<double>& U);

}  // namespace fe
}  // namespace aqamd

struct aq_molecule {
    aqamd::fe::MoleculeData d;
};
