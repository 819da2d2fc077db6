// FCIDUMP interchange (the Knowles-Handy integral file).
//
// Reader semantics follow the reference's FCIDUMP task
// (src/operator/fcidump.cxx:43-517): the header runs up to a line holding
// "/", "&END" or "$END" (any case) and gives NORB and NELEC; each later line
// is "value p q r s" with 1-based orbital indices:
//   p = 0            core energy (accumulated),
//   q = 0            orbital energy (ignored),
//   r = 0            one-electron h_pq (accumulated; mirrored to h_qp unless
//                    the file lists both triangles, the reference's
//                    "1eints full"),
//   otherwise        the Mulliken integral (pq|rs); with p >= q, r >= s a line
//                    whose pair (p,q) precedes (r,s) is skipped, as there, and
//                    the value fills all eight symmetric positions.
// The writer emits that canonical order: two-electron integrals with
// i >= j, k >= l, ij >= kl above a threshold, then h_ij (i >= j), then the
// core energy.
#include "internal.hpp"

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <fstream>
#include <sstream>
#include <string>

using namespace aqamd;

namespace {

std::string upper(std::string s) {
    for (auto& c : s) c = char(std::toupper(static_cast<unsigned char>(c)));
    return s;
}

// integer after "KEY=" (spaces allowed around '='); false if absent
bool header_int(const std::string& hdr, const std::string& key, long& out) {
    size_t pos = 0;
    while ((pos = hdr.find(key, pos)) != std::string::npos) {
        // the key must start a word (NORB vs e.g. XNORB)
        if (pos > 0 && (std::isalnum(static_cast<unsigned char>(hdr[pos - 1])) || hdr[pos - 1] == '_')) {
            pos += key.size();
            continue;
        }
        size_t i = pos + key.size();
        while (i < hdr.size() && std::isspace(static_cast<unsigned char>(hdr[i]))) ++i;
        if (i < hdr.size() && hdr[i] == '=') {
            ++i;
            while (i < hdr.size() && std::isspace(static_cast<unsigned char>(hdr[i]))) ++i;
            size_t j = i;
            if (j < hdr.size() && (hdr[j] == '-' || hdr[j] == '+')) ++j;
            while (j < hdr.size() && std::isdigit(static_cast<unsigned char>(hdr[j]))) ++j;
            if (j > i) {
                out = std::stol(hdr.substr(i, j - i));
                return true;
            }
        }
        pos += key.size();
    }
    return false;
}

struct Header {
    long norb = 0, nelec = 0, ms2 = 0;
};

Header read_header(std::istream& is, const std::string& path) {
    std::string hdr, line;
    bool ended = false;
    while (std::getline(is, line)) {
        hdr += " " + line;
        const std::string u = upper(line);
        if (u.find('/') != std::string::npos || u.find("&END") != std::string::npos ||
            u.find("$END") != std::string::npos) {
            ended = true;
            break;
        }
    }
    AQ_CHECK(ended, AQ_ERR_INVALID, path + ": FCIDUMP header has no &END");
    const std::string u = upper(hdr);
    Header h;
    AQ_CHECK(header_int(u, "NORB", h.norb) && h.norb > 0, AQ_ERR_INVALID, path + ": FCIDUMP header lacks NORB");
    AQ_CHECK(header_int(u, "NELEC", h.nelec) && h.nelec >= 0, AQ_ERR_INVALID, path + ": FCIDUMP header lacks NELEC");
    header_int(u, "MS2", h.ms2);
    return h;
}

}  // namespace

extern "C" {

int aq_fcidump_header(const char* path, int* norb, int* nelec, int* ms2) {
    return guarded([&] {
        AQ_CHECK(path, AQ_ERR_INVALID, "null path");
        std::ifstream is(path);
        AQ_CHECK(is.good(), AQ_ERR_INVALID, std::string("cannot open ") + path);
        const Header h = read_header(is, path);
        if (norb) *norb = int(h.norb);
        if (nelec) *nelec = int(h.nelec);
        if (ms2) *ms2 = int(h.ms2);
    });
}

int aq_fcidump_read(const char* path, int full_1e, double* h, double* eri, double* ecore) {
    return guarded([&] {
        AQ_CHECK(path && h && eri && ecore, AQ_ERR_INVALID, "null argument");
        std::ifstream is(path);
        AQ_CHECK(is.good(), AQ_ERR_INVALID, std::string("cannot open ") + path);
        const Header hd = read_header(is, path);
        const size_t n = size_t(hd.norb);
        std::fill(h, h + n * n, 0.0);
        std::fill(eri, eri + n * n * n * n, 0.0);
        *ecore = 0.0;
        auto E = [&](size_t p, size_t q, size_t r, size_t s) -> double& { return eri[p + n * (q + n * (r + n * s))]; };
        std::string line;
        long lineno = 0;
        while (std::getline(is, line)) {
            ++lineno;
            std::istringstream ls(line);
            std::string vtok;
            if (!(ls >> vtok)) continue;   // blank line
            for (auto& c : vtok)
                if (c == 'D' || c == 'd') c = 'E';
            double val = 0.0;
            long p = 0, q = 0, r = 0, s = 0;
            try {
                val = std::stod(vtok);
            } catch (...) {
                throw Error(AQ_ERR_INVALID, std::string(path) + ": bad value on integral line " + std::to_string(lineno));
            }
            AQ_CHECK(bool(ls >> p >> q >> r >> s), AQ_ERR_INVALID,
                     std::string(path) + ": integral line " + std::to_string(lineno) + " needs four indices");
            const long N = hd.norb;
            AQ_CHECK(p >= 0 && q >= 0 && r >= 0 && s >= 0 && p <= N && q <= N && r <= N && s <= N, AQ_ERR_INVALID,
                     std::string(path) + ": orbital index out of range on line " + std::to_string(lineno));
            if (p == 0) {
                *ecore += val;
            } else if (q == 0) {
                continue;
            } else if (r == 0) {
                h[size_t(p - 1) + n * size_t(q - 1)] += val;
                if (p != q && !full_1e) h[size_t(q - 1) + n * size_t(p - 1)] += val;
            } else {
                if (p < q) std::swap(p, q);
                if (r < s) std::swap(r, s);
                if (p < r || (p == r && q < s)) continue;
                const size_t a = size_t(p - 1), b = size_t(q - 1), c = size_t(r - 1), d = size_t(s - 1);
                E(a, b, c, d) = E(b, a, c, d) = E(a, b, d, c) = E(b, a, d, c) = val;
                E(c, d, a, b) = E(d, c, a, b) = E(c, d, b, a) = E(d, c, b, a) = val;
            }
        }
    });
}

int aq_fcidump_write(const char* path, int norb, int nelec, int ms2, const double* h, const double* eri, double ecore,
                     double threshold) {
    return guarded([&] {
        AQ_CHECK(path && h && eri && norb > 0 && nelec >= 0, AQ_ERR_INVALID, "bad argument");
        FILE* f = std::fopen(path, "w");
        AQ_CHECK(f, AQ_ERR_INVALID, std::string("cannot create ") + path);
        const size_t n = size_t(norb);
        std::fprintf(f, " &FCI NORB=%d,NELEC=%d,MS2=%d,\n  ORBSYM=", norb, nelec, ms2);
        for (int i = 0; i < norb; ++i) std::fprintf(f, "1,");
        std::fprintf(f, "\n  ISYM=1,\n &END\n");
        for (size_t i = 0; i < n; ++i)
            for (size_t j = 0; j <= i; ++j)
                for (size_t k = 0; k <= i; ++k)
                    for (size_t l = 0; l <= (k == i ? j : k); ++l) {
                        const double v = eri[i + n * (j + n * (k + n * l))];
                        if (std::fabs(v) >= threshold)
                            std::fprintf(f, "%24.16E %4zu %4zu %4zu %4zu\n", v, i + 1, j + 1, k + 1, l + 1);
                    }
        for (size_t i = 0; i < n; ++i)
            for (size_t j = 0; j <= i; ++j) {
                const double v = h[i + n * j];
                if (std::fabs(v) >= threshold) std::fprintf(f, "%24.16E %4zu %4zu %4d %4d\n", v, i + 1, j + 1, 0, 0);
            }
        std::fprintf(f, "%24.16E %4d %4d %4d %4d\n", ecore, 0, 0, 0, 0);
        const bool ok = std::ferror(f) == 0;
        std::fclose(f);
        AQ_CHECK(ok, AQ_ERR_INVALID, std::string("write error on ") + path);
    });
}

}  // extern "C"
