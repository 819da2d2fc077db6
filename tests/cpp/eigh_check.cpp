// Checks the host eigensolver of the SCF (aqamd::fe::eigh): residual
// |A u - w u|, orthogonality and ascending order on random symmetric matrices
// of several sizes, with and without a degenerate spectrum.  Exit status =
// number of failed cases.
#include "../../aquarius_amd/csrc/frontend.hpp"
#include <cmath>
#include <cstdio>
#include <random>
#include <vector>
int main() {
    std::mt19937_64 rng(5);
    std::normal_distribution<double> N01;
    int fails = 0;
    for (int n : {1, 2, 3, 7, 24, 57, 192}) {
        for (int deg = 0; deg < 2; ++deg) {
            std::vector<double> A(n * n);
            for (int i = 0; i < n; ++i)
                for (int j = 0; j <= i; ++j) A[i * n + j] = A[j * n + i] = N01(rng);
            if (deg && n >= 4) {   // degenerate spectrum: Q diag(1,1,2,2,...) Q^T via a reflection
                std::vector<double> v(n); double vv = 0;
                for (auto& x : v) { x = N01(rng); vv += x * x; }
                for (int i = 0; i < n; ++i)
                    for (int j = 0; j < n; ++j) {
                        double s = 0;
                        for (int k = 0; k < n; ++k) {
                            const double hik = (i == k) - 2 * v[i] * v[k] / vv, hjk = (j == k) - 2 * v[j] * v[k] / vv;
                            s += hik * double(1 + k / 2) * hjk;
                        }
                        A[i * n + j] = s;
                    }
            }
            std::vector<double> A0 = A, w, U;
            aqamd::fe::eigh(n, A, w, U);
            double res = 0, orth = 0, scale = 0;
            for (double x : A0) scale = std::max(scale, std::fabs(x));
            for (int i = 0; i < n; ++i)
                for (int k = 0; k < n; ++k) {
                    double s = 0;
                    for (int j = 0; j < n; ++j) s += A0[i * n + j] * U[j * n + k];
                    res = std::max(res, std::fabs(s - w[k] * U[i * n + k]));
                    double o = 0;
                    for (int j = 0; j < n; ++j) o += U[j * n + i] * U[j * n + k];
                    orth = std::max(orth, std::fabs(o - (i == k)));
                }
            bool sorted = true;
            for (int k = 1; k < n; ++k) sorted &= w[k - 1] <= w[k];
            const bool ok = res < 1e-12 * std::max(1.0, scale) * n && orth < 1e-12 * n && sorted;
            fails += !ok;
            printf("n=%3d deg=%d residual %.2e orth %.2e sorted %d %s\n", n, deg, res, orth, int(sorted), ok ? "ok" : "FAIL");
        }
    }
    return fails;
}
