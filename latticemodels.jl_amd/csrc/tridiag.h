// Host helpers of the Krylov ket propagator: the exponential of a small real
// symmetric tridiagonal matrix (the Lanczos matrix T_j) applied to e_1.
#pragma once

#include <cmath>
#include <complex>
#include <utility>
#include <vector>

namespace lmh {

// Eigen-decomposition T = Z diag(d) Z^T of the symmetric tridiagonal matrix
// with diagonal d[0..n) and off-diagonal e[0..n-1) (e[i] = T[i+1][i]) by the
// implicit QL method with Wilkinson shifts (Bowdler, Martin, Reinsch and
// Wilkinson, "The QR and QL algorithms for symmetric matrices", Numer. Math.
// 11 (1968)).  On return d holds the eigenvalues, z (n x n, z[i * n + k] =
// component i of eigenvector k) the eigenvectors.  False if an eigenvalue
// needs more than 60 sweeps.  ENDS: only rows 0 and n - 1 of z (the first
// and last components of every eigenvector, z[k] and z[n + k]); a rotation
// updates each row from that row alone, so those rows and d are the same bit
// for bit as with all n rows, in O(n^2) instead of O(n^3).
template <bool ENDS>
inline bool tridiag_eigen_t(int n, std::vector<double> &d, std::vector<double> e, std::vector<double> &z) {
    const int rows = ENDS ? (n > 0 ? 2 : 0) : n;
    z.assign((size_t)rows * n, 0.0);
    if (ENDS) {
        if (n > 0) {
            z[0] = 1.0;
            z[(size_t)n + n - 1] = 1.0;
        }
    } else {
        for (int i = 0; i < n; ++i) z[(size_t)i * n + i] = 1.0;
    }
    e.resize(n, 0.0);
    if (n > 0) e[n - 1] = 0.0;
    for (int l = 0; l < n; ++l) {
        int sweeps = 0;
        for (;;) {
            // the first negligible off-diagonal at or after l splits the matrix
            int m = l;
            for (; m < n - 1; ++m) {
                const double dd = std::fabs(d[m]) + std::fabs(d[m + 1]);
                if (std::fabs(e[m]) + dd == dd) break;
            }
            if (m == l) break;
            if (++sweeps > 60) return false;
            // Wilkinson shift from the leading 2 x 2 block
            double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
            double r = std::hypot(g, 1.0);
            g = d[m] - d[l] + e[l] / (g + (g >= 0.0 ? r : -r));
            double s = 1.0, c = 1.0, p = 0.0;
            int i = m - 1;
            bool deflated = false;
            for (; i >= l; --i) {
                double f = s * e[i];
                const double b = c * e[i];
                r = std::hypot(f, g);
                e[i + 1] = r;
                if (r == 0.0) {
                    // a zero rotation: the block splits here
                    d[i + 1] -= p;
                    e[m] = 0.0;
                    deflated = true;
                    break;
                }
                s = f / r;
                c = g / r;
                g = d[i + 1] - p;
                r = (d[i] - g) * s + 2.0 * c * b;
                p = s * r;
                d[i + 1] = g + p;
                g = c * r - b;
                for (int k = 0; k < rows; ++k) {
                    f = z[(size_t)k * n + i + 1];
                    z[(size_t)k * n + i + 1] = s * z[(size_t)k * n + i] + c * f;
                    z[(size_t)k * n + i] = c * z[(size_t)k * n + i] - s * f;
                }
            }
            if (deflated) continue;
            d[l] -= p;
            e[l] = g;
            e[m] = 0.0;
        }
    }
    return true;
}

inline bool tridiag_eigen(int n, std::vector<double> &d, std::vector<double> e, std::vector<double> &z) {
    return tridiag_eigen_t<false>(n, d, std::move(e), z);
}

// eigenvalues and rows 0 and n - 1 of the eigenvectors (z is 2 x n)
inline bool tridiag_eigen_ends(int n, std::vector<double> &d, std::vector<double> e, std::vector<double> &z) {
    return tridiag_eigen_t<true>(n, d, std::move(e), z);
}

// y = scale * exp(-i tau T) e_1 from the eigen-decomposition of T
inline void tridiag_expv(int n, const std::vector<double> &lam, const std::vector<double> &z, double tau, double scale,
                         std::complex<double> *y) {
    for (int i = 0; i < n; ++i) y[i] = 0.0;
    for (int k = 0; k < n; ++k) {
        const double a = -tau * lam[k];
        const std::complex<double> w = std::complex<double>(std::cos(a), std::sin(a)) * (scale * z[k]);
        for (int i = 0; i < n; ++i) y[i] += w * z[(size_t)i * n + k];
    }
}

// component n - 1 of tridiag_expv's y, from tridiag_eigen_ends' two rows (the
// same sum in the same order)
inline std::complex<double> tridiag_expv_last(int n, const std::vector<double> &lam, const std::vector<double> &z2,
                                              double tau, double scale) {
    std::complex<double> y = 0.0;
    for (int k = 0; k < n; ++k) {
        const double a = -tau * lam[k];
        const std::complex<double> w = std::complex<double>(std::cos(a), std::sin(a)) * (scale * z2[k]);
        y += w * z2[(size_t)n + k];
    }
    return y;
}

}  // namespace lmh
