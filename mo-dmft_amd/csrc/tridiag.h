// Host-side dense eigensolvers for the small matrices of the Lanczos method.
//
//  * tridiag_ql: symmetric tridiagonal eigenproblem by the implicit QL method with Wilkinson
//    shifts (EISPACK tql2; Press et al., Numerical Recipes 2nd ed., section 11.3), accumulating the
//    rotations into the rows of Z that the caller asks for (all rows, or only the last row for
//    residual estimates).  Eigenvalues are returned ascending with Z columns permuted to match.
//  * jacobi_eig: cyclic Jacobi method for a small dense symmetric matrix (Golub & Van Loan 8.5).
#pragma once
#include <algorithm>
#include <cmath>
#include <numeric>
#include <vector>

namespace moed {

// d[0..m-1] diagonal, e[0..m-2] off-diagonal.  On return d holds eigenvalues ascending.
// If zrows > 0, Z is zrows x m (row-major) and on input holds the selected rows of the identity
// (row r = e_{rowsel[r]}); rotations are applied to every row.  Returns false if not converged.
inline bool tridiag_ql(std::vector<double>& d, std::vector<double> e, int zrows, std::vector<double>* Z) {
    const int m = (int)d.size();
    e.resize(m, 0.0);
    e[m - 1] = 0.0;
    for (int l = 0; l < m; ++l) {
        int iter = 0;
        while (true) {
            int mm = l;
            for (; mm < m - 1; ++mm) {
                const double dd = std::fabs(d[mm]) + std::fabs(d[mm + 1]);
                if (std::fabs(e[mm]) <= 1e-300 || std::fabs(e[mm]) <= 2.220446049250313e-16 * dd) break;
            }
            if (mm == l) break;
            if (++iter > 60) return false;
            // Wilkinson shift from the 2x2 block at l
            double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
            double r = std::hypot(g, 1.0);
            g = d[mm] - d[l] + e[l] / (g + (g >= 0 ? r : -r));
            double s = 1.0, c = 1.0, p = 0.0;
            int i = mm - 1;
            bool underflow = false;
            for (; i >= l; --i) {
                double f = s * e[i];
                const double b = c * e[i];
                r = std::hypot(f, g);
                e[i + 1] = r;
                if (r == 0.0) {
                    d[i + 1] -= p;
                    e[mm] = 0.0;
                    underflow = true;
                    break;
                }
                s = f / r;
                c = g / r;
                g = d[i + 1] - p;
                r = (d[i] - g) * s + 2.0 * c * b;
                p = s * r;
                d[i + 1] = g + p;
                g = c * r - b;
                if (zrows > 0) {
                    double* z = Z->data();
                    for (int k = 0; k < zrows; ++k) {
                        f = z[(size_t)k * m + i + 1];
                        z[(size_t)k * m + i + 1] = s * z[(size_t)k * m + i] + c * f;
                        z[(size_t)k * m + i] = c * z[(size_t)k * m + i] - s * f;
                    }
                }
            }
            if (underflow && i >= l) continue;
            d[l] -= p;
            e[l] = g;
            e[mm] = 0.0;
        }
    }
    // sort ascending
    std::vector<int> idx(m);
    std::iota(idx.begin(), idx.end(), 0);
    std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return d[a] < d[b]; });
    std::vector<double> d2(m);
    for (int k = 0; k < m; ++k) d2[k] = d[idx[k]];
    d = d2;
    if (zrows > 0) {
        std::vector<double> Z2((size_t)zrows * m);
        for (int r = 0; r < zrows; ++r)
            for (int k = 0; k < m; ++k) Z2[(size_t)r * m + k] = (*Z)[(size_t)r * m + idx[k]];
        *Z = Z2;
    }
    return true;
}

// A: n x n symmetric (row-major), overwritten.  w ascending, V column j = eigenvector j (row-major n x n).
inline void jacobi_eig(std::vector<double> A, int n, std::vector<double>& w, std::vector<double>& V) {
    V.assign((size_t)n * n, 0.0);
    for (int i = 0; i < n; ++i) V[(size_t)i * n + i] = 1.0;
    for (int sweep = 0; sweep < 100; ++sweep) {
        double off = 0.0, tot = 0.0;
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < n; ++j) {
                const double a = A[(size_t)i * n + j];
                tot += a * a;
                if (i != j) off += a * a;
            }
        if (off <= 1e-32 * tot || off == 0.0) break;
        for (int p = 0; p < n; ++p)
            for (int q = p + 1; q < n; ++q) {
                const double apq = A[(size_t)p * n + q];
                if (apq == 0.0) continue;
                const double theta = (A[(size_t)q * n + q] - A[(size_t)p * n + p]) / (2.0 * apq);
                const double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
                const double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
                for (int k = 0; k < n; ++k) {
                    const double akp = A[(size_t)k * n + p], akq = A[(size_t)k * n + q];
                    A[(size_t)k * n + p] = c * akp - s * akq;
                    A[(size_t)k * n + q] = s * akp + c * akq;
                }
                for (int k = 0; k < n; ++k) {
                    const double apk = A[(size_t)p * n + k], aqk = A[(size_t)q * n + k];
                    A[(size_t)p * n + k] = c * apk - s * aqk;
                    A[(size_t)q * n + k] = s * apk + c * aqk;
                }
                for (int k = 0; k < n; ++k) {
                    const double vkp = V[(size_t)k * n + p], vkq = V[(size_t)k * n + q];
                    V[(size_t)k * n + p] = c * vkp - s * vkq;
                    V[(size_t)k * n + q] = s * vkp + c * vkq;
                }
            }
    }
    std::vector<int> idx(n);
    std::iota(idx.begin(), idx.end(), 0);
    std::stable_sort(idx.begin(), idx.end(),
                     [&](int a, int b) { return A[(size_t)a * n + a] < A[(size_t)b * n + b]; });
    w.resize(n);
    std::vector<double> V2((size_t)n * n);
    for (int k = 0; k < n; ++k) {
        w[k] = A[(size_t)idx[k] * n + idx[k]];
        for (int r = 0; r < n; ++r) V2[(size_t)r * n + k] = V[(size_t)r * n + idx[k]];
    }
    V = V2;
}

}  // namespace moed
