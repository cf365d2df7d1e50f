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

// Eigenvectors of the tridiagonal T (diagonal a[0..m), off-diagonal b[0..m-1)) for its lowest nvec
// eigenvalues w[0..nvec) (ascending, as tridiag_ql returns them), by inverse iteration: T - w_i I
// factored by Gaussian elimination with partial pivoting (Golub & Van Loan, section 4.3.3, band
// form: U gains a second superdiagonal where rows are exchanged; a pivot below eps |T| is replaced
// by +-eps |T|), three solves from a fixed start vector, each followed by modified Gram-Schmidt
// against the vectors already found for eigenvalues closer than 1e-3 |T| (the cluster tolerance of
// LAPACK's dstein); shifts closer than 10 eps |T| to the previous one are moved up by that much.
// V[i * m + j] = component j of vector i, unit norm, sign such that its largest component is
// positive.  O(nvec m) work against O(m^3) for accumulating every rotation of the QL iteration.
inline void tridiag_vectors(const double* a, const double* b, int m, const double* w, int nvec,
                            std::vector<double>& V) {
    V.assign((size_t)nvec * m, 0.0);
    double tn = 0.0;
    for (int j = 0; j < m; ++j)
        tn = std::max(tn, std::fabs(a[j]) + (j > 0 ? std::fabs(b[j - 1]) : 0.0) + (j < m - 1 ? std::fabs(b[j]) : 0.0));
    const double eps = 2.220446049250313e-16;
    const double tiny = std::max(eps * tn, 1e-300), ortol = 1e-3 * tn;
    std::vector<double> dl(m), dg(m), du(m), du2(m), x(m);
    std::vector<int> piv(m);
    double prev = 0.0;
    int c0 = 0;  // first vector of the current cluster
    for (int i = 0; i < nvec; ++i) {
        double th = w[i];
        if (i > 0) {
            if (th - w[i - 1] > ortol) c0 = i;
            if (th - prev < 10.0 * tiny) th = prev + 10.0 * tiny;
        }
        prev = th;
        for (int j = 0; j < m; ++j) {
            dg[j] = a[j] - th;
            if (j < m - 1) dl[j] = du[j] = b[j];
            du2[j] = 0.0;
        }
        for (int j = 0; j < m - 1; ++j) {
            if (std::fabs(dg[j]) >= std::fabs(dl[j])) {
                piv[j] = 0;
                const double f = dg[j] != 0.0 ? dl[j] / dg[j] : 0.0;
                dl[j] = f;
                dg[j + 1] -= f * du[j];
            } else {  // exchange rows j and j + 1
                piv[j] = 1;
                const double f = dg[j] / dl[j];
                dg[j] = dl[j];
                dl[j] = f;
                const double t = du[j];
                du[j] = dg[j + 1];
                dg[j + 1] = t - f * dg[j + 1];
                if (j < m - 2) {
                    du2[j] = du[j + 1];
                    du[j + 1] = -f * du[j + 1];
                }
            }
        }
        for (int j = 0; j < m; ++j)
            if (std::fabs(dg[j]) < tiny) dg[j] = dg[j] < 0.0 ? -tiny : tiny;
        for (int j = 0; j < m; ++j) x[j] = 1.0 + 0.25 * std::sin(0.7 * j + 0.3 * i);  // fixed start
        for (int it = 0; it < 3; ++it) {
            for (int j = 0; j < m - 1; ++j) {  // P, L
                if (piv[j] == 0) {
                    x[j + 1] -= dl[j] * x[j];
                } else {
                    const double t = x[j];
                    x[j] = x[j + 1];
                    x[j + 1] = t - dl[j] * x[j];
                }
            }
            for (int j = m - 1; j >= 0; --j) {  // U
                double r = x[j];
                if (j < m - 1) r -= du[j] * x[j + 1];
                if (j < m - 2) r -= du2[j] * x[j + 2];
                x[j] = r / dg[j];
            }
            for (int q = c0; q < i; ++q) {  // against the cluster's vectors found so far
                const double* v = V.data() + (size_t)q * m;
                double dt = 0.0;
                for (int j = 0; j < m; ++j) dt += v[j] * x[j];
                for (int j = 0; j < m; ++j) x[j] -= dt * v[j];
            }
            double mx = 0.0;
            for (int j = 0; j < m; ++j) mx = std::max(mx, std::fabs(x[j]));
            if (!(mx > 0.0) || !std::isfinite(mx)) {  // (lost in the cluster: restart from a unit vector)
                for (int j = 0; j < m; ++j) x[j] = j == (i + it) % m ? 1.0 : 0.0;
                continue;
            }
            for (int j = 0; j < m; ++j) x[j] /= mx;  // (keeps the next solve in range)
            double n2 = 0.0;
            for (int j = 0; j < m; ++j) n2 += x[j] * x[j];
            const double sc = 1.0 / std::sqrt(n2);
            for (int j = 0; j < m; ++j) x[j] *= sc;
        }
        int jm = 0;
        for (int j = 1; j < m; ++j)
            if (std::fabs(x[j]) > std::fabs(x[jm])) jm = j;
        const double sg = x[jm] < 0.0 ? -1.0 : 1.0;
        for (int j = 0; j < m; ++j) V[(size_t)i * m + j] = sg * x[j];
    }
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
