// Convergence test of the lowest Ritz pairs of a Lanczos tridiagonal T_m on the device, for one wave
// (the persistent eigensolver and the step-by-step engine): the host's test |beta_m z_{m,i}| <= tol
// |theta_i| for the lowest nev eigenvalues theta_i of T_m, z_{m,i} the bottom component of the
// normalised eigenvector.
//   theta_i   multisection on Sturm counts (the signs of the LDL^T pivots of T_m - x I), a group of
//             lanes per eigenvalue, from the Gershgorin interval to adjacent doubles
//   z_{m,i}   two steps of inverse iteration with the LDL^T factors of T_m - theta_i I (a pivot
//             below eps |T| is replaced by +-eps |T|)
#pragma once
#include <hip/hip_runtime.h>

namespace moed {

// number of eigenvalues of the tridiagonal (a[0..m), squared off-diagonals b2[0..m-1)) below x
__device__ __forceinline__ int sturm_count(const double* a, const double* b2, int m, double x, double tiny) {
    double q = a[0] - x;
    int c = q < 0.0;
    for (int j = 1; j < m; ++j) {
        if (q == 0.0) q = -tiny;
        q = (a[j] - x) - b2[j - 1] / q;
        c += q < 0.0;
    }
    return c;
}

// Called by all 64 lanes of one wave.  alpha[0..m), beta[0..m) in global memory (beta[m-1] is the
// residual norm of the step); ta, tb, tb2 hold m doubles each (LDS or global) for the staged T_m;
// ws holds nev * 3 * ldh doubles of inverse-iteration workspace (LDS where it fits).  Returns the
// same verdict in every lane: true when every one of the lowest min(nev, m) Ritz pairs passes, or
// lies above `above` with its whole residual interval at a residual below kAboveTol |theta_i| (a
// caller asking only which levels lie below `above` needs no more of it; the residual bound keeps
// an early, unresolved Ritz value from passing: a level that the Krylov space has not yet resolved
// still weighs about 1/sqrt(dim) in the Ritz vector, and its gap times that in the residual).
// Eigenvalues by multisection: the lanes form one group of P (a power of two) per wanted
// eigenvalue, and each round evaluates P Sturm counts inside the group's bracket and keeps the
// sub-interval that holds the eigenvalue, a factor P + 1 per round, down to adjacent doubles.
constexpr double kAboveTol = 1e-8;

__device__ bool ritz_converged(const double* alpha, const double* beta, int m, int nev, double tol, double* ta,
                               double* tb, double* tb2, double* ws, int ldh, double above = INFINITY) {
    const int lane = threadIdx.x & 63;
    const double eps = 0x1p-52;
    for (int j = lane; j < m; j += 64) {
        ta[j] = alpha[j];
        tb[j] = beta[j];
        tb2[j] = beta[j] * beta[j];
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    double lo = 1e300, hi = -1e300;  // Gershgorin interval
    for (int j = lane; j < m; j += 64) {
        const double r = (j > 0 ? fabs(tb[j - 1]) : 0.0) + (j < m - 1 ? fabs(tb[j]) : 0.0);
        lo = fmin(lo, ta[j] - r);
        hi = fmax(hi, ta[j] + r);
    }
    for (int off = 32; off > 0; off >>= 1) {
        lo = fmin(lo, __shfl_xor(lo, off, 64));
        hi = fmax(hi, __shfl_xor(hi, off, 64));
    }
    const double tiny = eps * fmax(fmax(fabs(lo), fabs(hi)), 1e-300);
    const int ne = min(nev, m);
    bool ok = true;
    for (int e0 = 0; e0 < ne; e0 += 64) {  // (more than 64 wanted eigenvalues: one per lane, in turns)
        const int nb = min(64, ne - e0);
        int P = 1;
        while (2 * P * nb <= 64) P *= 2;  // lanes per eigenvalue
        const int g = lane / P, jj = lane % P, ei = e0 + g;
        const bool act = g < nb;
        double l = lo - tiny, u = hi + tiny;
        // each round narrows the bracket by P + 1 (P = 1: bisection); 2^60 covers the Gershgorin
        // width down to adjacent doubles
        int rounds = 2;
        for (double f = 1.0; f < 0x1p60; f *= (double)(P + 1)) ++rounds;
        for (int it = 0; it < rounds; ++it) {
            const double x = l + (u - l) * ((double)(jj + 1) / (double)(P + 1));
            const bool right = act && x > l && x < u && sturm_count(ta, tb2, m, x, tiny) > ei;
            double nu = (act && right) ? x : u;
            double nl = (act && !right && x > l && x < u) ? x : l;
            for (int off = 1; off < P; off <<= 1) {  // within the group of P lanes
                nu = fmin(nu, __shfl_xor(nu, off, 64));
                nl = fmax(nl, __shfl_xor(nl, off, 64));
            }
            u = nu;
            l = nl;
        }
        if (act && jj == 0) {
            const double th = 0.5 * (l + u);
            // bottom component of the eigenvector: inverse iteration, two solves with the LDL^T
            // factors of T_m - theta I
            double* ll = ws + (size_t)g * 3 * ldh;
            double* dd = ll + ldh;
            double* r = dd + ldh;
            double dj = ta[0] - th;
            for (int j = 0; j < m; ++j) {
                if (fabs(dj) < tiny) dj = dj < 0.0 ? -tiny : tiny;
                dd[j] = dj;
                if (j < m - 1) {
                    ll[j] = tb[j] / dj;
                    dj = (ta[j + 1] - th) - ll[j] * tb[j];
                }
                r[j] = 1.0;
            }
            double zb = 0.0;
            for (int it = 0; it < 2; ++it) {
                double prev = r[0];
                for (int j = 1; j < m; ++j) {  // L y = r
                    prev = r[j] - ll[j - 1] * prev;
                    r[j] = prev;
                }
                for (int j = 0; j < m; ++j) r[j] = r[j] / dd[j];  // D z = y
                prev = r[m - 1];
                for (int j = m - 2; j >= 0; --j) {  // L^T x = z
                    prev = r[j] - ll[j] * prev;
                    r[j] = prev;
                }
                double n2 = 0.0;
                for (int j = 0; j < m; ++j) n2 = fma(r[j], r[j], n2);
                const double sc = 1.0 / sqrt(n2);
                for (int j = 0; j < m; ++j) r[j] *= sc;
                zb = r[m - 1];
            }
            const double res = fabs(beta[m - 1] * zb);
            ok = ok && (res <= tol * fmax(fabs(th), 1e-300) ||
                        (res <= kAboveTol * fmax(fabs(th), 1.0) && th - res > above));
        }
    }
    return __all(ok);
}

}  // namespace moed
