/* CPU oracle: plain-C restatement of the reference ED solver's hot path.
 *
 * TEST INFRASTRUCTURE ONLY -- the checker, never the product.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.
 *
 * Each function restates the reference routine cited above it, keeping the loop
 * order and the floating-point operation order, so that a strict-IEEE build of
 * the reference produces the same bits.  Build with -ffp-contract=off (no FMA).
 *
 * Single-precision literals of the reference are kept as (double)(float) values:
 *   ed_lanczos.F90:62 1.e-12, frprmn.f90:91 TOL=1.e-8, frprmn.f90:126 CGOLD/ZEPS,
 *   frprmn.f90:247 GOLD/TINY.
 * Fortran cmplx(x,y) without KIND returns a default-kind (REAL*4) complex, so the
 * reference evaluates G(i w) and Delta(i w) at single-precision-rounded E and w_n
 * (ed_green.F90:161,177,268; func.f90:25,80,96,110).  compat=1 reproduces that.
 * Complex arithmetic follows the reference build (flang): a/b by the textbook formula
 * ((ar*br + ai*bi)/(br^2+bi^2), (ai*br - ar*bi)/(br^2+bi^2)), a real operand promoted to
 * (x, 0), abs() = hypot.  The cx_* helpers below spell this out.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef struct {
    double re, im;
} cx;
static inline cx cx_mk(double re, double im) {
    cx z = {re, im};
    return z;
}
static inline cx cx_add(cx a, cx b) { return cx_mk(a.re + b.re, a.im + b.im); }
static inline cx cx_sub(cx a, cx b) { return cx_mk(a.re - b.re, a.im - b.im); }
static inline cx cx_subr(cx a, double r) { return cx_mk(a.re - r, a.im); }
static inline cx cx_addr(cx a, double r) { return cx_mk(a.re + r, a.im); }
static inline cx cx_div(cx a, cx b) {
    double den = b.re * b.re + b.im * b.im;
    return cx_mk((a.re * b.re + a.im * b.im) / den, (a.im * b.re - a.re * b.im) / den);
}
static inline cx cx_rdiv(double r, cx b) { return cx_div(cx_mk(r, 0.0), b); }
static inline cx cx_mul(cx a, cx b) { return cx_mk(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re); }
static inline cx cx_mulr(cx a, double r) { return cx_mul(a, cx_mk(r, 0.0)); }
static inline cx cx_conj(cx a) { return cx_mk(a.re, -a.im); }

/* ---------------------------------------------------------------- combinatorics */

/* ifact / iP / iCom, ed_utils.f90:28-69 (int32 arithmetic, wraps like the reference) */
static int32_t or_ifact(int32_t n) {
    uint32_t f = 1;
    for (int32_t i = 1; i <= n; i++) f *= (uint32_t)i;
    return (int32_t)f;
}
static int32_t or_iP(int32_t n, int32_t m) {
    uint32_t p = 1;
    for (int32_t i = n - m + 1; i <= n; i++) p *= (uint32_t)i;
    return (int32_t)p;
}
int32_t or_icom(int32_t n, int32_t m) {
    if (n == m || m == 0) return 1;
    if (m > n / 2) {
        int32_t k = n - m;
        return or_iP(n, k) / or_ifact(k);
    }
    return or_iP(n, m) / or_ifact(m);
}

/* ---------------------------------------------------------------- basis */

/* generate_basis, ed_basis.F90:35-107 (one rank: nloc = ntot, offsets = 0) */
int or_basis_make(int32_t nsite, int32_t ne_up, int32_t ne_dn, or_basis* b) {
    memset(b, 0, sizeof(*b));
    if (nsite < 1 || nsite > 15 || ne_up < 0 || ne_dn < 0 || ne_up > nsite || ne_dn > nsite) return -1;
    b->nsite = nsite;
    b->ne_up = ne_up;
    b->ne_dn = ne_dn;
    b->nup = or_icom(nsite, ne_up);
    b->ndn = or_icom(nsite, ne_dn);
    b->ntot = b->nup * b->ndn;
    b->up = (int32_t*)calloc((size_t)b->nup, sizeof(int32_t));
    b->dn = (int32_t*)calloc((size_t)b->ndn, sizeof(int32_t));
    b->idx_up = (int32_t*)calloc((size_t)1 << nsite, sizeof(int32_t));
    b->idx_dn = (int32_t*)calloc((size_t)1 << nsite, sizeof(int32_t));
    int nud[2] = {ne_up, ne_dn};
    for (int ispin = 0; ispin < 2; ispin++) {
        int32_t minrange = 0, maxrange = 0;
        for (int i = 1; i <= nud[ispin]; i++) {
            minrange += 1 << (i - 1);
            maxrange += 1 << (nsite - i);
        }
        int32_t counts = 0;
        for (int32_t i = minrange; i <= maxrange; i++) {
            int nbit = 0;
            for (int j = 0; j < nsite; j++)
                if ((i >> j) & 1) nbit++;
            if (nbit == nud[ispin]) {
                counts++;
                if (ispin == 0) {
                    b->up[counts - 1] = i;
                    b->idx_up[i] = counts;
                } else {
                    b->dn[counts - 1] = i;
                    b->idx_dn[i] = counts;
                }
            }
        }
    }
    return 0;
}

void or_basis_free(or_basis* b) {
    free(b->up);
    free(b->dn);
    free(b->idx_up);
    free(b->idx_dn);
    memset(b, 0, sizeof(*b));
}

/* ed_basis_get, ed_basis.F90:110-124 */
int64_t or_basis_get(const or_basis* b, int64_t idx) {
    int64_t iup = (idx - 1) % b->nup + 1;
    int64_t idown = (idx - 1) / b->nup + 1;
    return (int64_t)b->up[iup - 1] + ((int64_t)1 << b->nsite) * (int64_t)b->dn[idown - 1];
}

/* ed_basis_idx, ed_basis.F90:127-139 */
int64_t or_basis_idx(const or_basis* b, int64_t state) {
    int64_t p = (int64_t)1 << b->nsite;
    int64_t su = state % p, sd = state / p;
    return ((int64_t)b->idx_dn[sd] - 1) * b->nup + b->idx_up[su];
}

/* ---------------------------------------------------------------- operators */

/* get_bitidx, ed_operators.F90:63-66 */
static inline int bitidx_of(const or_basis* b, int isite, int ispin) { return (ispin - 1) * b->nsite + isite - 1; }

/* number_op, ed_operators.F90:273-290 */
static void or_number_op(const or_basis* b, int64_t in, int isite, int ispin, int64_t* out, int* sgn) {
    int bit = bitidx_of(b, isite, ispin);
    if ((in >> bit) & 1) {
        *sgn = 1;
        *out = in;
    } else {
        *sgn = 0;
        *out = 0;
    }
}

/* destruction_op, ed_operators.F90:292-319 */
void or_destruction_op(const or_basis* b, int64_t in, int isite, int ispin, int64_t* out, int* sgn) {
    int bit = bitidx_of(b, isite, ispin);
    if ((in >> bit) & 1) {
        int n = 0;
        for (int i = 0; i <= bit - 1; i++)
            if ((in >> i) & 1) n++;
        *out = in & ~((int64_t)1 << bit);
        *sgn = (n % 2 == 1) ? -1 : 1;
    } else {
        *out = 0;
        *sgn = 0;
    }
}

/* creation_op, ed_operators.F90:321-348 */
void or_creation_op(const or_basis* b, int64_t in, int isite, int ispin, int64_t* out, int* sgn) {
    int bit = bitidx_of(b, isite, ispin);
    if (!((in >> bit) & 1)) {
        int n = 0;
        for (int i = 0; i <= bit - 1; i++)
            if ((in >> i) & 1) n++;
        *out = in | ((int64_t)1 << bit);
        *sgn = (n % 2 == 1) ? -1 : 1;
    } else {
        *out = 0;
        *sgn = 0;
    }
}

static inline double vk_of(const or_model* m, int iorb, int ksite) {
    /* vk(iorb, ksite), ksite in Norb+1..Nsite, column-major (norb, nbath) */
    return m->vk[(iorb - 1) + m->norb * (ksite - m->norb - 1)];
}

/* onsite, ed_operators.F90:68-84 */
static void or_onsite(const or_model* m, const or_basis* b, int64_t in, int isite, int ispin, int64_t* out,
                      double* coeff) {
    int sgn;
    or_number_op(b, in, isite, ispin, out, &sgn);
    if (sgn == 0) {
        *coeff = 0.0;
        return;
    }
    *coeff = m->ek[isite - 1];
}

/* hybridization1, ed_operators.F90:86-107: c^+_{iorb} c_{bath} */
static void or_hybridization1(const or_model* m, const or_basis* b, int64_t in, int iorb, int ibath, int ispin,
                              int64_t* out, double* coeff) {
    int sgn;
    *coeff = 1.0;
    or_destruction_op(b, in, m->norb + ibath, ispin, out, &sgn);
    if (sgn == 0) {
        *coeff = 0.0;
        return;
    }
    *coeff = *coeff * sgn;
    or_creation_op(b, *out, iorb, ispin, out, &sgn);
    if (sgn == 0) {
        *coeff = 0.0;
        return;
    }
    *coeff = *coeff * sgn;
    *coeff = *coeff * vk_of(m, iorb, m->norb + ibath);
}

/* hybridization2, ed_operators.F90:109-130: c^+_{bath} c_{iorb} */
static void or_hybridization2(const or_model* m, const or_basis* b, int64_t in, int iorb, int ibath, int ispin,
                              int64_t* out, double* coeff) {
    int sgn;
    *coeff = 1.0;
    or_destruction_op(b, in, iorb, ispin, out, &sgn);
    if (sgn == 0) {
        *coeff = 0.0;
        return;
    }
    *coeff = *coeff * sgn;
    or_creation_op(b, *out, m->norb + ibath, ispin, out, &sgn);
    if (sgn == 0) {
        *coeff = 0.0;
        return;
    }
    *coeff = *coeff * sgn;
    *coeff = *coeff * vk_of(m, iorb, m->norb + ibath);
}

/* impurity hopping c^+_{iorb} c_{jorb} with t(iorb,jorb): extension, no reference counterpart;
 * written in the style of hybridization1 */
static void or_imp_hop(const or_model* m, const or_basis* b, int64_t in, int iorb, int jorb, int ispin,
                       int64_t* out, double* coeff) {
    int sgn;
    *coeff = 1.0;
    or_destruction_op(b, in, jorb, ispin, out, &sgn);
    if (sgn == 0) {
        *coeff = 0.0;
        return;
    }
    *coeff = *coeff * sgn;
    or_creation_op(b, *out, iorb, ispin, out, &sgn);
    if (sgn == 0) {
        *coeff = 0.0;
        return;
    }
    *coeff = *coeff * sgn;
    *coeff = *coeff * m->tij[(iorb - 1) + m->norb * (jorb - 1)];
}

/* dens_dens, ed_operators.F90:132-165 */
static void or_dens_dens(const or_model* m, const or_basis* b, int64_t in, int iorb, int jorb, int ispin,
                         int jspin, int64_t* out, double* coeff) {
    int n1, n2;
    or_number_op(b, in, iorb, ispin, out, &n1);
    or_number_op(b, in, jorb, jspin, out, &n2);
    if (n1 * n2 == 0) {
        *coeff = 0;
        return;
    }
    if (iorb == jorb) {
        if (ispin == jspin) {
            *coeff = 0.0;
            *out = 0;
            return;
        }
        *coeff = m->U;
    } else if (ispin == jspin) {
        *coeff = m->U - 3 * m->J;
    } else {
        *coeff = m->U - 2 * m->J;
    }
}

/* spin_flip, ed_operators.F90:167-218 */
static void or_spin_flip(const or_model* m, const or_basis* b, int64_t in, int iorb, int jorb, int64_t* out,
                         double* coeff) {
    int sgn_tot = 1, sgn;
    *coeff = 0.0;
    if (iorb == jorb) {
        *out = 0;
        return;
    }
    or_destruction_op(b, in, jorb, 1, out, &sgn);
    if (sgn == 0) return;
    sgn_tot *= sgn;
    or_creation_op(b, *out, jorb, 2, out, &sgn);
    if (sgn == 0) return;
    sgn_tot *= sgn;
    or_destruction_op(b, *out, iorb, 2, out, &sgn);
    if (sgn == 0) return;
    sgn_tot *= sgn;
    or_creation_op(b, *out, iorb, 1, out, &sgn);
    if (sgn == 0) return;
    sgn_tot *= sgn;
    *coeff = -m->J * sgn_tot;
}

/* pair_exchange, ed_operators.F90:220-271 */
static void or_pair_exchange(const or_model* m, const or_basis* b, int64_t in, int iorb, int jorb, int64_t* out,
                             double* coeff) {
    int sgn_tot = 1, sgn;
    *coeff = 0.0;
    if (iorb == jorb) {
        *out = 0;
        return;
    }
    or_destruction_op(b, in, jorb, 2, out, &sgn);
    if (sgn == 0) return;
    sgn_tot *= sgn;
    or_destruction_op(b, *out, jorb, 1, out, &sgn);
    if (sgn == 0) return;
    sgn_tot *= sgn;
    or_creation_op(b, *out, iorb, 2, out, &sgn);
    if (sgn == 0) return;
    sgn_tot *= sgn;
    or_creation_op(b, *out, iorb, 1, out, &sgn);
    if (sgn == 0) return;
    sgn_tot *= sgn;
    *coeff = -m->J * sgn_tot;
}

/* ---------------------------------------------------------------- H v */

/* diagonal part of multiply_H, ed_hamiltonian.F90:110-139 */
static double or_diag_of_state(const or_model* m, const or_basis* b, int64_t s) {
    int64_t sj;
    double coeff, coeff_sum = 0.0;
    for (int ispin = 1; ispin <= 2; ispin++) {
        for (int iorb = 1; iorb <= m->norb; iorb++) {
            or_onsite(m, b, s, iorb, ispin, &sj, &coeff);
            coeff_sum = coeff_sum + coeff;
        }
        for (int ibath = 1; ibath <= m->nbath; ibath++) {
            or_onsite(m, b, s, m->norb + ibath, ispin, &sj, &coeff);
            coeff_sum = coeff_sum + coeff;
        }
    }
    for (int iorb = 1; iorb <= m->norb; iorb++) {
        or_dens_dens(m, b, s, iorb, iorb, 1, 2, &sj, &coeff);
        coeff_sum = coeff_sum + coeff;
    }
    for (int iorb = 1; iorb <= m->norb; iorb++)
        for (int jorb = iorb + 1; jorb <= m->norb; jorb++)
            for (int ispin = 1; ispin <= 2; ispin++)
                for (int jspin = 1; jspin <= 2; jspin++) {
                    or_dens_dens(m, b, s, iorb, jorb, ispin, jspin, &sj, &coeff);
                    coeff_sum = coeff_sum + coeff;
                }
    return coeff_sum;
}

double or_diag_element(const or_model* m, const or_basis* b, int64_t i0) {
    return or_diag_of_state(m, b, or_basis_get(b, i0 + 1));
}

/* row i of H: the same terms as multiply_H below, collected instead of applied */
int or_h_row(const or_model* m, const or_basis* b, int64_t i0, int64_t* cols, double* vals, int cap) {
    int64_t s = or_basis_get(b, i0 + 1), sj;
    double coeff;
    int n = 0;
#define PUSH(col, v)                  \
    do {                              \
        if (n < cap) {                \
            cols[n] = (col);          \
            vals[n] = (v);            \
        }                             \
        n++;                          \
    } while (0)
    PUSH(i0, or_diag_of_state(m, b, s));
    for (int ispin = 1; ispin <= 2; ispin++)
        for (int iorb = 1; iorb <= m->norb; iorb++)
            for (int ibath = 1; ibath <= m->nbath; ibath++) {
                or_hybridization1(m, b, s, iorb, ibath, ispin, &sj, &coeff);
                if (sj != 0 && coeff != 0.0) PUSH(or_basis_idx(b, sj) - 1, coeff);
                or_hybridization2(m, b, s, iorb, ibath, ispin, &sj, &coeff);
                if (sj != 0 && coeff != 0.0) PUSH(or_basis_idx(b, sj) - 1, coeff);
            }
    if (m->tij)
        for (int ispin = 1; ispin <= 2; ispin++)
            for (int iorb = 1; iorb <= m->norb; iorb++)
                for (int jorb = 1; jorb <= m->norb; jorb++) {
                    if (iorb == jorb) continue;
                    or_imp_hop(m, b, s, iorb, jorb, ispin, &sj, &coeff);
                    if (sj != 0 && coeff != 0.0) PUSH(or_basis_idx(b, sj) - 1, coeff);
                }
    for (int iorb = 1; iorb <= m->norb; iorb++)
        for (int jorb = 1; jorb <= m->norb; jorb++) {
            if (iorb == jorb) continue;
            or_spin_flip(m, b, s, iorb, jorb, &sj, &coeff);
            if (sj != 0 && coeff != 0.0) PUSH(or_basis_idx(b, sj) - 1, coeff);
            or_pair_exchange(m, b, s, iorb, jorb, &sj, &coeff);
            if (sj != 0 && coeff != 0.0) PUSH(or_basis_idx(b, sj) - 1, coeff);
        }
#undef PUSH
    return n;
}

/* multiply_H, ed_hamiltonian.F90:88-191 (rows split over threads instead of MPI ranks) */
void or_multiply_H(const or_model* m, const or_basis* b, const double* A, double* B, int nthreads) {
    int64_t n = b->ntot;
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for schedule(static) num_threads(nthreads)
#endif
    for (int64_t i = 1; i <= n; i++) {
        int64_t s = or_basis_get(b, i), sj;
        double coeff;
        double Bi = or_diag_of_state(m, b, s) * A[i - 1];
        for (int ispin = 1; ispin <= 2; ispin++)
            for (int iorb = 1; iorb <= m->norb; iorb++)
                for (int ibath = 1; ibath <= m->nbath; ibath++) {
                    or_hybridization1(m, b, s, iorb, ibath, ispin, &sj, &coeff);
                    if (sj != 0) Bi = Bi + coeff * A[or_basis_idx(b, sj) - 1];
                    or_hybridization2(m, b, s, iorb, ibath, ispin, &sj, &coeff);
                    if (sj != 0) Bi = Bi + coeff * A[or_basis_idx(b, sj) - 1];
                }
        if (m->tij) /* extension: impurity-impurity hopping */
            for (int ispin = 1; ispin <= 2; ispin++)
                for (int iorb = 1; iorb <= m->norb; iorb++)
                    for (int jorb = 1; jorb <= m->norb; jorb++) {
                        if (iorb == jorb) continue;
                        or_imp_hop(m, b, s, iorb, jorb, ispin, &sj, &coeff);
                        if (sj != 0) Bi = Bi + coeff * A[or_basis_idx(b, sj) - 1];
                    }
        for (int iorb = 1; iorb <= m->norb; iorb++)
            for (int jorb = 1; jorb <= m->norb; jorb++) {
                if (iorb == jorb) continue;
                or_spin_flip(m, b, s, iorb, jorb, &sj, &coeff);
                if (sj != 0) Bi = Bi + coeff * A[or_basis_idx(b, sj) - 1];
            }
        for (int iorb = 1; iorb <= m->norb; iorb++)
            for (int jorb = 1; jorb <= m->norb; jorb++) {
                if (iorb == jorb) continue;
                or_pair_exchange(m, b, s, iorb, jorb, &sj, &coeff);
                if (sj != 0) Bi = Bi + coeff * A[or_basis_idx(b, sj) - 1];
            }
        B[i - 1] = Bi;
    }
}

/* apply_c, ed_operators.F90:10-61 (target sector basis passed in) */
void or_apply_c(const or_basis* b, const double* vec, int pm, int iorb, int ispin, const or_basis* bout,
                double* vec_out) {
    for (int64_t i = 1; i <= bout->ntot; i++) {
        int64_t si = or_basis_get(bout, i), sj;
        int sgn;
        vec_out[i - 1] = 0.0;
        if (pm == 1)
            or_destruction_op(b, si, iorb, ispin, &sj, &sgn);
        else
            or_creation_op(b, si, iorb, ispin, &sj, &sgn);
        if (sgn == 0) continue;
        int64_t j = or_basis_idx(b, sj);
        vec_out[i - 1] = vec_out[i - 1] + vec[j - 1] * sgn;
    }
}

/* ---------------------------------------------------------------- Lanczos */

/* mpi_dot_product at one rank, ed_utils.f90:7-18: sum(A*B) in index order */
static double or_dot(const double* a, const double* b, int64_t n) {
    double s = 0.0;
    for (int64_t i = 0; i < n; i++) s = s + a[i] * b[i];
    return s;
}

/* lanczos_iteration, ed_lanczos.F90:14-68 */
int or_lanczos_iteration(const or_model* m, const or_basis* b, const double* vec, int nstep, int* nstep_calc,
                         double* a, double* bb, int nthreads) {
    int64_t n = b->ntot;
    if (nstep < 1) return -1;
    double* v1 = (double*)malloc(sizeof(double) * n);
    double* v2 = (double*)calloc(n, sizeof(double));
    double* v3 = (double*)malloc(sizeof(double) * n);
    double* w = (double*)calloc(n, sizeof(double));
    const double brk = (double)1.e-12f; /* ed_lanczos.F90:62, REAL*4 literal */
    for (int j = 0; j <= nstep; j++) a[j] = bb[j] = 0.0;
    memcpy(v1, vec, sizeof(double) * n);
    double norm_v = or_dot(v1, v1, n);
    for (int64_t i = 0; i < n; i++) v1[i] = v1[i] / sqrt(norm_v);
    or_multiply_H(m, b, v1, w, nthreads);
    a[0] = or_dot(v1, w, n);
    bb[0] = 0.0;
    for (int64_t i = 0; i < n; i++) v2[i] = w[i] - a[0] * v1[i];
    bb[1] = or_dot(v2, v2, n);
    or_multiply_H(m, b, v2, w, nthreads);
    a[1] = or_dot(w, v2, n);
    a[1] = a[1] / bb[1];
    for (int64_t i = 0; i < n; i++) v2[i] = v2[i] / sqrt(bb[1]);
    *nstep_calc = nstep;
    for (int j = 2; j <= nstep; j++) {
        for (int64_t i = 0; i < n; i++) v3[i] = w[i] / sqrt(bb[j - 1]) - a[j - 1] * v2[i] - sqrt(bb[j - 1]) * v1[i];
        bb[j] = or_dot(v3, v3, n);
        or_multiply_H(m, b, v3, w, nthreads);
        a[j] = or_dot(v3, w, n);
        a[j] = a[j] / bb[j];
        for (int64_t i = 0; i < n; i++) v3[i] = v3[i] / sqrt(bb[j]);
        if (bb[j] < brk) {
            *nstep_calc = j;
            break;
        }
        double* t = v1;
        v1 = v2;
        v2 = v3;
        v3 = t;
    }
    free(v1);
    free(v2);
    free(v3);
    free(w);
    return 0;
}

/* ---------------------------------------------------------------- ed_solve helpers */

/* boltzmann_factor, ed_solver.F90:193-219 */
void or_boltzmann_factor(const double* elv, int n, double beta, double* p) {
    for (int i = 0; i < n; i++) {
        double Z = 0.0, en = elv[i];
        int cut = 0;
        for (int k = 0; k < n; k++) {
            if (-beta * (elv[k] - en) > 10.0) {
                p[i] = 0.0;
                cut = 1;
                break;
            }
            Z = Z + exp(-beta * (elv[k] - en));
        }
        if (!cut) p[i] = 1.0 / Z;
    }
}

/* SORT (Numerical Recipes heapsort index), func.f90:301-353; indx is 1-based */
void or_sort(int n, const double* arrin, int* indx) {
    const double* A = arrin - 1; /* 1-based views */
    int* X = indx - 1;
    if (n == 1) {
        X[1] = 1;
        return;
    }
    for (int j = 1; j <= n; j++) X[j] = j;
    int l = n / 2 + 1, ir = n, indxt, i, j;
    double q;
    for (;;) {
        if (l > 1) {
            l = l - 1;
            indxt = X[l];
            q = A[indxt];
        } else {
            indxt = X[ir];
            q = A[indxt];
            X[ir] = X[1];
            ir = ir - 1;
            if (ir == 1) {
                X[1] = indxt;
                return;
            }
        }
        i = l;
        j = l + l;
        while (j <= ir) {
            if (j < ir && A[X[j]] < A[X[j + 1]]) j = j + 1;
            if (q < A[X[j]]) {
                X[i] = X[j];
                i = j;
                j = j + j;
            } else {
                j = ir + 1;
            }
        }
        X[i] = indxt;
    }
}

/* ---------------------------------------------------------------- Green's function */

static const double OR_PI = 3.141592653589793115997963468544185161590576171875; /* 4*atan(1d0) */

/* ed_green_init frequency grid, ed_green.F90:45-51 at one rank (ishift = 0) */
void or_omega_grid(int nw, double beta, double* omega) {
    for (int i = 1; i <= nw; i++) omega[i - 1] = (2.0 * (double)(float)(i - 1) + 1) * OR_PI / beta;
}

static inline double cround(double x, int compat) { return compat ? (double)(float)x : x; }

/* continued fractions, ed_green.F90:160-168 (sign=+1) and :176-184 (sign=-1) */
void or_green_cf(int sign, double E, int n, const double* a, const double* b, int nw, const double* omega,
                 int compat, double* g) {
    for (int i = 0; i < nw; i++) {
        cx z, grx;
        if (sign > 0) {
            z = cx_mk(cround(E, compat), cround(omega[i], compat));
            grx = cx_rdiv(b[n], cx_subr(z, a[n]));
            for (int j = n - 1; j >= 0; j--) grx = cx_rdiv(b[j], cx_sub(cx_subr(z, a[j]), grx));
        } else {
            z = cx_mk(cround(-E, compat), cround(omega[i], compat));
            grx = cx_rdiv(b[n], cx_addr(z, a[n]));
            for (int j = n - 1; j >= 0; j--) grx = cx_rdiv(b[j], cx_sub(cx_addr(z, a[j]), grx));
        }
        g[2 * i] = grx.re;
        g[2 * i + 1] = grx.im;
    }
}

/* dos_diag, dos.f90:65-106: A(w) = -Im(G+(w) + G-(w)) / pi on the real axis for one record, with
 * z+ = cmplx(ev + w, small) and z- = cmplx(w - ev, small) (default-kind cmplx: REAL*4 parts when
 * compat); the recurrences are green_diag's.  pi = 4 atan(1) (ed_config.f90:7). */
void or_dos_diag(int nw, const double* omega_real, double small, double ev, int nsp, int nsn, const double* ap,
                 const double* bp, const double* an, const double* bn, int compat, double* aff) {
    const double pi = 4.0 * atan(1.0);
    for (int i = 0; i < nw; i++) {
        cx z = cx_mk(cround(ev + omega_real[i], compat), cround(small, compat));
        cx gp = cx_rdiv(bp[nsp], cx_subr(z, ap[nsp]));
        for (int j = nsp - 1; j >= 0; j--) gp = cx_rdiv(bp[j], cx_sub(cx_subr(z, ap[j]), gp));
        z = cx_mk(cround(omega_real[i] - ev, compat), cround(small, compat));
        cx gn = cx_rdiv(bn[nsn], cx_addr(z, an[nsn]));
        for (int j = nsn - 1; j >= 0; j--) gn = cx_rdiv(bn[j], cx_sub(cx_addr(z, an[j]), gn));
        const cx grx = cx_add(gp, gn);
        aff[i] = -grx.im / pi;
    }
}

/* green_diag, ed_green.F90:136-185 */
int or_green_diag(const or_model* m, const or_basis* b, const double* eigvec, double E, double pev, double factor,
                  int iorb, int ispin, int nstep, int nw, const double* omega, int compat, double* Gr, int* nsp,
                  int* nsn, double* ap, double* bp, double* an, double* bn, int nthreads) {
    or_basis bo;
    int dup = (ispin == 1), ddn = (ispin == 2);
    double* g = (double*)malloc(sizeof(double) * 2 * nw);
    /* one more particle */
    if (or_basis_make(b->nsite, b->ne_up + dup, b->ne_dn + ddn, &bo)) return -1;
    double* v = (double*)malloc(sizeof(double) * bo.ntot);
    or_apply_c(b, eigvec, 1, iorb, ispin, &bo, v);
    or_lanczos_iteration(m, &bo, v, nstep, nsp, ap, bp, nthreads);
    double normsq = or_dot(v, v, bo.ntot);
    bp[0] = normsq;
    or_green_cf(+1, E, *nsp, ap, bp, nw, omega, compat, g);
    for (int i = 0; i < nw; i++) {
        Gr[2 * i] = Gr[2 * i] + pev * factor * g[2 * i];
        Gr[2 * i + 1] = Gr[2 * i + 1] + pev * factor * g[2 * i + 1];
    }
    free(v);
    or_basis_free(&bo);
    /* one less particle */
    if (or_basis_make(b->nsite, b->ne_up - dup, b->ne_dn - ddn, &bo)) return -1;
    v = (double*)malloc(sizeof(double) * bo.ntot);
    or_apply_c(b, eigvec, 2, iorb, ispin, &bo, v);
    or_lanczos_iteration(m, &bo, v, nstep, nsn, an, bn, nthreads);
    bn[0] = 1.0 - normsq;
    or_green_cf(-1, E, *nsn, an, bn, nw, omega, compat, g);
    for (int i = 0; i < nw; i++) {
        Gr[2 * i] = Gr[2 * i] + pev * factor * g[2 * i];
        Gr[2 * i + 1] = Gr[2 * i + 1] + pev * factor * g[2 * i + 1];
    }
    free(v);
    or_basis_free(&bo);
    free(g);
    return 0;
}

/* calc_dev, ed_green.F90:260-273; ef = impurity levels */
void or_calc_dev(const or_model* m, const double* ef, int nw, const double* omega, int compat, double* Dev) {
    for (int i = 0; i < nw; i++)
        for (int korb = 1; korb <= m->norb; korb++) {
            cx d = cx_mk(0.0, 0.0);
            cx z = cx_mk(0.0, cround(omega[i], compat));
            for (int k = m->norb + 1; k <= m->nsite; k++) {
                double v = vk_of(m, korb, k);
                d = cx_add(d, cx_rdiv(v * v, cx_subr(z, m->ek[k - 1])));
            }
            d = cx_addr(d, ef[korb - 1]);
            size_t o = 2 * ((size_t)(korb - 1) + (size_t)m->norb * i);
            Dev[o] = d.re;
            Dev[o + 1] = d.im;
        }
}

/* ---------------------------------------------------------------- bath fit */

/* x_to_ev, func.f90:193-216 */
void or_x_to_ev(const double* x, int nxsize, int ns, int nbath, int norb, double* ef, double* ek, double* vk) {
    for (int i = 1; i <= norb; i++) {
        ek[i - 1] = x[nxsize - norb + i - 1];
        ef[i - 1] = ek[i - 1];
    }
    for (int i = norb + 1; i <= norb + nbath; i++) ek[i - 1] = x[i - norb - 1];
    for (int iorb = 1; iorb <= norb; iorb++) {
        int ia = (iorb - 1) * nbath;
        for (int i = norb + 1; i <= ns; i++) vk[(iorb - 1) + norb * (i - norb - 1)] = x[i + nbath + ia - norb - 1];
    }
}

/* ev_to_x, func.f90:165-191 */
void or_ev_to_x(const double* ek, const double* vk, const double* ef, int ns, int nbath, int nxsize, double* x) {
    int norb = ns - nbath;
    for (int i = 1; i <= nbath; i++) x[i - 1] = ek[norb + i - 1];
    for (int iorb = 1; iorb <= norb; iorb++) {
        int iinit = nbath + 1 + (iorb - 1) * nbath, ifina = nbath + iorb * nbath;
        for (int k = iinit; k <= ifina; k++) x[k - 1] = vk[(iorb - 1) + norb * (k - iinit)];
    }
    for (int i = 1; i <= norb; i++) x[nxsize - norb + i - 1] = ef[i - 1];
}

/* Delta(iw) of FUNC/dfunc, func.f90:19-29 and :74-84 */
static void or_delta(const double* x, int ns, int nw, int nxsize, const double* omega, int nbath, int norb,
                     int compat, cx* delta, double* ek, double* vk, double* ef) {
    or_x_to_ev(x, nxsize, ns, nbath, norb, ef, ek, vk);
    for (int iw = 0; iw < nw; iw++)
        for (int korb = 1; korb <= norb; korb++) {
            cx d = cx_mk(0.0, 0.0);
            cx z = cx_mk(0.0, cround(omega[iw], compat));
            for (int ks = norb + 1; ks <= ns; ks++) {
                double v = vk[(korb - 1) + norb * (ks - norb - 1)];
                d = cx_add(d, cx_rdiv(v * v, cx_subr(z, ek[ks - 1])));
            }
            delta[(korb - 1) + norb * iw] = cx_addr(d, ef[korb - 1]);
        }
}

/* FUNC, func.f90:1-45 (one rank) */
double or_func(const double* x, int ns, int nw, int nxsize, const double* omega, const double* Dev, int nbath,
               int norb, int compat) {
    cx* delta = (cx*)malloc(sizeof(cx) * norb * nw);
    double* ek = (double*)malloc(sizeof(double) * ns);
    double* vk = (double*)malloc(sizeof(double) * norb * (nbath > 0 ? nbath : 1));
    double* ef = (double*)malloc(sizeof(double) * norb);
    or_delta(x, ns, nw, nxsize, omega, nbath, norb, compat, delta, ek, vk, ef);
    double f = 0.0;
    for (int iw = 0; iw < nw; iw++)
        for (int korb = 0; korb < norb; korb++) {
            size_t o = (size_t)korb + (size_t)norb * iw;
            cx c = cx_sub(delta[o], cx_mk(Dev[2 * o], Dev[2 * o + 1]));
            f = f + hypot(c.re, c.im) * hypot(c.re, c.im) / omega[iw];
        }
    f = f / (double)(float)(norb * norb * nw);
    free(delta);
    free(ek);
    free(vk);
    free(ef);
    return f;
}

/* dfunc, func.f90:47-142 (one rank) */
void or_dfunc(const double* x, double* p, int nw, int ns, int nxsize, const double* omega, const double* Dev,
              int nbath, int norb, int compat) {
    cx* delta = (cx*)malloc(sizeof(cx) * norb * nw);
    double* ek = (double*)malloc(sizeof(double) * ns);
    double* vk = (double*)malloc(sizeof(double) * norb * (nbath > 0 ? nbath : 1));
    double* ef = (double*)malloc(sizeof(double) * norb);
    or_delta(x, ns, nw, nxsize, omega, nbath, norb, compat, delta, ek, vk, ef);
#define DEV(o) cx_mk(Dev[2 * (o)], Dev[2 * (o) + 1])
    int nek = nbath;
    for (int ibath = 1; ibath <= nek; ibath++) {
        double s = 0.0;
        for (int korb = 1; korb <= norb; korb++)
            for (int iw = 0; iw < nw; iw++) {
                size_t o = (size_t)(korb - 1) + (size_t)norb * iw;
                cx z = cx_mk(0.0, cround(omega[iw], compat));
                double v = vk[(korb - 1) + norb * (ibath - 1)];
                cx den = cx_subr(z, ek[norb + ibath - 1]);
                cx t = cx_div(cx_div(cx_mulr(cx_mulr(cx_conj(cx_sub(delta[o], DEV(o))), v), v), den), den);
                s = s + 2.0 * t.re / omega[iw];
            }
        p[ibath - 1] = s;
    }
    for (int iorb = 1; iorb <= norb; iorb++)
        for (int ks = norb + 1; ks <= ns; ks++) {
            int i = nek + ks - norb + (iorb - 1) * nbath;
            double s = 0.0;
            for (int iw = 0; iw < nw; iw++) {
                size_t o = (size_t)(iorb - 1) + (size_t)norb * iw;
                cx z = cx_mk(0.0, cround(omega[iw], compat));
                double v = vk[(iorb - 1) + norb * (ks - norb - 1)];
                cx den = cx_subr(z, ek[ks - 1]);
                cx cd = cx_conj(cx_sub(delta[o], DEV(o)));
                cx t = cx_add(cx_div(cx_mulr(cd, v), den), cx_div(cx_mulr(cd, v), den));
                s = s + 2.0 * t.re / omega[iw];
            }
            p[i - 1] = s;
        }
    for (int iorb = 1; iorb <= norb; iorb++) {
        int i = nxsize - norb + iorb;
        double s = 0.0;
        for (int iw = 0; iw < nw; iw++) {
            size_t o = (size_t)(iorb - 1) + (size_t)norb * iw;
            s = s + 2.0 * cx_sub(delta[o], DEV(o)).re / omega[iw];
        }
        p[i - 1] = s;
    }
#undef DEV
    for (int i = 0; i < nxsize; i++) p[i] = p[i] / (double)(float)(norb * norb) / (double)(float)nw;
    free(delta);
    free(ek);
    free(vk);
    free(ef);
}

/* ---------------------------------------------------------------- CG minimiser (frprmn.f90) */

typedef struct {
    int n, nw, ns, nbath, norb, compat;
    const double *omega, *Dev;
    const double *pcom, *xicom;
    double* xt;
    int* nfev;
} or_f1ctx;

/* f1dim, frprmn.f90:212-234 */
static double or_f1dim(or_f1ctx* c, double x) {
    for (int j = 0; j < c->n; j++) c->xt[j] = c->pcom[j] + x * c->xicom[j];
    (*c->nfev)++;
    return or_func(c->xt, c->ns, c->nw, c->n, c->omega, c->Dev, c->nbath, c->norb, c->compat);
}

static inline double f_sign(double a, double b) { return b >= 0.0 ? fabs(a) : -fabs(a); }

/* mnbrak, frprmn.f90:236-310 */
static void or_mnbrak(or_f1ctx* c, double* ax, double* bx, double* cx, double* fa, double* fb, double* fc) {
    const double GOLD = (double)1.618034f, GLIMIT = 100.0, TINY = (double)1.e-20f;
    double dum, fu, q, r, u, ulim;
    *fa = or_f1dim(c, *ax);
    *fb = or_f1dim(c, *bx);
    if (*fb > *fa) {
        dum = *ax;
        *ax = *bx;
        *bx = dum;
        dum = *fb;
        *fb = *fa;
        *fa = dum;
    }
    *cx = *bx + GOLD * (*bx - *ax);
    *fc = or_f1dim(c, *cx);
    while (*fb >= *fc) {
        r = (*bx - *ax) * (*fb - *fc);
        q = (*bx - *cx) * (*fb - *fa);
        u = *bx - ((*bx - *cx) * q - (*bx - *ax) * r) / (2. * f_sign(fmax(fabs(q - r), TINY), q - r));
        ulim = *bx + GLIMIT * (*cx - *bx);
        if ((*bx - u) * (u - *cx) > 0.) {
            fu = or_f1dim(c, u);
            if (fu < *fc) {
                *ax = *bx;
                *fa = *fb;
                *bx = u;
                *fb = fu;
                return;
            } else if (fu > *fb) {
                *cx = u;
                *fc = fu;
                return;
            }
            u = *cx + GOLD * (*cx - *bx);
            fu = or_f1dim(c, u);
        } else if ((*cx - u) * (u - ulim) > 0.) {
            fu = or_f1dim(c, u);
            if (fu < *fc) {
                *bx = *cx;
                *cx = u;
                u = *cx + GOLD * (*cx - *bx);
                *fb = *fc;
                *fc = fu;
                fu = or_f1dim(c, u);
            }
        } else if ((u - ulim) * (ulim - *cx) >= 0.) {
            u = ulim;
            fu = or_f1dim(c, u);
        } else {
            u = *cx + GOLD * (*cx - *bx);
            fu = or_f1dim(c, u);
        }
        *ax = *bx;
        *bx = *cx;
        *cx = u;
        *fa = *fb;
        *fb = *fc;
        *fc = fu;
    }
}

/* brent, frprmn.f90:119-210; returns 0 or -2 on ITMAX */
static int or_brent(or_f1ctx* c, double ax, double bx, double cx, double tol, double* xmin, double* fret) {
    const int ITMAX = 100;
    const double CGOLD = (double).3819660f, ZEPS = (double)1.0e-10f;
    double a, b, d = 0.0, e, etemp, fu, fv, fw, fx, p, q, r, tol1, tol2, u, v, w, x, xm;
    a = fmin(ax, cx);
    b = fmax(ax, cx);
    v = bx;
    w = v;
    x = v;
    e = 0.;
    fx = or_f1dim(c, x);
    fv = fx;
    fw = fx;
    for (int iter = 1; iter <= ITMAX; iter++) {
        xm = 0.5 * (a + b);
        tol1 = tol * fabs(x) + ZEPS;
        tol2 = 2. * tol1;
        if (fabs(x - xm) <= (tol2 - .5 * (b - a))) {
            *xmin = x;
            *fret = fx;
            return 0;
        }
        int golden = 1;
        if (fabs(e) > tol1) {
            r = (x - w) * (fx - fv);
            q = (x - v) * (fx - fw);
            p = (x - v) * q - (x - w) * r;
            q = 2. * (q - r);
            if (q > 0.) p = -p;
            q = fabs(q);
            etemp = e;
            e = d;
            if (!(fabs(p) >= fabs(.5 * q * etemp) || p <= q * (a - x) || p >= q * (b - x))) {
                d = p / q;
                u = x + d;
                if (u - a < tol2 || b - u < tol2) d = f_sign(tol1, xm - x);
                golden = 0;
            }
        }
        if (golden) {
            if (x >= xm)
                e = a - x;
            else
                e = b - x;
            d = CGOLD * e;
        }
        if (fabs(d) >= tol1)
            u = x + d;
        else
            u = x + f_sign(tol1, d);
        fu = or_f1dim(c, u);
        if (fu <= fx) {
            if (u >= x)
                a = x;
            else
                b = x;
            v = w;
            fv = fw;
            w = x;
            fw = fx;
            x = u;
            fx = fu;
        } else {
            if (u < x)
                a = u;
            else
                b = u;
            if (fu <= fw || w == x) {
                v = w;
                fv = fw;
                w = u;
                fw = fu;
            } else if (fu <= fv || v == x || v == w) {
                v = u;
                fv = fu;
            }
        }
    }
    return -2;
}

/* frprmn + linmin, frprmn.f90:1-117 */
int or_frprmn(double* p, int n, double ftol, int* iter, double* fret, int nw, int ns, const double* omega,
              const double* Dev, int nbath, int norb, int compat, int* nfev) {
    const int ITMAX = 20000;
    const double EPS = 1.e-10;
    const double LTOL = (double)1.e-8f; /* linmin TOL, frprmn.f90:91 */
    double *g = (double*)malloc(sizeof(double) * n), *h = (double*)malloc(sizeof(double) * n);
    double *xi = (double*)malloc(sizeof(double) * n), *pcom = (double*)malloc(sizeof(double) * n);
    double *xicom = (double*)malloc(sizeof(double) * n), *xt = (double*)malloc(sizeof(double) * n);
    int dummy = 0, rc = -1;
    if (!nfev) nfev = &dummy;
    or_f1ctx c = {n, nw, ns, nbath, norb, compat, omega, Dev, pcom, xicom, xt, nfev};
    double fp = or_func(p, ns, nw, n, omega, Dev, nbath, norb, compat);
    (*nfev)++;
    or_dfunc(p, xi, nw, ns, n, omega, Dev, nbath, norb, compat);
    for (int j = 0; j < n; j++) {
        g[j] = -xi[j];
        h[j] = g[j];
        xi[j] = h[j];
    }
    for (int its = 1; its <= ITMAX; its++) {
        *iter = its;
        /* linmin */
        for (int j = 0; j < n; j++) {
            pcom[j] = p[j];
            xicom[j] = xi[j];
        }
        double ax = 0., xx = 1., bx, fa, fx, fb, xmin;
        or_mnbrak(&c, &ax, &xx, &bx, &fa, &fx, &fb);
        if (or_brent(&c, ax, xx, bx, LTOL, &xmin, fret)) {
            rc = -2;
            goto done;
        }
        for (int j = 0; j < n; j++) {
            xi[j] = xmin * xi[j];
            p[j] = p[j] + xi[j];
        }
        /* end linmin */
        if (2.0 * fabs(*fret - fp) <= ftol * (fabs(*fret) + fabs(fp) + EPS)) {
            rc = 0;
            goto done;
        }
        fp = *fret;
        or_dfunc(p, xi, nw, ns, n, omega, Dev, nbath, norb, compat);
        double gg = 0.0, dgg = 0.0;
        for (int j = 0; j < n; j++) {
            gg = gg + g[j] * g[j];
            dgg = dgg + (xi[j] + g[j]) * xi[j];
        }
        if (gg == 0.0) {
            rc = 0;
            goto done;
        }
        double gam = dgg / gg;
        for (int j = 0; j < n; j++) {
            g[j] = -xi[j];
            h[j] = g[j] + gam * h[j];
            xi[j] = h[j];
        }
    }
done:
    free(g);
    free(h);
    free(xi);
    free(pcom);
    free(xicom);
    free(xt);
    return rc;
}

/* ---------------------------------------------------------------- lattice step */

/* ed_set_band_structure + find_hk (ed_hamiltonian.F90:29-85) */
void or_band_tb(int norb, int nq, double* hk) {
    const double pi = 4.0 * atan(1.0);                  /* ed_config.f90:7 */
    const int nk = (int)sqrt((double)nq);               /* int(dsqrt(dfloat(Nq))) */
    const double dq = 2.0 * pi / (double)(float)nk;     /* float(nk): REAL*4, exact */
    int ik = 0;
    for (int i = 1; i <= nk; i++)
        for (int j = 1; j <= nk; j++) {
            const double kx = dq * (double)(float)(i - 1);
            const double ky = dq * (double)(float)(j - 1);
            const double h = -0.5 * (cos(kx) + cos(ky));
            for (int o = 0; o < norb; o++) hk[(size_t)ik * norb + o] = h;
            ik++;
        }
}

/* ed_delta_new (ed_green.F90:187-215), node-local frequency slice = all nw */
void or_delta_new(int norb, int nw, int nq, double rmu, const double* hk, const double* gr, double* dev,
                  double* gksum) {
    for (int iw = 0; iw < nw; iw++)
        for (int o = 0; o < norb; o++) {
            const size_t e = (size_t)iw * norb + o;
            const cx G = cx_mk(gr[2 * e], gr[2 * e + 1]);
            const cx D = cx_mk(dev[2 * e], dev[2 * e + 1]);
            cx s = cx_mk(0.0, 0.0);
            for (int q = 0; q < nq; q++) {
                const cx tmp = cx_sub(cx_add(cx_add(cx_rdiv(1.0, G), D), cx_mk(rmu, 0.0)),
                                      cx_mk(hk[(size_t)q * norb + o], 0.0));
                s = cx_add(s, cx_rdiv(1.0, tmp));
            }
            const double fq = (double)(float)nq;
            s = cx_div(s, cx_mk(fq, 0.0));
            gksum[2 * e] = s.re;
            gksum[2 * e + 1] = s.im;
            const cx dn = cx_add(cx_sub(cx_rdiv(1.0, G), cx_rdiv(1.0, s)), D);
            dev[2 * e] = dn.re;
            dev[2 * e + 1] = dn.im;
        }
}

/* n_from_gksum (ed_green.F90:222-253) */
void or_n_from_gksum(int norb, int nw, double beta, const double* gksum, double* nocc) {
    for (int o = 0; o < norb; o++) {
        double s = 0.0;
        for (int iw = 0; iw < nw; iw++) s = s + gksum[2 * ((size_t)iw * norb + o)];
        nocc[o] = 2.0 * s / beta + 0.50;
    }
}
