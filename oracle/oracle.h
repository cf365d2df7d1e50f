/* CPU oracle: a plain-C restatement of the reference ED impurity solver's hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library links, loads or calls
 * this code; only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * use it, and only as the checker.
 *
 * Every function follows one routine of the reference (yw-choi/MO-DMFT) step by
 * step, with the same loop order and the same floating-point operation order, so
 * its outputs are bit-identical to a strict-IEEE build of the reference.  The
 * citation above each function names the reference lines it restates.
 *
 * Conventions (as in the reference):
 *   - sites are 1-based in the reference; here isite is 1-based in arguments that
 *     mirror reference arguments and 0-based in array indexing;
 *   - a Fock state is the integer up + 2^nsite * down (ed_basis.F90:123);
 *   - vk is column-major (norb, nbath): V(orb o, bath b) = vk[o + norb*b];
 *   - vectors are indexed i = iup + nup * idown (0-based).
 */
#ifndef MOED_ORACLE_H
#define MOED_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int32_t norb, nbath, nsite;
    double U, J;
    const double* ek;   /* [nsite] */
    const double* vk;   /* [norb*nbath] column-major */
    const double* tij;  /* [norb*norb] column-major or NULL: impurity hopping extension (not in reference) */
} or_model;

typedef struct {
    int32_t nsite, ne_up, ne_dn;
    int64_t nup, ndn, ntot;
    int32_t* up;      /* [nup] states, ascending */
    int32_t* dn;      /* [ndn] */
    int32_t* idx_up;  /* [2^nsite]: 1-based rank or 0 */
    int32_t* idx_dn;
} or_basis;

int32_t or_icom(int32_t n, int32_t m);
int or_basis_make(int32_t nsite, int32_t ne_up, int32_t ne_dn, or_basis* b);
void or_basis_free(or_basis* b);
int64_t or_basis_get(const or_basis* b, int64_t i1);      /* 1-based index -> state */
int64_t or_basis_idx(const or_basis* b, int64_t state);   /* state -> 1-based index */

/* elementary operators (ed_operators.F90:273-348); isite 1-based, ispin 1..2 */
void or_destruction_op(const or_basis* b, int64_t in, int isite, int ispin, int64_t* out, int* sgn);
void or_creation_op(const or_basis* b, int64_t in, int isite, int ispin, int64_t* out, int* sgn);

/* B = H A over the full sector (ed_hamiltonian.F90:88-191); nthreads<=0: all */
void or_multiply_H(const or_model* m, const or_basis* b, const double* A, double* B, int nthreads);
/* the diagonal element of row i (0-based), same summation order */
double or_diag_element(const or_model* m, const or_basis* b, int64_t i0);
/* nonzero elements of row i (0-based) of H in the terms of multiply_H, column indices 0-based:
 * returns the count (<= cap) written to cols/vals, diagonal first */
int or_h_row(const or_model* m, const or_basis* b, int64_t i0, int64_t* cols, double* vals, int cap);

/* vec_out = c^+ vec (pm=1) or c vec (pm=2) into basis_out (ed_operators.F90:10-61) */
void or_apply_c(const or_basis* b, const double* vec, int pm, int iorb, int ispin,
                const or_basis* bout, double* vec_out);

/* lanczos_iteration (ed_lanczos.F90:14-68); a,b have nstep+1 entries */
int or_lanczos_iteration(const or_model* m, const or_basis* b, const double* vec, int nstep,
                         int* nstep_calc, double* a, double* bb, int nthreads);

/* boltzmann_factor (ed_solver.F90:193-219) and SORT (func.f90:301-353, 1-based indx) */
void or_boltzmann_factor(const double* elv, int n, double beta, double* p);
void or_sort(int n, const double* arrin, int* indx);

/* Matsubara grid of ed_green_init (ed_green.F90:45,50), 1 rank */
void or_omega_grid(int nw, double beta, double* omega);

/* continued fractions of green_diag (ed_green.F90:160-168, 176-184).
 * sign=+1: particle part b/(z-a-...), z=cmplx(E,w); sign=-1: hole part b/(z+a-...), z=cmplx(-E,w).
 * compat=1 reproduces the default-kind (REAL*4) rounding of cmplx(); g is interleaved re,im [nw] */
void or_green_cf(int sign, double E, int nstep_calc, const double* a, const double* b,
                 int nw, const double* omega, int compat, double* g);

/* green_diag (ed_green.F90:136-185): accumulates pev*factor*(G+ + G-) into Gr[nw] (interleaved) and
 * returns the Lanczos records.  ap..bn have nstep+1 entries. */
int or_green_diag(const or_model* m, const or_basis* b, const double* eigvec, double E, double pev,
                  double factor, int iorb, int ispin, int nstep, int nw, const double* omega, int compat,
                  double* Gr, int* nsp, int* nsn, double* ap, double* bp, double* an, double* bn,
                  int nthreads);

/* calc_dev (ed_green.F90:260-273): Dev[norb][nw] interleaved, column-major (norb, nw) */
void or_calc_dev(const or_model* m, const double* ef, int nw, const double* omega, int compat, double* Dev);

/* bath-fit objective (func.f90:1-216), x packed as in ev_to_x */
void or_x_to_ev(const double* x, int nxsize, int ns, int nbath, int norb, double* ef, double* ek, double* vk);
void or_ev_to_x(const double* ek, const double* vk, const double* ef, int ns, int nbath, int nxsize, double* x);
double or_func(const double* x, int ns, int nw, int nxsize, const double* omega, const double* Dev,
               int nbath, int norb, int compat);
void or_dfunc(const double* x, double* p, int nw, int ns, int nxsize, const double* omega,
              const double* Dev, int nbath, int norb, int compat);
/* frprmn driver (frprmn.f90): returns 0 ok, -1 itmax exceeded, -2 brent itmax exceeded.
 * nfev counts FUNC evaluations. */
int or_frprmn(double* p, int n, double ftol, int* iter, double* fret, int nw, int ns,
              const double* omega, const double* Dev, int nbath, int norb, int compat, int* nfev);

/* lattice step.  ed_set_band_structure / find_hk (ed_hamiltonian.F90:29-85): hk[q*norb + o] =
 * Hk(o,o,q) = -0.5 (cos kx + cos ky) on the nk x nk grid, nk = int(sqrt(Nq)).
 * ed_delta_new (ed_green.F90:187-215): Gksum = (1/Nq) sum_q 1/((1/Gr + D_ev + rMu) - Hk(q)),
 * D_ev <- 1/Gr - 1/Gksum + D_ev; gr, dev, gksum interleaved complex, column-major (norb, nw).
 * Gksum/float(Nq) is the complex division by (Nq, 0), as the reference build evaluates it.
 * n_from_gksum (ed_green.F90:222-253): nocc[o] = 2 sum_w Re Gksum(o, w) / beta + 0.5 */
void or_band_tb(int norb, int nq, double* hk);
void or_delta_new(int norb, int nw, int nq, double rmu, const double* hk, const double* gr, double* dev,
                  double* gksum);
void or_n_from_gksum(int norb, int nw, double beta, const double* gksum, double* nocc);

#ifdef __cplusplus
}
#endif
void or_dos_diag(int nw, const double* omega_real, double small, double ev, int nsp, int nsn, const double* ap,
                 const double* bp, const double* an, const double* bn, int compat, double* aff);
#endif
