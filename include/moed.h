/* moed.h -- C ABI of the MI355X exact-diagonalisation impurity solver (libmoed.so).
 *
 * The entry points replace the hot-path procedures of the reference DMFT-ED code
 * (yw-choi/MO-DMFT); each declaration cites the reference interface it stands in for.
 * The reference keeps its state in Fortran module globals (ed_config.f90:11-37); here the
 * same state lives in an opaque handle created from moed_params.
 *
 * Conventions
 *   - every pointer argument is a HOST pointer unless the name says _device;
 *   - arrays are Fortran column-major, indices are int32 and 1-based where the reference's
 *     are (orbital, spin, sector, level), vectors are 0-based C arrays of doubles laid out
 *     as in the reference: element (iup, idown) at iup + nup*idown (ed_basis.F90:120-123);
 *   - complex arrays are interleaved (re, im) doubles (Fortran double complex layout);
 *   - every call is blocking (the stream is synchronised before return) except the _device
 *     variants, which are ordered on the stream passed in;
 *   - return value: MOED_OK (0) or a negative MOED_E* code; the message is available from
 *     moed_last_error().  No exception crosses the ABI.
 *   - one handle = one HIP stream = one host thread at a time (not re-entrant).
 */
#ifndef MOED_H
#define MOED_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MOED_OK 0
#define MOED_EARG -1     /* bad argument */
#define MOED_ESECTOR -2  /* sector out of range or empty target sector */
#define MOED_EZERO -3    /* zero start vector (c+|psi> = 0) */
#define MOED_EHIP -4     /* HIP runtime error */
#define MOED_ENOMEM -5   /* device allocation failed */
#define MOED_ESTATE -6   /* call order: e.g. moed_green before moed_solve */
#define MOED_ECONV -7    /* iteration limit reached (eigensolver, CG, Brent) */

/* flags */
#define MOED_F_EXACT_CMPLX 1u   /* evaluate G(iw), Delta(iw) at exact double E, w_n instead of the
                                   reference's default-kind cmplx() (REAL*4-rounded) values */
/* (bits 2u and 4u are reserved) */
#define MOED_F_BATCHED 8u       /* moed_solve: all sectors in one multi-problem Lanczos run (one launch
                                   per operation for every sector) instead of sector by sector */
#define MOED_F_FIT_TREE 16u     /* bath fit: sum chi^2 and its gradient over frequencies in a fixed tree
                                   order instead of the reference's sequential order: deterministic and
                                   much faster, rounded differently (the CG path then differs in the
                                   last bits from the reference's) */
#define MOED_F_NO_DEGENERACY_CHECK 64u  /* moed_solve with nev >= 2: skip the check for levels a
                                   single-start Lanczos run misses (each degenerate level is otherwise
                                   found with its multiplicity by deflated restarts) */
#define MOED_F_FULL_REORTH 32u  /* eigensolver: full reorthogonalisation, two classical Gram-Schmidt
                                   passes against the whole Krylov basis at every step (CGS2), the
                                   contractions and the Ritz product on the FP64 matrix cores, instead
                                   of the default partial reorthogonalisation */

typedef struct moed_handle moed_handle;

/* Configuration: the reference's ed_config.f90 globals. */
typedef struct moed_params {
    int32_t norb;          /* Norb  (ed_config.f90:53) */
    int32_t nsite;         /* Nsite = Norb + Nbath (ed_config.f90:55), <= 15 */
    double U, J;           /* U, Jex (ed_config.f90:82-83) */
    double beta;           /* beta (ed_config.f90:85) */
    int32_t nstep;         /* Nstep: continued-fraction Lanczos steps (ed_config.f90:118) */
    int32_t nev;           /* Nev: eigenpairs per sector (ed_config.f90:124) */
    int32_t nw;            /* Nw: Matsubara frequencies (ed_config.f90:126) */
    const double* ek;      /* ek(1:Nsite) */
    const double* vk;      /* vk(1:Norb, Norb+1:Nsite), column-major */
    const double* tij;     /* Norb x Norb impurity hopping, column-major, or NULL (not in reference) */
    int32_t nsector;       /* Nsector and the Sectors block: nelec(:), nup(:) (ed_config.f90:66-80) */
    const int32_t* nelec;
    const int32_t* nup;
    uint32_t flags;        /* MOED_F_* */
    double gs_tol;         /* relative residual tolerance of the eigensolver (0 -> 1e-12) */
    int32_t gs_maxvec;     /* Krylov basis cap of the eigensolver (0 -> 400) */
    int32_t device;        /* HIP device ordinal */
} moed_params;

/* create / destroy.  Replaces the module initialisation ed_read_options + ed_green_init
 * (ed_config.f90:43, ed_green.F90:25). */
int moed_create(const moed_params* p, moed_handle** h);
int moed_destroy(moed_handle* h);
/* copies the last error message (NUL-terminated, truncated to len). */
int moed_last_error(const moed_handle* h, char* buf, int32_t len);

/* new bath (ek, vk), as the DMFT loop's mpi_bcast of ek/vk after the fit (main.f90:103-106). */
int moed_set_bath(moed_handle* h, const double* ek, const double* vk);

/* sector dimension and basis tables: generate_basis (ed_basis.F90:35-107).
 * up_states[C(Nsite,nup_e)], dn_states[C(Nsite,ndn_e)] ascending; either may be NULL. */
int moed_sector_dim(moed_handle* h, int32_t nup_e, int32_t ndn_e, int64_t* dim);
int moed_basis(moed_handle* h, int32_t nup_e, int32_t ndn_e, int32_t* up_states, int32_t* dn_states);

/* y = H x on sector (nup_e, ndn_e): multiply_H (ed_hamiltonian.F90:88-191). */
int moed_hv(moed_handle* h, int32_t nup_e, int32_t ndn_e, const double* x, double* y);
/* same on device pointers, ordered on `stream` (hipStream_t, NULL = the handle's stream). */
int moed_hv_device(moed_handle* h, int32_t nup_e, int32_t ndn_e, const double* x_device, double* y_device,
                   void* stream);

/* vec_out = c+_{iorb,ispin} vec (pm=1) or c_{iorb,ispin} vec (pm=2) into the N+-1 sector:
 * apply_c (ed_operators.F90:10-61).  y has the target sector's dimension. */
int moed_apply_c(moed_handle* h, int32_t pm, int32_t iorb, int32_t ispin, int32_t nup_e, int32_t ndn_e,
                 const double* x, double* y);

/* lanczos_iteration (ed_lanczos.F90:14-117): a(0:nstep), b(0:nstep) with b = squared norms,
 * b(0) = 0, breakdown when b(j) < 1.e-12 (REAL*4 literal) sets nstep_calc = j. */
int moed_lanczos(moed_handle* h, int32_t nup_e, int32_t ndn_e, const double* vec, int32_t nstep,
                 int32_t* nstep_calc, double* a, double* b);

/* ed_solve (ed_solver.F90:28-105): lowest nev eigenpairs of every sector (diag, :107-191),
 * SORT + boltzmann_factor, keep p > 1e-3 among the lowest nev, renormalise.
 * Outputs have room for nev entries; sector and level are 1-based.  Eigenvectors stay on
 * the device in the handle for moed_green.  Unfilled eigenvalues are DBL_MAX (the reference's
 * huge(1d0) for unconverged pairs). */
int moed_solve(moed_handle* h, int32_t* nev_calc, double* val, double* prob, int32_t* sector,
               int32_t* level);
/* eigenvalues of one sector after moed_solve: ev[nev] (eigval_sec of ed_solver.F90:53). */
int moed_sector_eigenvalues(moed_handle* h, int32_t isector, double* ev);
/* copy eigenvector (isector, ilevel) to the host: import_eigvec (ed_io.F90:105-139). */
int moed_eigvec(moed_handle* h, int32_t isector, int32_t ilevel, double* vec);

/* ed_calc_green_ftn (ed_green.F90:60-134) over the states kept by moed_solve:
 * Gr(Norb, nw) complex, accumulated as in green_diag (:136-185) at w = omega(1:nw). */
int moed_green(moed_handle* h, int32_t nw, const double* omega, double* Gr);
/* Lanczos records written to apbpanbn.dat by ed_calc_green_ftn (ed_green.F90:112-129):
 * number of records, then record irec (0-based): iev, iorb, ispin, E, p, even, nsp, nsn and
 * ap,bp,an,bn (room for nstep+1 each). */
int moed_green_nrecords(moed_handle* h, int32_t* nrec);
int moed_green_record(moed_handle* h, int32_t irec, int32_t* iev, int32_t* iorb, int32_t* ispin, double* E,
                      double* p, int32_t* even, int32_t* nsp, int32_t* nsn, double* ap, double* bp, double* an,
                      double* bn);

/* continued fraction of green_diag (ed_green.F90:160-168 sign=+1; :176-184 sign=-1) for one
 * coefficient set; g[nw] complex. */
int moed_green_cf(moed_handle* h, int32_t sign, double E, int32_t nsc, const double* a, const double* b,
                  int32_t nw, const double* omega, double* g);
/* On-disk records, in the reference's Fortran unformatted sequential layout (4-byte record
 * lengths around every record).
 * moed_eigvec_filename: eigvec_filename (ed_io.F90:141-147), "eigvec-" I2.2 "-" I3.3 "-" I4.4; a
 *   field that does not fit prints as asterisks, as in the reference (sectors >= 100 give "**").
 * moed_export_eigvec / moed_import_eigvec: export_eigvec / import_eigvec (ed_io.F90:13-33,
 *   105-139): records (nloc) and (eigvec(1:nloc)) in dir/<filename>; import with vec = NULL
 *   returns nloc only.  Host only.
 * moed_export_green: apbpanbn.dat as ed_calc_green_ftn writes it (ed_green.F90:75, 112-129): per
 *   record (nstep_calcp, nstep_calcn), then (iev, iorb, E, p, even, ap, bp, an, bn), from the
 *   records of the last moed_green.
 * moed_import_green: reads such a file (written by the reference or by moed_export_green) into the
 *   handle's records, for moed_dos and moed_green_record; records with iorb > Norb or more than
 *   nstep steps are rejected (MOED_EARG). */
int moed_eigvec_filename(int32_t isector, int32_t iev, int32_t node, char* buf, int32_t len);
int moed_export_eigvec(const char* dir, int32_t isector, int32_t iev, int32_t node, int64_t nloc, const double* vec);
int moed_import_eigvec(const char* dir, int32_t isector, int32_t iev, int32_t node, int64_t* nloc, double* vec);
int moed_export_green(moed_handle* h, const char* path);
int moed_import_green(moed_handle* h, const char* path);

/* Real-axis spectral function (dos.f90).
 * moed_dos_diag: dos_diag (dos.f90:65-106) for one record: aff[w] = -Im(G+ + G-)/pi with
 *   z+ = cmplx(ev + w, small), z- = cmplx(w - ev, small) and the green_diag recurrences over
 *   ap, bp (0..nsp) and an, bn (0..nsn).
 * moed_dos: dos (dos.f90:1-63) over the records of the last moed_green, in apbpanbn.dat order:
 *   Aff(Norb, nw) column-major.  weighted = 0 reproduces the reference, where each record overwrites
 *   its orbital's row with no probability or spin weight (dos.f90:102), so the last record per
 *   orbital remains; weighted = 1 sums p * factor * A over the records (factor 1 for N_up = N_dn,
 *   else 1/2, as ed_calc_green_ftn weights G).  The reference's grid is w_i = (i - nw/2) * 0.01
 *   (main.f90:187-193) and small = DMFT.small (0.05, ed_config.f90:116). */
int moed_dos_diag(moed_handle* h, int32_t nw, const double* omega_real, double small, double ev, int32_t nsp,
                  int32_t nsn, const double* ap, const double* bp, const double* an, const double* bn, double* aff);
int moed_dos(moed_handle* h, int32_t nw, const double* omega_real, double small, int32_t weighted, double* Aff);
/* adjoint of moed_green_cf: given gbar = dL/dRe g + i dL/dIm g (complex[nw]), returns
 * ga[j] = dL/da_j, gb[j] = dL/db_j for j = 0..nsc.  No reference counterpart: it
 * differentiates G(iw) with respect to the Lanczos coefficients. */
int moed_green_cf_adjoint(moed_handle* h, int32_t sign, double E, int32_t nsc, const double* a,
                          const double* b, int32_t nw, const double* omega, const double* gbar, double* ga,
                          double* gb);

/* calc_dev (ed_green.F90:260-273): Dev(Norb, nw) complex from (ek, vk) and impurity levels ef. */
int moed_calc_dev(moed_handle* h, const double* ek, const double* vk, const double* ef, int32_t nw,
                  const double* omega, double* Dev);

/* batched FUNC + dfunc (func.f90:1-142) for nfit parameter vectors x[nfit][nx] (nx = Nbath*(1+Norb)+Norb,
 * packed as ev_to_x, func.f90:165-191), one shared target Dev(Norb, nw): f[nfit], g[nfit][nx]. */
int moed_chi2_grad(moed_handle* h, int32_t nfit, int32_t nx, const double* x, int32_t nw, const double* omega,
                   const double* Dev, double* f, double* g);
/* minimization (func.f90:144-163 -> frprmn.f90) for nfit independent starting points at once:
 * x[nfit][nx] in/out, fret[nfit], iter[nfit], status[nfit] (0 ok, MOED_ECONV on itmax, where the
 * reference stops the program: frprmn.f90:83).  The whole minimiser runs on the device, one
 * workgroup per fit; omega must be positive and omega, Dev, x finite. */
int moed_minimize(moed_handle* h, int32_t nfit, int32_t nx, double* x, int32_t nw, const double* omega,
                  const double* Dev, double ftol, double* fret, int32_t* iter, int32_t* status);
/* moed_minimize on device buffers, allocating no device memory.  No counterpart in the reference:
 * it is for callers that cannot allocate (code under HIP stream capture, serving loops, harnesses
 * that forbid hipMalloc).  x_device[nfit][nx] in/out, omega_device[nw], Dev_device (complex
 * (Norb, nw)) in; fret_device[nfit], iter_device[nfit], status_device[nfit] out (any may be NULL),
 * all on the handle's device, ordered on `stream` (NULL = the handle's stream); the call returns
 * when the results are written.  At most 1024 fits per call (MOED_EARG beyond).  The CG state lives
 * in one fixed device array per GPU (~1.6 MB, held until the process ends) that every call on that
 * GPU shares: calls that overlap on one GPU race.  Same results as moed_minimize bit for bit;
 * moed_fit_nfev is not updated. */
int moed_minimize_device_noalloc(moed_handle* h, int32_t nfit, int32_t nx, double* x_device, int32_t nw,
                                 const double* omega_device, const double* Dev_device, double ftol,
                                 double* fret_device, int32_t* iter_device, int32_t* status_device, void* stream);
/* CG iterations each fit runs per kernel launch inside moed_minimize (0: default).  It bounds the
 * duration of one launch; results do not depend on it.  No counterpart in the reference. */
int moed_set_fit_launch(moed_handle* h, int32_t iterations);
/* FUNC evaluations (f1dim calls plus the first FUNC) of each of the first nfit fits of the last
 * moed_minimize.  No counterpart in the reference (diagnostics). */
int moed_fit_nfev(moed_handle* h, int32_t nfit, int32_t* nfev);

/* Lattice step of the DMFT loop (next to the solver: main.f90:76-79).
 * moed_band_tb: the diagonal of H(k) of ed_set_band_structure / find_hk (ed_hamiltonian.F90:29-85),
 *   hk[q*norb + o] = -0.5 (cos kx + cos ky) on the nk x nk grid, nk = int(sqrt(nq)); host only.
 * moed_delta_new: ed_delta_new + n_from_gksum (ed_green.F90:187-253) on the device:
 *   Gksum(o,w) = (1/Nq) sum_q 1/((1/Gr + Dev + rmu) - hk(q,o)),  Dev <- 1/Gr - 1/Gksum + Dev (in place),
 *   nocc[o] = 2 sum_w Re Gksum(o,w) / beta + 1/2.  Gr, Dev, Gksum: complex (norb, nw) column-major;
 *   Gksum and nocc may be NULL. */
int moed_band_tb(int32_t norb, int32_t nq, double* hk);
int moed_delta_new(moed_handle* h, int32_t nw, int32_t nq, double rmu, const double* hk, const double* Gr,
                   double* Dev, double* Gksum, double* nocc);

/* Matsubara grid of ed_green_init (ed_green.F90:45-51) at one rank. */
int moed_omega_grid(int32_t nw, double beta, double* omega);

/* counter-based start vector x_i = 2*u(splitmix64(seed ^ i)) - 1 on the device. */
int moed_fill_start_device(moed_handle* h, uint64_t seed, int64_t n, double* x_device, void* stream);

/* launch shape of the H v kernels (no counterpart in the reference): row pass threads per block
 * (0 = default: one resident 1024-thread block per compute unit that keeps its columns' records in
 * registers, where the sector has 2048 to 4093 up strings and at least two rows per compute unit,
 * else one 1024-thread block per row; 256, 512 or 1024 = one block of that size per row) and strip
 * pass width in columns (0 = default: 128 where the up-string count is even, else 64; 64 or 128).
 * Results are bit-identical for every shape.  Vectors passed to the H v need 8-byte alignment. */
int moed_set_hv_shape(moed_handle* h, int32_t threads, int32_t strip);
/* sectors of at most max_dim elements that moed_solve with MOED_F_BATCHED runs in its persistent
 * work-queue launch (no counterpart in the reference): the whole Lanczos iteration of each such
 * sector inside one workgroup, the sectors pulled from a queue, largest first; larger sectors run
 * step by step, concurrently.  -1 = default, 0 = none, at most 2^20. */
int moed_set_persistent(moed_handle* h, int64_t max_dim);
/* vectors per allocation of the eigensolver's Krylov basis (no counterpart in the reference):
 * 0 = default (what fits in 4 GiB).  The basis is allocated as the iteration reaches it, up to
 * gs_maxvec + 1 vectors, and the handle keeps it for later calls; results do not depend on the
 * chunk size.  Releases the handle's current basis. */
int moed_set_eig_chunk(moed_handle* h, int32_t vectors);
/* device pointer of the handle's stream (hipStream_t). */
void* moed_stream(moed_handle* h);

#ifdef __cplusplus
}
#endif
#endif
