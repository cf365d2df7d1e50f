// Internal declarations of libmoed: device table layout, kernel argument blocks and the
// host-side launchers shared by moed_kernels.hip and moed_host.hip.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/moed.h"

namespace moed {

constexpr int kMaxSite = 15;   // strings fit in uint16; reference state integer fits int32 (ed_config.f90:8)
constexpr int kMaxOrb = 4;     // impurity orbitals
constexpr int kMaxCoef = 128;  // distinct hopping coefficients (V_ob and t_ij)
constexpr int kMaxHop = 72;    // one-body moves per string (<= Norb*Nbath + Norb*(Norb-1))
constexpr int kXcd = 8;        // MI355X: 8 XCDs, workgroups dealt round-robin

// One popcount class of Nsite-bit strings ("up" or "down" strings of a sector), on the device.
// Structure-of-arrays over the string rank r (0-based, ascending integer order = the reference's
// generate_basis order, ed_basis.F90:87-105):
//   str[r]                 the bit string
//   nhop[r]                number of one-body moves with a nonzero coefficient (V_ob, t_ij)
//   mv[r*mh + k], mvc[r*mh + k]
//                          move k of string r (k < nhop[r]): target rank and signed coefficient;
//                          rows padded to mh (a multiple of 8) with target r and coefficient 0.0
//   imp[(a*norb+b)*count+r] impurity move b -> a (c+_a c_b), a != b: target rank | sign << 16, or -1
//   rec[r*rw .. r*rw+rw)   row-pass record of string r, rw = record_dwords(nslot) dwords:
//                          uint16 [0,4) the onsite prefix (sum of ek over the occupied sites in
//                          ascending site order from 0.0, the first partial sum of multiply_H's
//                          diagonal, ed_hamiltonian.F90:112-123) as a double; [4] the string;
//                          [5 + k] for slot k < nslot: the staged-row entry the move reads, rank
//                          (sign +), row_neg(count) + rank (sign -, the negated copy), or count
//                          (zero) when the slot does not apply.  A slot is a fixed (orbital, bath)
//                          pair or an ordered impurity pair with coefficient slotc[k].
// A row staged for the row pass (LDS, row_len(nU) doubles): x[k] at entry k < nU, 0.0 at nU, -x[k]
// at row_neg(nU) + k, 0.0 at row_neg(nU) + nU.  The negated copy starts at a multiple of 32
// doubles (the period of the LDS banks), so a gather's bank is the same whichever copy its sign
// selects: the up moves of neighbouring columns reach nearby ranks, and at 14 sites this cuts the
// bank-conflict cycles of the slot gathers by a quarter against a copy at nU + 1.
__host__ __device__ constexpr int row_neg(int nU) { return (nU + 32) & ~31; }
__host__ __device__ constexpr int row_len(int nU) { return row_neg(nU) + nU + 1; }

struct StrDev {
    const uint16_t* str;
    const int32_t* nhop;
    const uint32_t* mv;
    const double* mvc;
    const int32_t* imp;
    const uint4* rec;
    const uint4* rec8;  // the records with slot entries as byte offsets (8 entry); nullptr unless 8 row_len(count) <= 65536
    int32_t count;
    int32_t mh;
    int32_t nslot;
    int32_t rw;   // dwords per row-pass record (record_dwords(nslot)), the records' stride
    int32_t nel;  // electrons per string
    // the strings ordered by component under the orbital-bath moves, by rank within a component:
    // such a move keeps every orbital's count of electrons on it and its bath sites, so where each
    // bath site couples to one orbital the strings split into classes closed under those moves.
    // The order the strip pass sweeps rows in; null (rank order) for models with impurity hopping.
    const int32_t* cperm;
};

// Sector H v in two passes (moed_hv.hip).
//   row pass:   y  = diag x + T_up x           (one block stages whole rows of x in LDS)
//   strip pass: y += T_dn x + spin flip + pair exchange, then y *= 1/yscale and the x.y partials
//               (a 64-column strip of x over all rows stays in one XCD's L2)
struct HvArgs {
    const double* x;
    double* y;
    StrDev up, dn;
    const double* ek;     // [nsite]
    const double* coef;   // [kMaxCoef]: V_ob at o + norb*b, then t_ij; the last entry is 0.0
    const double* slotc;  // [up.nslot] coefficients of the slots
    double U, UmJ3, Um2J, J;  // U, U-3J, U-2J, J exactly as the reference forms them (ed_operators.F90:156-162)
    double* dot_part;         // optional: per-block partial of sum_i x_i y_i (nullptr = off)
    const double* yscale;     // optional device scalar: y *= 1/yscale[0] (nullptr = off)
    const int32_t* stop;      // optional device flag: both passes are no-ops when *stop != 0
    int32_t rows_b;           // strip pass: rows per block
    int32_t items_per_xcd;    // strip pass: (strip, row block) items per XCD
    int64_t ldx;              // row stride of x and y in doubles (nU, or nU + 1 for padded odd rows)
};

constexpr int kStripThreads = 256;

// launch geometry of the two passes for a sector with nU x nD elements
struct HvPlan {
    int threads_a, grid_a;              // row pass: one block per row
    int grid_p;                         // persistent row pass: blocks (0: not used)
    int cpl;                            // strip pass: columns per lane (strip width 64 * cpl)
    int rows_b, items_per_xcd, grid_b;  // grid_b = number of x.y partials
};
// ldx: row stride of the vectors (nU, or more for padded rows); x and y 8-byte aligned
HvPlan hv_plan(int nU, int nD, int threads_a, int strip, int ncu, int64_t ldx);
hipError_t launch_hv(const HvArgs& a, const HvPlan& plan, int norb, hipStream_t s);
// record dwords the row pass reads for nslot move slots (3, 5, 7, 9, 12, 16, 24 or 32; 0 for too many)
int record_dwords(int nslot);

// ------------------------------------------------------------------ multi-problem engine (moed_eig.hip)
//
// Several independent problems (sectors, or Lanczos runs in different sectors) advance together:
// their vectors are packed end to end (problem p at offset off[p], length dim[p]) and one launch
// serves all of them.  Vector work is split into items of kEigChunk elements that never straddle
// two problems; the H v of small sectors is one item per (problem, row).

constexpr int kEigChunk = 1024;   // elements per vector item (256 threads x 4)
constexpr int kEigThreads = 256;

struct BatchProbDev {
    StrDev up, dn;
    int64_t off;
};

// single-pass H v over the rows of many small sectors: y_p = H_p x_p (/ scale_p)
struct HvBatchArgs {
    const BatchProbDev* prob;
    const int32_t* row_prob;   // [nrow items]
    const int32_t* row_id;     // [nrow items] down-string rank of the item
    const double* x;           // packed
    double* y;                 // packed
    const double* ek;
    const double* slotc;
    double U, UmJ3, Um2J, J;
    const double* scale;       // optional: y_p *= 1 / scale[p * scale_stride]
    int32_t scale_stride;
    const int32_t* skip;       // optional: problem p is a no-op when skip[p] != 0
};
hipError_t launch_hv_batch(const HvBatchArgs& a, int nitems, int max_nU, int norb, int nslot, hipStream_t s);

// vector items and per-problem bookkeeping of the engine
struct EigArgs {
    int32_t P, ni;
    int32_t ldh;               // row length of the per-problem scalar tables (>= M + 2)
    const int32_t* item_prob;  // [ni]
    const int64_t* item_start; // [ni] packed index of the item's first element
    const int32_t* item_len;   // [ni]
    const int32_t* pitem0;     // [P] first item of problem p
    const int32_t* pnitem;     // [P]
    const int64_t* poff;       // [P] packed offset of problem p
    const int64_t* pdim;       // [P] its length
    const int32_t* mmax;       // [P] basis size cap
    int32_t* done;             // [P] 1 when the problem has stopped
    int32_t* pass2;            // [P] DGKS second pass requested
    int32_t* mfinal;           // [P] basis size at stop
    double* alpha;             // [P][ldh]
    double* beta;              // [P][ldh]
    double* nrm;               // [P][ldh] norms of the unnormalised basis vectors
    double* h;                 // [P][ldh] inner products of the current pass
    double* c;                 // [P][ldh] projection coefficients
    double* aux;               // [P][4]: alpha accumulator, |w|^2, |w'|^2
    double* part;              // [(M + 2) * ni] per-item partial sums
    double* om;                // [P][3][ldh] estimated overlaps omega_{k-1}, omega_k, omega_{k+1} (rotating)
    int32_t* full;             // [P] 1: the current step projects against the whole basis
    int32_t* force;            // [P] 1: the next step does too (the step after a triggered one)
    int32_t full_mode;         // 1: every step projects twice against the whole basis (CGS2) on the
                               // matrix cores, and the Ritz product runs there too
    const double* above;       // [P] or null: a Ritz pair whose residual interval lies above above[p]
                               // needs no convergence (the caller asks only what lies below)
    int32_t* cflag;            // [P] basis size at which the device convergence test passed (0: not
                               // yet; -1: a problem of the persistent launch, tested there)
};
// part[item] = sum over the item of a*b ; then out[p * stride] = sum over p's items (ordered) (op 1: sqrt)
hipError_t launch_eig_segdot(const EigArgs& e, const double* a, const double* b, double* out, int stride, int op,
                             hipStream_t s);
// one Gram-Schmidt pass at step k (V: device table of the basis vectors' addresses, each vector
// packed like w; vk1: the address of vector k + 1), c_i = <V_i, src> / |V_i|^2:
//   pass 0: V_{k+1} = w - sum_{i = k-1, k} c_i V_i (the three-term part); |V_{k+1}|^2 is the DGKS
//           reference; the overlap estimates decide whether the step needs pass 1 (partial
//           reorthogonalisation) — if not, the step's alpha, beta and stop test
//   pass 1: (problems flagged) V_{k+1} -= sum_{i <= k} c_i V_i in place; DGKS test against the reference
//   pass 2: (problems flagged) pass 1 again; after pass 1 or 2 the step's alpha, beta and stop test
hipError_t launch_eig_pass(const EigArgs& e, const double* const* V, double* vk1, const double* w, int k, int pass,
                           double eta2, double tiny, hipStream_t s);
// deflated operator H' = H + sigma_p sum_i L_i L_i^T on the step's product: w += sigma_p sum_i
// <L_i, V_k / |V_k|> L_i for the locked vectors L_0..L_{nl-1} (packed like w; zero vectors where a
// problem has fewer), w = H V_k / |V_k| on entry
hipError_t launch_eig_deflate(const EigArgs& e, const double* const* L, int nl, const double* sigma,
                              const double* vk, double* w, int k, hipStream_t s);
// convergence after a step with basis size m, on the device: for every running problem, the lowest
// nev Ritz pairs of T_m pass |beta_m z_{m,i}| <= tol |theta_i| -> done = 1, mfinal = m
// (zs: [P][nev][3 ldh] workspace)
hipError_t launch_eig_conv(const EigArgs& e, int m, int nev, double tol, double* zs, hipStream_t s);
// the verdicts of launch_eig_conv in stream order: done = 1, mfinal = cflag where cflag > 0 and the
// problem has not stopped at or before that basis size
hipError_t launch_eig_conv_apply(const EigArgs& e, hipStream_t s);
// Ritz vectors: out[p] + j*dim_p = sum_{i < mfinal[p]} S[(p*nev + j)*lds + i] V_i, j < nf[p], normalised
hipError_t launch_eig_ritz(const EigArgs& e, const double* const* V, const double* S, int lds, int nev,
                           const int32_t* nf, double* const* out, double* n2, hipStream_t s);

// Persistent work-queue eigensolver (moed_persist.hip): one launch runs the whole Lanczos iteration
// with partial reorthogonalisation and the convergence test of every queued small problem, one
// workgroup per problem at a time, pulling problems from *head (0 at launch) in `order`.
constexpr int kPersistThreads = 1024;
constexpr int kPersistMaxBasis = 402;  // ldh cap: 400 + 2
constexpr int kPersistMaxU = 1024;     // up strings per sector row
constexpr int64_t kPersistMaxDim = (int64_t)1 << 20;
constexpr int64_t kPersistDefaultDim = 8192;  // default size limit of the persistent launch
struct PersistArgs {
    EigArgs e;                  // alpha, beta, nrm, om, done, mfinal, mmax, pdim, ldh (zeroed tables)
    const BatchProbDev* prob;
    double* const* V;           // addresses of the basis vectors 0..mmax (packed like w)
    double* W;                  // packed work vector
    const double* ek;
    const double* slotc;
    double U, UmJ3, Um2J, J;
    const uint64_t* seed;       // [P] start-vector seeds
    const int32_t* order;       // [nq] queued problems, largest first
    int32_t nq;
    int32_t* head;              // queue head
    int32_t nev;
    int32_t max_nU;             // largest up-string count of the queued problems
    double tol;                 // |beta_m z_{m,i}| <= tol |theta_i|
    double* zs;                 // [P][nev][3 ldh] inverse-iteration workspace
};
hipError_t launch_eig_persistent(const PersistArgs& a, int grid, int norb, int nslot, hipStream_t s);

// many reference recurrences at once (the Green's-function chains), packed like the eigensolver's
// problems; A, B, stop, nsc per chain (A, B with row length ldh)
hipError_t launch_lzb_update(const EigArgs& e, int j, const double* w, const double* u1, const double* u0,
                             double* out, const double* A, const double* B, int ldh, const int32_t* stop,
                             hipStream_t s);
// which = 0: B[p][j] = sum of part over p's items; 1: A[p][j] = sum / B[p][j] and the breakdown test
hipError_t launch_lzb_finish(const EigArgs& e, int j, int which, double* A, double* B, int ldh, int32_t* stop,
                             int32_t* nsc, double thr, hipStream_t s);
hipError_t launch_lzb_dot(const EigArgs& e, const double* a, const double* b, const int32_t* stop, hipStream_t s);

// apply_c: per target string, the source rank and sign of the elementary operator
struct ApplyCArgs {
    const double* x;
    double* y;
    const int32_t* map;  // [n_target strings of the changed spin]: src rank | (sign bit 30), or -1
    int32_t nU_out, nD_out, nU_in;
    int32_t spin;  // 0: up strings change (column map), 1: down strings change (row map)
};

// launchers (moed_kernels.hip)
hipError_t launch_apply_c(const ApplyCArgs& a, hipStream_t s);
hipError_t launch_fill_start(uint64_t seed, int64_t n, double* x, hipStream_t s);
hipError_t launch_fill(double v, int64_t n, double* x, hipStream_t s);
// dst[id * ldx + iu] = src[id * nU + iu]: rows of nU doubles copied to a stride of ldx (the padding
// columns of dst are left as they are)
hipError_t launch_pad_rows(const double* src, int64_t nU, int64_t nD, double* dst, int64_t ldx, hipStream_t s);
hipError_t launch_scale_copy(const double* x, double alpha, int64_t n, double* y, hipStream_t s);

// deterministic reductions: partial sums per block then one final block in fixed order
constexpr int kRedBlocks = 1024;
hipError_t launch_dot_partial(const double* a, const double* b, int64_t n, double* part, hipStream_t s);
// out[slot] = op(sum(part[0:npart]))  op: 0 plain, 1 divide by scal[div_slot], 2 sqrt
hipError_t launch_reduce_final(const double* part, int npart, double* out, int slot, int op, const double* divp,
                               hipStream_t s);

// Lanczos (reference recurrence, ed_lanczos.F90:37-68) on unnormalised vectors u_j:
//   u_j = (w - A[j-1] u_{j-1}) / sqrt(B[j-1]) - sqrt(B[j-1]/B[j-2]) u_{j-2}
// scal layout: A[0..nstep], B[0..nstep] (B[0] = |u_0|^2 here), stop flag and nstep_calc as int at the end
struct LzScal {
    double* A;
    double* B;
    int32_t* stop;
    int32_t* nsc;
};
hipError_t launch_lz_update(int j, const double* w, const double* u1, const double* u0, double* out, int64_t n,
                            LzScal sc, double* part, hipStream_t s);
// B[j] = sum(part); for j >= 2 nothing else.  A[j] = sum(part)/B[j]; if j>=2 and B[j] < thr: stop=1, nsc=j.
hipError_t launch_lz_finish_b(int j, const double* part, int npart, LzScal sc, hipStream_t s);
hipError_t launch_lz_finish_a(int j, const double* part, int npart, LzScal sc, double thr, hipStream_t s);

// continued fraction over all frequencies, accumulating w*g into G (interleaved complex, stride)
struct CfArgs {
    const double* a;
    const double* b;
    int32_t n;      // nstep_calc
    int32_t sign;   // +1 particle, -1 hole
    double E;
    double weight;  // pev*factor
    const double* omega;
    int32_t nw;
    int32_t compat;
    double* G;      // interleaved complex, element i at G[2*i*stride]
    int32_t stride;
};
hipError_t launch_cf(const CfArgs& a, hipStream_t s);
// the continued fractions of many chains: G(o, w_i) += weight_c g_c(w_i) over the chains c of
// orbital o in chain order (the order of the sequential accumulation).  Chain c: coefficients
// A[c][0..nsc[c]], B[c][1..nsc[c]] (row length ldh) with b0[c] in place of B[c][0], sign[c] (+1
// particle, -1 hole), E[c], weight[c]; chains with use[c] == 0 contribute nothing.
struct CfBatchArgs {
    int32_t nchain, ldh, norb;
    const double* A;
    const double* B;
    const double* b0;
    const int32_t* nsc;
    const int32_t* sign;
    const int32_t* iorb;  // 0-based
    const int32_t* use;
    const double* E;
    const double* weight;
    const double* omega;
    int32_t nw, compat;
    double* G;  // interleaved complex (norb, nw) column-major
};
hipError_t launch_cf_batch(const CfBatchArgs& a, hipStream_t s);

// real-axis spectral function of one record (dos_diag, dos.f90:65-106):
// out[i * stride] = -Im(G+ + G-)(w_i) / pi, or out += weight * that when accumulate
struct DosArgs {
    const double* ap;  // ap[0..nsp], bp[0..nsp], an[0..nsn], bn[0..nsn] (device)
    const double* bp;
    const double* an;
    const double* bn;
    int32_t nsp, nsn;
    double ev, small, weight;
    const double* omega;  // real frequencies
    int32_t nw, compat, accumulate;
    double* out;
    int32_t stride;
};
hipError_t launch_dos(const DosArgs& a, hipStream_t s);
hipError_t launch_cf_adjoint(const CfArgs& a, const double* gbar, double* scratch, double* ga, double* gb,
                             hipStream_t s);

// lattice step ed_delta_new (ed_green.F90:187-215); complex arrays interleaved, column-major (norb, nw)
struct DeltaArgs {
    int32_t norb, nw, nq;
    double rmu;
    const double* hk;  // [nq][norb] diagonal of H(k)
    const double* gr;
    double* dev;       // in: D_ev, out: new D_ev
    double* gksum;     // optional
};
hipError_t launch_delta_new(const DeltaArgs& a, hipStream_t s);

// Delta(i w) - Dev of one parameter vector (func.f90:19-29), used by calc_dev with a zero target
struct FitArgs {
    const double* x;  // [nx], packed as ev_to_x
    int32_t nx, norb, nbath;
    const double* omega;
    const double* dev;  // interleaved complex (norb, nw) column-major
    int32_t nw, compat;
    double* work;       // out: interleaved complex (norb, nw)
};
hipError_t launch_fit_residual(const FitArgs& a, hipStream_t s);

}  // namespace moed
