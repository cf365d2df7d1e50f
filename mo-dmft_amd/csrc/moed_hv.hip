// The sector H v product (multiply_H, ed_hamiltonian.F90:88-191) for gfx950, in two passes.
//
// Vector layout in HBM: x[iu + nU*id] (up-string rank fastest, ed_basis.F90:120-123); a "row" is
// the nU doubles of one down string id, a "column" the nD doubles of one up string iu.
//   H = diag + T_up (x) 1 + 1 (x) T_dn + (spin flip, pair exchange)
// T_up moves an up electron: it mixes the elements of one row; T_dn mixes the elements of one
// column.  No single traversal order keeps both neighbourhoods on chip (a row's down-move
// neighbours are rows all over the sector), so the product runs as
//   row pass    y  = diag x + T_up x   one block stages one whole row of x in LDS and gathers
//                                      the up moves there
//   strip pass  y += T_dn x + SF + PE  a 128-column strip of x over all nD rows (nD * 1 KB,
//                                      3.5 MB at 14 sites) is worked on by one XCD at a time, so
//                                      every down-move read after the first hits that XCD's L2;
//                                      this pass also applies 1/yscale and forms the x.y partials
// HBM traffic per element: 16 bytes (row pass) + 24 bytes (strip pass: x, y in, y out).
//
// Every element's terms are summed in one fixed order, whatever the launch shape: the diagonal
// product (rounded), then fused multiply-adds of the up-move slots in slot order, the down moves
// in move-list order and spin flip / pair exchange per ordered impurity pair.  The diagonal itself is summed in the reference's order
// (ed_hamiltonian.F90:110-139); terms a state does not have are added as +0.0, which leaves the
// sum bit-identical (the sum starts at +0.0 and can never become -0.0).
#include <hip/hip_runtime.h>

#include <algorithm>

#include "moed_device.h"
#include "moed_hvops.h"
#include "moed_internal.h"

namespace moed {

// two consecutive doubles as one 16-byte access; gfx950 serves it at any 8-byte alignment (the HSA
// runtime enables unaligned access), and this type tells the compiler the address is only 8-byte
// aligned
typedef double d2u __attribute__((ext_vector_type(2), aligned(8)));

// ------------------------------------------------------------------ row pass

// Row pass, one block per row (down string id).  The row is staged in LDS twice, as x and -x:
//   entry k (0 <= k < nU): x[id][k];  entry nU: 0.0;  entry row_neg(nU) + k: -x[id][k]
// A move with sign -1 reads the negated copy, so every slot term is one fma with the slot's
// coefficient, and slots that do not apply (and the padding slots past nslot, coefficient 0.0)
// read the zero: the KMAX = 2 NW - 5 slot reads carry no branches and are all in flight
// together.  One column per thread and iteration; the column's record (onsite prefix, string,
// slot entries) is read from global memory (L2-resident, shared by all rows), one column ahead
// (the first before the row is staged), so that each wait for a record overlaps a whole column's
// work and the previous column's store.
//   BYTES   the up strings have byte-offset records (nU <= kFastMaxU): the staged row is then the
//           kernel's only LDS object, at LDS address 0, and a slot entry (8 x the entry) is the
//           address it reads, which saves the index arithmetic per slot
//   NDP     the energies of the row's occupied down sites sit in scalar registers, padded with +0.0
//           to NDP adds (S + 0.0 == S: the sum is never -0.0)
// The interaction terms use the row's RowDiag coefficients.  The diagonal is diag_element's, term
// for term.
constexpr int kFastMaxU = 4095;

template <int NORB, int NW, int NDP, bool BYTES>
__global__ void __launch_bounds__(1024) hv_rows_kernel(HvArgs P) {
    constexpr int KMAX = 2 * NW - 5;
    __shared__ __attribute__((aligned(16))) double xstat[BYTES ? row_len(kFastMaxU) : 1];
    extern __shared__ __attribute__((aligned(16))) double xdyn[];
    double* const xs = BYTES ? xstat : xdyn;

    if (P.stop && *P.stop) return;
    const int NT = blockDim.x;
    const int nU = P.up.count;
    const int id = blockIdx.x;
    const int neg = row_neg(nU);
    const int tid = threadIdx.x;
    const double* __restrict__ xr = P.x + (size_t)id * P.ldx;
    double* __restrict__ yr = P.y + (size_t)id * P.ldx;
    const uint32_t* __restrict__ recw = reinterpret_cast<const uint32_t*>(BYTES ? P.up.rec8 : P.up.rec);
    const int rstride = P.up.rw;
    uint32_t wn[NW];
    if (tid < nU) {
#pragma unroll
        for (int q = 0; q < NW; ++q) wn[q] = recw[(size_t)tid * rstride + q];
    }

    // ---- stage the row and its negation: 16-byte pairs (rows at either 8-byte phase), the last
    //      element of an odd row alone; each thread issues all its loads of a round before it writes
    //      any of them to LDS
    {
        constexpr int R = 4;
        const int np = nU / 2;
        for (int p0 = tid; p0 < np; p0 += R * NT) {
            d2u v[R];
#pragma unroll
            for (int u = 0; u < R; ++u)
                if (p0 + u * NT < np) v[u] = reinterpret_cast<const d2u*>(xr)[p0 + u * NT];
#pragma unroll
            for (int u = 0; u < R; ++u) {
                const int p = p0 + u * NT;
                if (p < np) {
                    reinterpret_cast<double2*>(xs)[p] = make_double2(v[u].x, v[u].y);
                    reinterpret_cast<double2*>(xs + neg)[p] = make_double2(-v[u].x, -v[u].y);
                }
            }
        }
        if ((nU & 1) && tid == NT - 1) {
            const double v = xr[nU - 1];
            xs[nU - 1] = v;
            xs[neg + nU - 1] = -v;
        }
    }
    const double U = P.U, UmJ3 = P.UmJ3, Um2J = P.Um2J;
    const uint32_t sd = scalar_u16(P.dn.str, id);
    double ekr[NDP];
    {
        const cptr<double> ekc = as_const(P.ek);
        uint32_t m = sd;
#pragma unroll
        for (int q = 0; q < NDP; ++q) {
            const double e = ekc[m ? __builtin_ctz(m) : 0];
            ekr[q] = m ? e : 0.0;
            m &= m - 1;
        }
    }
    const RowDiag<NORB> rd = row_diag<NORB>(sd, U, UmJ3, Um2J);
    if (tid == 0) {
        xs[nU] = 0.0;
        xs[neg + nU] = 0.0;
    }
    __syncthreads();

    const cptr<double> slotc = as_const(P.slotc);
    for (int iu = tid; iu < nU; iu += NT) {
        uint32_t w[NW];
#pragma unroll
        for (int q = 0; q < NW; ++q) w[q] = wn[q];
        if (iu + NT < nU) {
#pragma unroll
            for (int q = 0; q < NW; ++q) wn[q] = recw[(size_t)(iu + NT) * rstride + q];
        }
        double S = __hiloint2double((int)w[1], (int)w[0]);
#pragma unroll
        for (int q = 0; q < NDP; ++q) S += ekr[q];
        double acc = __dmul_rn(diag_interaction<NORB>(S, w[2] & 0xffffu, rd, UmJ3), xs[iu]);
#pragma unroll
        for (int q = 0; q < KMAX; ++q) {
            const uint32_t e = rec_u16(w, 5 + q);
            const double v = BYTES ? *reinterpret_cast<const double*>(reinterpret_cast<const char*>(xstat) + e) : xdyn[e];
            acc = fma(slotc[q], v, acc);
        }
        yr[iu] = acc;
    }
}

// Resident row pass for sectors whose row and its negation fit twice in LDS (nU <= 4093): one
// 1024-thread block per compute unit works through rows blockIdx.x, blockIdx.x + gridDim.x, ...
// Each thread keeps the byte-offset records of its columns (iu = tid + 1024 c, c < 4) in
// registers, loaded once, instead of reading them from L2 for every row (at 14 sites the per-row
// kernel reads 425 MB of records per product).  The row is staged in two LDS buffers of
// row_len(nU) doubles (layout as in hv_rows_kernel).  Step k writes the registers holding row k + 1 to the
// buffer step k - 1 read, issues the loads of row k + 2 into them, and computes row k from the
// other buffer, so a row's loads are in flight for a whole step; one barrier per step.  Every
// element is computed by hv_rows_kernel's operations in its order (bit-identical).
constexpr int kResRowsMaxU = 4093;
constexpr int kResRowsBuf = 8190;  // doubles per buffer: 65,520 bytes, so the second buffer is an
                                   // immediate LDS offset, and 16-byte aligned
constexpr int kResRowsCpt = 4;     // columns per thread: ceil(4093 / 1024)
static_assert(row_len(kResRowsMaxU) <= kResRowsBuf, "a staged row fits its buffer");

// a thread's share of one row: the 16-byte pairs p = tid + 1024 j (j < 2: nU <= 4093 has at most
// 2046 pairs) and, for an odd nU, thread 1023 also the last element
struct RowRegs {
    d2u v[2];
    double t;
};

__device__ __forceinline__ void persist_load(RowRegs& r, const double* __restrict__ xr, int nU, int tid) {
#pragma unroll
    for (int j = 0; j < 2; ++j) {
        const int p = tid + 1024 * j;
        if (p < nU / 2) r.v[j] = reinterpret_cast<const d2u*>(xr)[p];
    }
    if ((nU & 1) && tid == 1023) r.t = xr[nU - 1];
}

__device__ __forceinline__ void persist_store(const RowRegs& r, double* xs, int nU, int tid) {
    const int neg = row_neg(nU);
#pragma unroll
    for (int j = 0; j < 2; ++j) {
        const int p = tid + 1024 * j;
        if (p < nU / 2) {
            reinterpret_cast<double2*>(xs)[p] = make_double2(r.v[j].x, r.v[j].y);
            reinterpret_cast<double2*>(xs + neg)[p] = make_double2(-r.v[j].x, -r.v[j].y);
        }
    }
    if ((nU & 1) && tid == 1023) {
        xs[nU - 1] = r.t;
        xs[neg + nU - 1] = -r.t;
    }
}

// one step: stage row idn (registers -> buffer 1 - B), load row idl into registers, compute row id
// from buffer B (idn, idl < 0: none)
template <int NORB, int NW, int NDP, int B>
__device__ __forceinline__ void persist_step(const HvArgs& P, double* xb, const uint32_t (&w)[kResRowsCpt][NW],
                                             RowRegs& rr, int id, int idn, int idl) {
    constexpr int KMAX = 2 * NW - 5;
    const int nU = P.up.count, tid = threadIdx.x;
    if (idn >= 0) persist_store(rr, xb + (1 - B) * kResRowsBuf, nU, tid);
    if (idl >= 0) persist_load(rr, P.x + (size_t)idl * P.ldx, nU, tid);

    const double U = P.U, UmJ3 = P.UmJ3, Um2J = P.Um2J;
    const uint32_t sd = scalar_u16(P.dn.str, id);
    double ekr[NDP];
    {
        const cptr<double> ekc = as_const(P.ek);
        uint32_t m = sd;
#pragma unroll
        for (int q = 0; q < NDP; ++q) {
            const double e = ekc[m ? __builtin_ctz(m) : 0];
            ekr[q] = m ? e : 0.0;
            m &= m - 1;
        }
    }
    const RowDiag<NORB> rd = row_diag<NORB>(sd, U, UmJ3, Um2J);
    const cptr<double> slotc = as_const(P.slotc);
    const char* base = reinterpret_cast<const char*>(xb + B * kResRowsBuf);
    double* __restrict__ yr = P.y + (size_t)id * P.ldx;
#pragma unroll
    for (int c = 0; c < kResRowsCpt; ++c) {
        const int iu = tid + 1024 * c;
        if (iu < nU) {
            double S = __hiloint2double((int)w[c][1], (int)w[c][0]);
#pragma unroll
            for (int q = 0; q < NDP; ++q) S += ekr[q];
            double acc = __dmul_rn(diag_interaction<NORB>(S, w[c][2] & 0xffffu, rd, UmJ3),
                                   *reinterpret_cast<const double*>(base + 8 * iu));
#pragma unroll
            for (int q = 0; q < KMAX; ++q) {
                const uint32_t e = rec_u16(w[c], 5 + q);
                acc = fma(slotc[q], *reinterpret_cast<const double*>(base + e), acc);
            }
            yr[iu] = acc;
        }
    }
    __syncthreads();
}

template <int NORB, int NW, int NDP>
__global__ void __launch_bounds__(1024) hv_rows_persist_kernel(HvArgs P) {
    __shared__ __attribute__((aligned(16))) double xb[2 * kResRowsBuf];
    if (P.stop && *P.stop) return;
    const int nU = P.up.count, nD = P.dn.count, tid = threadIdx.x;
    const int G = gridDim.x, r0 = blockIdx.x;
    const int K = (nD - r0 + G - 1) / G;  // rows of this block: r0 + G k, k < K (K >= 1)
    const uint32_t* __restrict__ recw = reinterpret_cast<const uint32_t*>(P.up.rec8);
    const int rstride = P.up.rw;
    uint32_t w[kResRowsCpt][NW];
#pragma unroll
    for (int c = 0; c < kResRowsCpt; ++c) {
        const int iu = min(tid + 1024 * c, nU - 1);
#pragma unroll
        for (int q = 0; q < NW; ++q) w[c][q] = recw[(size_t)iu * rstride + q];
    }
    // prologue: row 0 of the block into buffer 0, row 1 into registers, the zero entries of both buffers
    RowRegs rr;
    persist_load(rr, P.x + (size_t)r0 * P.ldx, nU, tid);
    persist_store(rr, xb, nU, tid);
    if (K > 1) persist_load(rr, P.x + (size_t)(r0 + G) * P.ldx, nU, tid);
    if (tid < 2) {
        xb[tid * kResRowsBuf + nU] = 0.0;
        xb[tid * kResRowsBuf + row_neg(nU) + nU] = 0.0;
    }
    __syncthreads();
    for (int k = 0; k < K; k += 2) {
        persist_step<NORB, NW, NDP, 0>(P, xb, w, rr, r0 + G * k, k + 1 < K ? r0 + G * (k + 1) : -1,
                                              k + 2 < K ? r0 + G * (k + 2) : -1);
        if (k + 1 >= K) break;
        persist_step<NORB, NW, NDP, 1>(P, xb, w, rr, r0 + G * (k + 1), k + 2 < K ? r0 + G * (k + 2) : -1,
                                              k + 3 < K ? r0 + G * (k + 3) : -1);
    }
}

// ------------------------------------------------------------------ strip pass

// CPL consecutive columns of one row as one load.  CPL = 2 is one 16-byte load, which gfx950
// serves at any 8-byte alignment (the HSA runtime enables unaligned access), so rows of odd
// stride take it too; the type below tells the compiler the address is only 8-byte aligned.
template <int CPL>
struct Cols {
    double v[CPL];
};
template <int CPL>
__device__ __forceinline__ Cols<CPL> load_cols(const double* p) {
    Cols<CPL> c;
    if constexpr (CPL == 2) {
        const d2u t = *reinterpret_cast<const d2u*>(p);
        c.v[0] = t.x;
        c.v[1] = t.y;
    } else {
        c.v[0] = *p;
    }
    return c;
}

// Work item (strip, row block): strip s = columns [W s, W s + W) with W = 64 CPL, row block b =
// rows [rows_b b, rows_b (b + 1)).  Items are numbered strip-major and each XCD (blocks are dealt
// to the 8 XCDs round-robin) takes a contiguous range, so the blocks resident on one XCD at any
// time share one or two strips.  Each wave handles whole rows of its item's strip, CPL columns
// per lane: the row's down moves and down impurity moves are wave-uniform scalar loads, and a
// move that does not apply issues no load.
template <int NORB, int CPL>
__global__ void __launch_bounds__(kStripThreads) hv_strip_kernel(HvArgs P) {
    constexpr int NW = kStripThreads / 64;
    constexpr int NP = NORB * NORB;
    constexpr int MU = NORB * NORB / 4;  // most ordered impurity pairs one up string can move
    constexpr int W = 64 * CPL;
    __shared__ double s_red[NW];

    if (P.stop && *P.stop) return;
    const int nU = P.up.count, nD = P.dn.count;
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = uniform_i32(tid >> 6);  // makes every per-row quantity below wave-uniform
    const int nstrip = (nU + W - 1) / W;
    const int bps = (nD + P.rows_b - 1) / P.rows_b;
    const int item = (blockIdx.x % kXcd) * P.items_per_xcd + blockIdx.x / kXcd;
    const double* __restrict__ X = P.x;
    double* __restrict__ Y = P.y;
    const cptr<uint32_t> mvp = as_const(P.dn.mv);
    const cptr<double> mvcp = as_const(P.dn.mvc);
    const cptr<int32_t> nhopc = as_const(P.dn.nhop);
    const cptr<int32_t> dimpc = as_const(P.dn.imp);
    const int mh = P.dn.mh;
    const double J = P.J;
    const double inv = P.yscale ? 1.0 / P.yscale[0] : 1.0;
    double dsum = 0.0;

    if (item < nstrip * bps) {
        const int strip = item / bps, rb = item % bps;
        const int c00 = strip * W + lane * CPL;
        const bool ok = c00 < nU;
        // with CPL = 2 and an odd nU the lane of the last column takes the pair (nU - 2, nU - 1) and
        // keeps only its second column (each column's terms do not depend on its place in the pair),
        // so no load reaches past a row; idle lanes read a valid pair and keep nothing
        const bool tail = CPL == 2 && ok && c00 + 1 >= nU;
        const int iu = ok ? (tail ? nU - 2 : c00) : ((nU - CPL) & ~(CPL - 1));
        const bool ok1 = CPL == 1 || iu + 1 < nU;
        const int r1 = min(nD, (rb + 1) * P.rows_b);
        // up half of the spin-flip / pair-exchange terms, compacted per column: pair p = (i, j) needs
        // up j occupied and up i empty (move j -> i), which holds for at most MU of the ordered pairs
        // of one up string (k (NORB - k) <= NORB^2 / 4 with k impurity electrons).  Slot s of column
        // c keeps the s-th applicable pair in ascending order as (p << 24) | target | sign << 16, or
        // -1: the terms are then summed in ascending pair order, as over the full pair list.
        int32_t ua[MU > 0 ? MU : 1][CPL];
        if constexpr (NORB > 1) {
#pragma unroll
            for (int c = 0; c < CPL; ++c) {
#pragma unroll
                for (int k = 0; k < MU; ++k) ua[k][c] = -1;
                int n = 0;
#pragma unroll
                for (int p = 0; p < NP; ++p) {
                    if (p / NORB == p % NORB) continue;
                    const int32_t v = (c == 0 || ok1) ? P.up.imp[(size_t)p * nU + iu + c] : -1;
                    if (v >= 0) {
#pragma unroll
                        for (int k = 0; k < MU; ++k)
                            if (n == k) ua[k][c] = (p << 24) | v;
                        ++n;
                    }
                }
            }
        }
        for (int pos = rb * P.rows_b + wave; pos < r1; pos += NW) {
            const int id = P.dn.cperm ? uniform_i32(as_const(P.dn.cperm)[pos]) : pos;
            const size_t ro = (size_t)id * P.ldx;
            // ---- scalar loads: the row's first eight down moves and its down impurity moves
            const int nh = nhopc[id];
            const size_t m0 = (size_t)id * mh;
            uint32_t t[8];
            double cf[8];
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                t[q] = mvp[m0 + q];
                cf[q] = mvcp[m0 + q];
            }
            // down half per pair p = (i, j): spin flip moves down i -> j, pair exchange down j -> i; one
            // row has at most one of them (down i occupied and j empty, or the reverse), and both give
            // the element +-J times the two signs, so the row keeps one entry per pair
            int32_t dh[NP > 1 ? NP : 1];
            if constexpr (NORB > 1) {
#pragma unroll
                for (int p = 0; p < NP; ++p) {
                    const int i = p / NORB, j = p % NORB;
                    if (i == j) {
                        dh[p] = -1;
                        continue;
                    }
                    const int32_t sf = dimpc[(size_t)(j * NORB + i) * nD + id];
                    dh[p] = sf >= 0 ? sf : dimpc[(size_t)(i * NORB + j) * nD + id];
                }
            }
            // ---- every vector load of the row before any use
            Cols<CPL> acc = load_cols<CPL>(Y + ro + iu);
            Cols<CPL> xo;
            if (P.dot_part) xo = load_cols<CPL>(X + ro + iu);
            Cols<CPL> xv[8];
#pragma unroll
            for (int q = 0; q < 8; ++q)
                if (q < nh) xv[q] = load_cols<CPL>(X + (size_t)t[q] * P.ldx + iu);
            // each column's applicable spin-flip / pair-exchange entries: the row's entry of the slot's pair
            int32_t dsel[MU > 0 ? MU : 1][CPL];
            double gx[MU > 0 ? MU : 1][CPL];
            if constexpr (NORB > 1) {
#pragma unroll
                for (int k = 0; k < MU; ++k)
#pragma unroll
                    for (int c = 0; c < CPL; ++c) {
                        const int32_t u = ua[k][c];
                        const int pu = u >> 24;  // -1 for an empty slot
                        int32_t d = -1;
#pragma unroll
                        for (int p = 0; p < NP; ++p)
                            if (p / NORB != p % NORB) d = pu == p ? dh[p] : d;
                        dsel[k][c] = d;
                        if (d >= 0) gx[k][c] = X[(size_t)(d & 0xffff) * P.ldx + (u & 0xffff)];
                    }
            }
            // ---- down-string moves in move order
#pragma unroll
            for (int q = 0; q < 8; ++q)
                if (q < nh) {
#pragma unroll
                    for (int c = 0; c < CPL; ++c) acc.v[c] = fma(cf[q], xv[q].v[c], acc.v[c]);
                }
            for (int k0 = 8; k0 < nh; k0 += 8) {
#pragma unroll
                for (int q = 0; q < 8; ++q) {
                    t[q] = mvp[m0 + k0 + q];
                    cf[q] = mvcp[m0 + k0 + q];
                }
#pragma unroll
                for (int q = 0; q < 8; ++q)
                    if (k0 + q < nh) xv[q] = load_cols<CPL>(X + (size_t)t[q] * P.ldx + iu);
#pragma unroll
                for (int q = 0; q < 8; ++q)
                    if (k0 + q < nh) {
#pragma unroll
                        for (int c = 0; c < CPL; ++c) acc.v[c] = fma(cf[q], xv[q].v[c], acc.v[c]);
                    }
            }
            // ---- spin flip (i,j): up j->i, down i->j; pair exchange (i,j): up j->i, down j->i;
            //      element J * sign_up * sign_dn (-J of ed_operators.F90:215,268 times the operator signs)
            if constexpr (NORB > 1) {
#pragma unroll
                for (int k = 0; k < MU; ++k)
#pragma unroll
                    for (int c = 0; c < CPL; ++c) {
                        const int32_t d = dsel[k][c];
                        if (d >= 0) {
                            const double cu = ((ua[k][c] >> 16) & 1) ? -J : J;
                            acc.v[c] = fma((d >> 16) ? -cu : cu, gx[k][c], acc.v[c]);
                        }
                    }
            }
            if (P.yscale) {
#pragma unroll
                for (int c = 0; c < CPL; ++c) acc.v[c] *= inv;
            }
            if (ok) {
                if (CPL == 2 && !tail) {  // one 16-byte store (two 8-byte ones write ~2 % more)
                    const d2u v = {acc.v[0], acc.v[CPL - 1]};
                    __builtin_nontemporal_store(v, reinterpret_cast<d2u*>(Y + ro + iu));
                } else {
                    __builtin_nontemporal_store(acc.v[CPL - 1], Y + ro + iu + CPL - 1);
                }
                if (P.dot_part) {
                    if (CPL == 1 || !tail) dsum = fma(xo.v[0], acc.v[0], dsum);
                    if (CPL == 2) dsum = fma(xo.v[CPL - 1], acc.v[CPL - 1], dsum);
                }
            }
        }
    }
    if (P.dot_part) {
        const double s = block_sum<kStripThreads>(dsum, s_red);
        if (tid == 0) P.dot_part[blockIdx.x] = s;
    }
}

// ------------------------------------------------------------------ launch

// record dwords the row pass reads: the smallest of 3, 5, 7, 9, 12, 16, 24, 32 with room for
// the onsite prefix, the string and nslot slot entries (2 NW - 5 slots)
int record_dwords(int nslot) {
    const int need = (10 + 2 * nslot + 3) / 4;
    for (int nw : {3, 5, 7, 9, 12, 16, 24, 32})
        if (need <= nw) return nw;
    return 0;
}

// threads_a: row pass threads per block (0: resident blocks where the sector allows, else 1024);
// strip: strip width 64 or 128 (0: 128 where the row stride ldx is even, else 64 -- the 16-byte
// strip reads rows of odd stride at 8-byte alignment, and at the 15-site (7, 8) sector it measured
// 426 against 403 us for 64-column strips); ncu: compute units of the device
HvPlan hv_plan(int nU, int nD, int threads_a, int strip, int ncu, int64_t ldx) {
    HvPlan p{};
    // the default shape (threads_a 0) takes the resident row pass where its records fill the
    // threads (2048 <= nU <= kResRowsMaxU) and every compute unit gets two rows or more
    p.grid_p = (threads_a <= 0 && nU >= 2048 && nU <= kResRowsMaxU && ncu > 0 && nD >= 2 * ncu) ? ncu : 0;
    p.threads_a = threads_a > 0 ? threads_a : 1024;
    p.grid_a = nD;
    if (strip <= 0) strip = (ldx % 2) == 0 ? 128 : 64;
    p.cpl = (strip >= 128 && nU >= 2) ? 2 : 1;
    // rows per work item: 8 per wave with 128-column strips, 32 with 64-column strips (items of
    // 1-2 K elements; measured best at the 14-site (7, 7) and 15-site (7, 8) sectors)
    p.rows_b = (p.cpl == 2 ? 2 : 8) * (kStripThreads / 64);
    const int64_t w = 64 * p.cpl;
    const int64_t nstrip = (nU + w - 1) / w;
    const int64_t bps = (nD + p.rows_b - 1) / p.rows_b;
    p.items_per_xcd = (int)((nstrip * bps + kXcd - 1) / kXcd);
    p.grid_b = p.items_per_xcd * kXcd;
    return p;
}

template <int NORB, int NW, int NDP>
static hipError_t rows_launch_nd(const HvArgs& a, const HvPlan& pl, hipStream_t s) {
    // where 4 columns' records and the row's diagonal terms fit the 128 registers of a 1024-thread
    // block without spilling (larger records or more orbitals take the per-row blocks)
    if constexpr (NW + 3 * NORB <= 15) {
        if (pl.grid_p > 0 && a.up.rec8 && a.up.count <= kResRowsMaxU) {
            hipLaunchKernelGGL((hv_rows_persist_kernel<NORB, NW, NDP>), dim3(pl.grid_p), dim3(1024), 0, s, a);
            return hipGetLastError();
        }
    }
    if (a.up.rec8 && a.up.count <= kFastMaxU) {
        hipLaunchKernelGGL((hv_rows_kernel<NORB, NW, NDP, true>), dim3(pl.grid_a), dim3(pl.threads_a), 0, s, a);
        return hipGetLastError();
    }
    const size_t lds = (size_t)row_len(a.up.count) * sizeof(double);
    auto kern = hv_rows_kernel<NORB, NW, NDP, false>;
    if (lds > 64 * 1024) {
        hipError_t e = hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(kern, dim3(pl.grid_a), dim3(pl.threads_a), lds, s, a);
    return hipGetLastError();
}

// down-site energies: 8 adds for up to 8 down electrons, else kMaxSite
template <int NORB, int NW>
static hipError_t rows_launch_t(const HvArgs& a, const HvPlan& pl, hipStream_t s) {
    if (a.dn.nel < 0 || a.dn.nel > kMaxSite) return hipErrorInvalidValue;
    return a.dn.nel <= 8 ? rows_launch_nd<NORB, NW, 8>(a, pl, s) : rows_launch_nd<NORB, NW, kMaxSite>(a, pl, s);
}

template <int NORB>
static hipError_t rows_launch(const HvArgs& a, const HvPlan& pl, hipStream_t s) {
    switch (record_dwords(a.up.nslot)) {
        case 3:
            return rows_launch_t<NORB, 3>(a, pl, s);
        case 5:
            return rows_launch_t<NORB, 5>(a, pl, s);
        case 7:
            return rows_launch_t<NORB, 7>(a, pl, s);
        case 9:
            return rows_launch_t<NORB, 9>(a, pl, s);
        case 12:
            return rows_launch_t<NORB, 12>(a, pl, s);
        case 16:
            return rows_launch_t<NORB, 16>(a, pl, s);
        case 24:
            return rows_launch_t<NORB, 24>(a, pl, s);
        case 32:
            return rows_launch_t<NORB, 32>(a, pl, s);
        default:
            return hipErrorInvalidValue;
    }
}

template <int NORB>
static hipError_t hv_launch_norb(const HvArgs& a0, const HvPlan& pl, hipStream_t s) {
    if (a0.up.count == 0 || a0.dn.count == 0) return hipSuccess;
    HvArgs a = a0;
    a.rows_b = pl.rows_b;
    a.items_per_xcd = pl.items_per_xcd;
    hipError_t e = rows_launch<NORB>(a, pl, s);
    if (e != hipSuccess) return e;
    if (((reinterpret_cast<uintptr_t>(a.x) | reinterpret_cast<uintptr_t>(a.y)) & 7) != 0) return hipErrorInvalidValue;
    if (pl.cpl == 2)
        hipLaunchKernelGGL((hv_strip_kernel<NORB, 2>), dim3(pl.grid_b), dim3(kStripThreads), 0, s, a);
    else
        hipLaunchKernelGGL((hv_strip_kernel<NORB, 1>), dim3(pl.grid_b), dim3(kStripThreads), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_hv(const HvArgs& a, const HvPlan& pl, int norb, hipStream_t s) {
    switch (norb) {
        case 1:
            return hv_launch_norb<1>(a, pl, s);
        case 2:
            return hv_launch_norb<2>(a, pl, s);
        case 3:
            return hv_launch_norb<3>(a, pl, s);
        case 4:
            return hv_launch_norb<4>(a, pl, s);
        default:
            return hipErrorInvalidValue;
    }
}

}  // namespace moed
