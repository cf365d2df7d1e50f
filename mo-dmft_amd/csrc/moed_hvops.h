// Per-element pieces of the sector H v shared by the two-pass product (moed_hv.hip) and the
// batched single-pass product (moed_eig.hip): both sum every element's terms in the same order.
#pragma once
#include <hip/hip_runtime.h>

#include "moed_internal.h"

namespace moed {

// diagonal element of (up string su with onsite prefix Sup, down string sd), ed_hamiltonian.F90:110-139:
// Sup, then ek of each occupied down site in ascending order (ekd[0..nd), in LDS), then the
// interaction terms (ed_operators.F90:156-162) orbital by orbital and pair by pair.  fma(n, V, S)
// with n in {0, 1} equals S + (n ? V : 0.0) exactly, and adding +0.0 for a term the state does
// not have leaves the sum bit-identical (it starts at +0.0 and never becomes -0.0).
template <int NORB>
__device__ __forceinline__ double diag_element(double Sup, uint32_t su, uint32_t sd, const double* ekd, int nd,
                                               double U, double UmJ3, double Um2J) {
    double S = Sup;
    for (int q = 0; q < nd; ++q) S += ekd[q];
    double n[NORB];
#pragma unroll
    for (int o = 0; o < NORB; ++o) n[o] = (double)((su >> o) & 1u);
#pragma unroll
    for (int o = 0; o < NORB; ++o)
        if ((sd >> o) & 1u) S = fma(n[o], U, S);
#pragma unroll
    for (int o = 0; o < NORB; ++o)
#pragma unroll
        for (int p = o + 1; p < NORB; ++p) {
            S = fma(n[o] * n[p], UmJ3, S);
            if ((sd >> p) & 1u) S = fma(n[o], Um2J, S);
            if ((sd >> o) & 1u) S = fma(n[p], Um2J, S);
            if ((sd >> o) & (sd >> p) & 1u) S += UmJ3;
        }
    return S;
}

// The interaction part of diag_element with the down string's conditions turned into per-row
// coefficients: a term whose down orbital is empty is added with coefficient 0.0 instead of being
// skipped.  fma(n, 0.0, S) and S + 0.0 equal S exactly (S is never -0.0), so the result is
// bit-identical to diag_element's, with no per-element selects.
template <int NORB>
struct RowDiag {
    double Uo[NORB];         // U where down orbital o is occupied, else 0.0
    double A[NORB];          // U - 2J where down orbital o is occupied, else 0.0
    double B[NORB][NORB];    // U - 3J where down orbitals o and p are both occupied, else 0.0
};

template <int NORB>
__device__ __forceinline__ RowDiag<NORB> row_diag(uint32_t sd, double U, double UmJ3, double Um2J) {
    RowDiag<NORB> r;
#pragma unroll
    for (int o = 0; o < NORB; ++o) {
        const bool d = (sd >> o) & 1u;
        r.Uo[o] = d ? U : 0.0;
        r.A[o] = d ? Um2J : 0.0;
#pragma unroll
        for (int p = 0; p < NORB; ++p) r.B[o][p] = (d && ((sd >> p) & 1u)) ? UmJ3 : 0.0;
    }
    return r;
}

// diag_element's interaction terms, in its order, on top of S = onsite sum of the up and down strings
template <int NORB>
__device__ __forceinline__ double diag_interaction(double S, uint32_t su, const RowDiag<NORB>& r, double UmJ3) {
    double n[NORB];
#pragma unroll
    for (int o = 0; o < NORB; ++o) n[o] = (double)((su >> o) & 1u);
#pragma unroll
    for (int o = 0; o < NORB; ++o) S = fma(n[o], r.Uo[o], S);
#pragma unroll
    for (int o = 0; o < NORB; ++o)
#pragma unroll
        for (int p = o + 1; p < NORB; ++p) {
            S = fma(n[o] * n[p], UmJ3, S);
            S = fma(n[o], r.A[p], S);
            S = fma(n[p], r.A[o], S);
            S += r.B[o][p];
        }
    return S;
}

// uint16 number i of a record held in registers
template <int NW>
__device__ __forceinline__ uint32_t rec_u16(const uint32_t (&w)[NW], int i) {
    return (i & 1) ? (w[i >> 1] >> 16) : (w[i >> 1] & 0xffffu);
}

}  // namespace moed
