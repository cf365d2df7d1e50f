// The single-pass H v of one row of a small sector (ed_hamiltonian.F90:88-191), shared by the
// batched engine's one-block-per-row kernel and the persistent eigensolver.
#pragma once
#include <hip/hip_runtime.h>

#include "moed_device.h"
#include "moed_hvops.h"
#include "moed_internal.h"

namespace moed {

// Row id of a small sector, single pass: thread t of nthr stages the row and its negation in xs
// (the layout of row_neg / row_len: 0.0 at nU, -x[k] at row_neg(nU) + k) and thread 0 the energies of the row's
// occupied down sites in ekd.  The caller synchronises before hv_row_compute.
__device__ __forceinline__ void hv_row_stage(const BatchProbDev& d, int id, const double* __restrict__ X, double* xs,
                                             double* ekd, const double* ek, int t, int nthr) {
    const int nU = d.up.count, neg = row_neg(nU);
    const double* __restrict__ xr = X + (size_t)id * nU;
    for (int k = t; k < nU; k += nthr) {
        const double v = xr[k];
        xs[k] = v;
        xs[neg + k] = -v;
    }
    if (t == 0) {
        xs[nU] = 0.0;
        xs[neg + nU] = 0.0;
        const cptr<double> ekc = as_const(ek);
        uint32_t m = scalar_u16(d.dn.str, id);
        for (int q = 0; m; m &= m - 1, ++q) ekd[q] = ekc[__builtin_ctz(m)];
    }
}

// y[id][iu] for the columns iu = t, t + nthr, ... of row id: the diagonal (diag_element, the
// reference's order), the up slots from the staged row, the down moves, spin flip and pair
// exchange -- the term order of the two-pass product, so both give bit-identical y; y *= inv
// when scaled.  id must be uniform over each wave.
template <int NORB, int NW>
__device__ __forceinline__ void hv_row_compute(const BatchProbDev& d, int id, const double* __restrict__ X,
                                               double* __restrict__ Y, const double* xs, const double* ekd,
                                               const double* slotcp, double U, double UmJ3, double Um2J, double J,
                                               bool scaled, double inv, int t, int nthr) {
    constexpr int KMAX = 2 * NW - 5;
    constexpr int NP = NORB * NORB;
    const int nU = d.up.count, nD = d.dn.count;
    const uint32_t sd = scalar_u16(d.dn.str, id);
    const int nd = __builtin_popcount(sd);
    const cptr<double> slotc = as_const(slotcp);
    const uint32_t* __restrict__ recw = reinterpret_cast<const uint32_t*>(d.up.rec);
    const int rstride = d.up.rw;
    const cptr<uint32_t> mvp = as_const(d.dn.mv);
    const cptr<double> mvcp = as_const(d.dn.mvc);
    const int nh = as_const(d.dn.nhop)[id];
    const size_t m0 = (size_t)id * d.dn.mh;
    int32_t msf[NP > 1 ? NP : 1], mpe[NP > 1 ? NP : 1];
    if constexpr (NORB > 1) {
        const cptr<int32_t> dimpc = as_const(d.dn.imp);
#pragma unroll
        for (int q = 0; q < NP; ++q) {
            const int i = q / NORB, j = q % NORB;
            msf[q] = i != j ? dimpc[(size_t)(j * NORB + i) * nD + id] : -1;
            mpe[q] = i != j ? dimpc[(size_t)(i * NORB + j) * nD + id] : -1;
        }
    }
    for (int iu = t; iu < nU; iu += nthr) {
        uint32_t w[NW];
#pragma unroll
        for (int q = 0; q < NW; ++q) w[q] = recw[(size_t)iu * rstride + q];
        const double Sup = __hiloint2double((int)w[1], (int)w[0]);
        double acc = __dmul_rn(diag_element<NORB>(Sup, w[2] & 0xffffu, sd, ekd, nd, U, UmJ3, Um2J), xs[iu]);
#pragma unroll
        for (int q = 0; q < KMAX; ++q) acc = fma(slotc[q], xs[rec_u16(w, 5 + q)], acc);
        for (int k = 0; k < nh; ++k) acc = fma(mvcp[m0 + k], X[(size_t)mvp[m0 + k] * nU + iu], acc);
        if constexpr (NORB > 1) {
#pragma unroll
            for (int q = 0; q < NP; ++q) {
                if (q / NORB == q % NORB) continue;
                const int32_t mu = d.up.imp[(size_t)q * nU + iu];
                if (mu >= 0) {
                    const int uu = mu & 0xffff;
                    const double cu = (mu >> 16) ? -J : J;
                    if (msf[q] >= 0) acc = fma((msf[q] >> 16) ? -cu : cu, X[(size_t)(msf[q] & 0xffff) * nU + uu], acc);
                    if (mpe[q] >= 0) acc = fma((mpe[q] >> 16) ? -cu : cu, X[(size_t)(mpe[q] & 0xffff) * nU + uu], acc);
                }
            }
        }
        Y[(size_t)id * nU + iu] = scaled ? acc * inv : acc;
    }
}


}  // namespace moed
