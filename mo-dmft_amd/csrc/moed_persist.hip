// Persistent work-queue eigensolver for the small sectors of a multi-problem solve: one launch runs
// the whole Lanczos iteration of every queued sector (the diag of ed_solver.F90:107-191 for all of
// the sector loop ed_solver.F90:38-60 at once).
//
// Each workgroup pulls the next sector from a global queue (largest first, one atomic add per
// sector) and runs that sector's iteration to the end before it pulls another, so sectors of very
// different sizes share the GPU without any host round trip: the launch ends when the queue is
// empty, and every workgroup reaches that exit.  Per step, inside the workgroup:
//   w = H V_k / |V_k|        four rows at a time, 256 threads per row, the single-pass row
//                            arithmetic of the batched H v (bit-identical products)
//   three-term projection    V_{k+1} = w - c_{k-1} V_{k-1} - c_k V_k
//   partial reorthogonalisation
//                            Simon's overlap estimates (the recurrence of eig_finish_kernel), and
//                            where they exceed sqrt(eps) classical Gram-Schmidt against the whole
//                            basis with the DGKS second pass
//   convergence              one wave: the lowest nev eigenvalues of T_m by bisection on Sturm
//                            counts (one lane per eigenvalue), the bottom component of each Ritz
//                            vector by two steps of inverse iteration on the LDL^T factors of
//                            T_m - theta I, and the test |beta_m z_{m,i}| <= tol |theta_i| --
//                            after steps m < 20 and every fifth step after, as the host test
// The alpha, beta, norms and basis land in the engine's tables, so the host forms the Ritz pairs
// exactly as for the step-by-step engine.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "moed_device.h"
#include "moed_hvops.h"
#include "moed_hvrow.h"
#include "moed_internal.h"
#include "moed_ritzconv.h"

namespace moed {
namespace {

constexpr int kPT = kPersistThreads;  // threads per workgroup
constexpr int kPQ = kPT / 256;        // rows staged at once (256 threads per row)
constexpr int kPW = kPT / 64;         // waves

__device__ __forceinline__ double start_value(uint64_t seed, int64_t i) {  // fill_start_kernel's hash
    uint64_t z = (seed ^ (uint64_t)i) + 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z = z ^ (z >> 31);
    const double u = (double)(z >> 11) * (1.0 / 9007199254740992.0);
    return 2.0 * u - 1.0;
}

// out[g] = sum over the workgroup of v[g], fixed order (wave trees, then waves in index order);
// every thread returns with out[] valid
template <int G>
__device__ __forceinline__ void wg_sums(double (&v)[G], double (*red)[8], double* out) {
#pragma unroll
    for (int g = 0; g < G; ++g) v[g] = wave_sum(v[g]);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) {
#pragma unroll
        for (int g = 0; g < G; ++g) red[wid][g] = v[g];
    }
    __syncthreads();
    if (threadIdx.x < G) {
        double s = 0.0;
        for (int w = 0; w < kPW; ++w) s += red[w][threadIdx.x];
        out[threadIdx.x] = s;
    }
    __syncthreads();
}

struct PersistShared {
    double cvec[kPersistMaxBasis];  // projection coefficients of the current pass
    const double* vptr[kPersistMaxBasis];  // addresses of the problem's basis vectors
    double sums[16];
    double scal[8];  // [0] 1/|V_k|, [1] alpha accumulator, [2] |w'|^2 of pass 0, [3] |w'|^2 now
    int32_t flag[8]; // [0] done, [1] full pass, [2] force, [3] pass 2, [4] converged
    double red[kPW][8];
    double ekd[kPQ][kMaxSite];
};

}  // namespace

template <int NORB, int NW>
__global__ void __launch_bounds__(kPT) eig_persist_kernel(PersistArgs A) {
    extern __shared__ double dyn[];  // kPQ staged rows of row_len(max_nU) doubles, then alpha, beta^2, beta
    __shared__ PersistShared S;
    __shared__ int s_next;
    const EigArgs& e = A.e;
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const int quarter = uniform_i32(tid >> 8), t256 = tid & 255;
    const int rowlen = row_len(A.max_nU);
    double* xs = dyn + (size_t)quarter * rowlen;
    double* ta = dyn + (size_t)kPQ * rowlen;  // T_m for the convergence test
    double* tb2 = ta + e.ldh;
    double* tb = tb2 + e.ldh;
    const double eps = 0x1p-52;
    const int ldh = e.ldh;

    while (true) {
        if (tid == 0) s_next = atomicAdd(A.head, 1);
        __syncthreads();
        const int q = uniform_i32(s_next);
        __syncthreads();
        if (q >= A.nq) return;
        const int p = uniform_i32(as_const(A.order)[q]);
        BatchProbDev d;  // scalar loads of the descriptor
        {
            const cptr<uint64_t> src = (cptr<uint64_t>)(const __attribute__((address_space(4))) void*)(A.prob + p);
            uint64_t wds[sizeof(BatchProbDev) / 8];
#pragma unroll
            for (int k = 0; k < (int)(sizeof(BatchProbDev) / 8); ++k) wds[k] = src[k];
            __builtin_memcpy(&d, wds, sizeof(d));
        }
        const int64_t off = d.off;
        const int64_t D = as_const(e.pdim)[p];
        const int M = as_const(e.mmax)[p];
        const int nD = d.dn.count;
        double* alpha = e.alpha + (size_t)p * ldh;
        double* beta = e.beta + (size_t)p * ldh;
        double* nrm = e.nrm + (size_t)p * ldh;
        double* om = e.om + (size_t)p * 3 * ldh;
        double* const* V = A.V;
        double* W = A.W + off;

        // ---- start vector and its norm
        {
            double* V0 = V[0] + off;
            const uint64_t seed = as_const(A.seed)[p];
            double s[1] = {0.0};
            for (int64_t i = tid; i < D; i += kPT) {
                const double v = start_value(seed, i);
                V0[i] = v;
                s[0] = fma(v, v, s[0]);
            }
            wg_sums<1>(s, S.red, S.sums);
            if (tid == 0) {
                nrm[0] = sqrt(S.sums[0]);
                S.flag[0] = 0;
                S.flag[2] = 0;
                S.flag[4] = 0;
            }
            __syncthreads();
        }
        int mfinal = M;
        for (int k = 0; k < M; ++k) {
            const double* Vk = V[k] + off;
            double* Vn = V[k + 1] + off;
            // ---- w = H V_k / |V_k|
            const double inv = 1.0 / nrm[k];
            for (int id0 = 0; id0 < nD; id0 += kPQ) {
                const int id = id0 + quarter;
                if (id < nD) hv_row_stage(d, id, Vk, xs, S.ekd[quarter], A.ek, t256, 256);
                __syncthreads();
                if (id < nD)
                    hv_row_compute<NORB, NW>(d, id, Vk, W, xs, S.ekd[quarter], A.slotc, A.U, A.UmJ3, A.Um2J, A.J,
                                             true, inv, t256, 256);
                __syncthreads();
            }
            // ---- pass 0: the three-term part against V_{k-1}, V_k
            const int i0 = k > 0 ? k - 1 : 0;
            {
                const double* Va = V[i0] + off;
                double s[2] = {0.0, 0.0};
                for (int64_t i = tid; i < D; i += kPT) {
                    const double wv = W[i];
                    s[0] = fma(Va[i], wv, s[0]);
                    s[1] = fma(Vk[i], wv, s[1]);
                }
                wg_sums<2>(s, S.red, S.sums);
                const double ca = k > 0 ? S.sums[0] / (nrm[i0] * nrm[i0]) : 0.0;
                const double ck = S.sums[1] / (nrm[k] * nrm[k]);
                double s2[1] = {0.0};
                for (int64_t i = tid; i < D; i += kPT) {
                    double v = W[i];
                    if (k > 0) v = fma(-ca, Va[i], v);
                    v = fma(-ck, Vk[i], v);
                    Vn[i] = v;
                    s2[0] = fma(v, v, s2[0]);
                }
                wg_sums<1>(s2, S.red, S.sums + 2);
                if (tid == 0) {
                    S.scal[1] = S.sums[1] / nrm[k];  // alpha_k so far
                    S.scal[2] = S.sums[2];           // |w'|^2 after pass 0
                    S.scal[3] = S.sums[2];
                }
                __syncthreads();
            }
            // ---- overlap estimates omega_{k+1,i} (eig_finish_kernel, pass 0); wave 0
            if (wid == 0) {
                double* wn = om + (size_t)((k + 1) % 3) * ldh;
                const double* wc = om + (size_t)(k % 3) * ldh;
                const double* wp = om + (size_t)((k + 2) % 3) * ldh;
                const double ak = S.scal[1], bk = sqrt(S.scal[2]);
                double tn = 0.0;
                for (int i = lane; i < k; i += 64) tn = fmax(tn, fabs(alpha[i]) + beta[i] + (i > 0 ? beta[i - 1] : 0.0));
                for (int off2 = 32; off2 > 0; off2 >>= 1) tn = fmax(tn, __shfl_xor(tn, off2, 64));
                tn = fmax(tn, fabs(ak) + bk + (k > 0 ? beta[k - 1] : 0.0));
                const double rt = eps * tn;
                double mx = 0.0;
                if (bk > 0.0) {
                    for (int i = lane; i < k; i += 64) {
                        double t = beta[i] * wc[i + 1] + (alpha[i] - ak) * wc[i] - beta[k - 1] * wp[i];
                        if (i > 0) t += beta[i - 1] * wc[i - 1];
                        t = (t + copysign(rt, t)) / bk;
                        wn[i] = t;
                        mx = fmax(mx, fabs(t));
                    }
                }
                for (int off2 = 32; off2 > 0; off2 >>= 1) mx = fmax(mx, __shfl_xor(mx, off2, 64));
                if (lane == 0) {
                    if (k == 0) om[0] = 1.0;
                    wn[k] = bk > 0.0 ? rt / bk : 0.0;
                    wn[k + 1] = 1.0;
                    const bool need = bk > 0.0 && (mx > sqrt(eps) || S.flag[2]);
                    S.flag[1] = need;
                    if (need) S.flag[2] = S.flag[2] ? 0 : 1;
                }
            }
            __syncthreads();
            // ---- passes 1 and 2: the whole basis (classical Gram-Schmidt, DGKS second pass)
            if (S.flag[1]) {
                // the addresses of V_0..V_k, staged once
                for (int j = tid; j <= k; j += kPT) S.vptr[j] = V[j] + off;
                __syncthreads();
                for (int pass = 1; pass <= 2; ++pass) {
                    const int k1 = k + 1;
                    for (int g0 = 0; g0 < k1; g0 += 8) {
                        const double* vg[8];
#pragma unroll
                        for (int g = 0; g < 8; ++g) vg[g] = S.vptr[g0 + g < k1 ? g0 + g : g0];
                        double s[8];
#pragma unroll
                        for (int g = 0; g < 8; ++g) s[g] = 0.0;
                        for (int64_t i = tid; i < D; i += kPT) {
                            const double wv = Vn[i];
                            double u[8];
#pragma unroll
                            for (int g = 0; g < 8; ++g) u[g] = vg[g][i];
#pragma unroll
                            for (int g = 0; g < 8; ++g)
                                if (g0 + g < k1) s[g] = fma(u[g], wv, s[g]);
                        }
                        wg_sums<8>(s, S.red, S.sums);
                        if (tid < 8 && g0 + tid < k1) {
                            const int i = g0 + tid;
                            S.cvec[i] = S.sums[tid] / (nrm[i] * nrm[i]);
                            if (i == k) S.scal[1] += S.sums[tid] / nrm[i];
                        }
                        __syncthreads();
                    }
                    double s2[1] = {0.0};
                    for (int64_t i = tid; i < D; i += kPT) {
                        double v = Vn[i];
                        int j = 0;
                        for (; j + 8 <= k1; j += 8) {  // eight vectors' loads in flight
                            double u[8];
#pragma unroll
                            for (int g = 0; g < 8; ++g) u[g] = S.vptr[j + g][i];
#pragma unroll
                            for (int g = 0; g < 8; ++g) v = fma(-S.cvec[j + g], u[g], v);
                        }
                        for (; j < k1; ++j) v = fma(-S.cvec[j], S.vptr[j][i], v);
                        Vn[i] = v;
                        s2[0] = fma(v, v, s2[0]);
                    }
                    wg_sums<1>(s2, S.red, S.sums);
                    const double wp2 = S.sums[0];
                    __syncthreads();
                    // second pass where w' lost more than half its norm
                    const bool again = pass == 1 && wp2 < 0.5 * S.scal[2];
                    if (tid == 0) {
                        S.scal[3] = wp2;
                        if (again) S.scal[2] = wp2;
                    }
                    __syncthreads();
                    if (!again) break;
                }
                if (wid == 0) {  // V_{k+1} is now orthogonal to the basis to rounding
                    double* wn = om + (size_t)((k + 1) % 3) * ldh;
                    for (int i = lane; i < k; i += 64) wn[i] = eps;
                }
            }
            // ---- close the step: alpha_k, beta_k, exhaustion or cap (eig_close_step)
            if (tid == 0) {
                const double wp2 = S.scal[3];
                alpha[k] = S.scal[1];
                const double bk = sqrt(wp2);
                double tnorm = 0.0;
                for (int i = 0; i <= k; ++i) tnorm = fmax(tnorm, fabs(alpha[i]) + (i < k ? beta[i] : 0.0));
                beta[k] = bk;
                if (bk <= 1e-14 * fmax(tnorm, 1e-300) || k + 1 >= M)
                    S.flag[0] = 1;
                else
                    nrm[k + 1] = bk;
            }
            __syncthreads();
            const int m = k + 1;
            if (S.flag[0]) {
                mfinal = m;
                break;
            }
            // ---- convergence of the lowest nev Ritz pairs (wave 0)
            if (m >= A.nev && (m < 20 || m % 5 == 0)) {
                if (wid == 0) {
                    const bool all = ritz_converged(alpha, beta, m, A.nev, A.tol, ta, tb, tb2,
                                                    A.zs + (size_t)p * A.nev * 3 * ldh, ldh);
                    if (lane == 0 && all) S.flag[0] = 1;
                }
                __syncthreads();
                if (S.flag[0]) {
                    mfinal = m;
                    break;
                }
            }
        }
        if (tid == 0) {
            e.mfinal[p] = mfinal;
            e.done[p] = 1;
        }
        __syncthreads();
    }
}

template <int NORB, int NW>
static hipError_t persist_t(const PersistArgs& a, int grid, hipStream_t s) {
    const size_t lds = ((size_t)kPQ * row_len(a.max_nU) + 3 * (size_t)a.e.ldh) * sizeof(double);
    auto kern = eig_persist_kernel<NORB, NW>;
    if (lds > 64 * 1024) {
        hipError_t err = hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (err != hipSuccess) return err;
    }
    hipLaunchKernelGGL(kern, dim3(grid), dim3(kPT), lds, s, a);
    return hipGetLastError();
}

template <int NORB>
static hipError_t persist_nw(const PersistArgs& a, int grid, int nslot, hipStream_t s) {
    switch (record_dwords(nslot)) {
        case 3:
            return persist_t<NORB, 3>(a, grid, s);
        case 5:
            return persist_t<NORB, 5>(a, grid, s);
        case 7:
            return persist_t<NORB, 7>(a, grid, s);
        case 9:
            return persist_t<NORB, 9>(a, grid, s);
        case 12:
            return persist_t<NORB, 12>(a, grid, s);
        case 16:
            return persist_t<NORB, 16>(a, grid, s);
        case 24:
            return persist_t<NORB, 24>(a, grid, s);
        case 32:
            return persist_t<NORB, 32>(a, grid, s);
        default:
            return hipErrorInvalidValue;
    }
}

hipError_t launch_eig_persistent(const PersistArgs& a, int grid, int norb, int nslot, hipStream_t s) {
    if (a.nq <= 0) return hipSuccess;
    if (a.max_nU > kPersistMaxU || a.e.ldh > kPersistMaxBasis) return hipErrorInvalidValue;
    switch (norb) {
        case 1:
            return persist_nw<1>(a, grid, nslot, s);
        case 2:
            return persist_nw<2>(a, grid, nslot, s);
        case 3:
            return persist_nw<3>(a, grid, nslot, s);
        case 4:
            return persist_nw<4>(a, grid, nslot, s);
        default:
            return hipErrorInvalidValue;
    }
}

}  // namespace moed
