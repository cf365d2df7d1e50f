// Multi-problem engine: the kernels that let many independent problems (all the sectors of a
// model, or many Lanczos runs) advance with one launch per operation.
//
//   hv_batch_kernel      single-pass sector H v, one block per (problem, row); for the small
//                        sectors of a batch the whole vector of a problem sits in L2, so the
//                        down-move rows are read directly.  Term order per element is that of
//                        the two-pass product (diagonal, up slots, down moves, spin flip /
//                        pair exchange), so both give bit-identical y.
//   eig_* kernels        Lanczos with partial reorthogonalisation over packed vectors: the
//                        three-term projection every step, the whole basis (classical
//                        Gram-Schmidt, DGKS second pass) where the estimated loss of
//                        orthogonality asks for it; per-item partial inner products,
//                        per-problem reductions in a fixed order, the step's alpha / beta and
//                        the stop test, all on the device.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "moed_device.h"
#include "moed_hvops.h"
#include "moed_hvrow.h"
#include "moed_internal.h"
#include "moed_ritzconv.h"

namespace moed {

// ------------------------------------------------------------------ batched H v

template <int NORB, int NW>
__global__ void __launch_bounds__(kEigThreads) hv_batch_kernel(HvBatchArgs A) {
    extern __shared__ double xs[];
    __shared__ double ekd[kMaxSite];

    const int item = blockIdx.x;
    const int p = as_const(A.row_prob)[item];
    if (A.skip && as_const(A.skip)[p]) return;
    const int id = as_const(A.row_id)[item];
    BatchProbDev d;  // scalar loads of the descriptor
    {
        static_assert(sizeof(BatchProbDev) % 8 == 0, "descriptor is read as 8-byte words");
        const cptr<uint64_t> src = (cptr<uint64_t>)(const __attribute__((address_space(4))) void*)(A.prob + p);
        uint64_t wds[sizeof(BatchProbDev) / 8];
#pragma unroll
        for (int q = 0; q < (int)(sizeof(BatchProbDev) / 8); ++q) wds[q] = src[q];
        __builtin_memcpy(&d, wds, sizeof(d));
    }
    const double* __restrict__ X = A.x + d.off;
    double* __restrict__ Y = A.y + d.off;
    hv_row_stage(d, id, X, xs, ekd, A.ek, threadIdx.x, kEigThreads);
    __syncthreads();
    const double inv = A.scale ? 1.0 / A.scale[(size_t)p * A.scale_stride] : 1.0;
    hv_row_compute<NORB, NW>(d, id, X, Y, xs, ekd, A.slotc, A.U, A.UmJ3, A.Um2J, A.J, A.scale != nullptr, inv,
                             threadIdx.x, kEigThreads);
}

template <int NORB, int NW>
static hipError_t hv_batch_t(const HvBatchArgs& a, int nitems, int max_nU, hipStream_t s) {
    const size_t lds = (size_t)row_len(max_nU) * sizeof(double);
    auto kern = hv_batch_kernel<NORB, NW>;
    if (lds > 64 * 1024) {
        hipError_t e = hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(kern, dim3(nitems), dim3(kEigThreads), lds, s, a);
    return hipGetLastError();
}

template <int NORB>
static hipError_t hv_batch_nw(const HvBatchArgs& a, int nitems, int max_nU, int nslot, hipStream_t s) {
    switch (record_dwords(nslot)) {
        case 3:
            return hv_batch_t<NORB, 3>(a, nitems, max_nU, s);
        case 5:
            return hv_batch_t<NORB, 5>(a, nitems, max_nU, s);
        case 7:
            return hv_batch_t<NORB, 7>(a, nitems, max_nU, s);
        case 9:
            return hv_batch_t<NORB, 9>(a, nitems, max_nU, s);
        case 12:
            return hv_batch_t<NORB, 12>(a, nitems, max_nU, s);
        case 16:
            return hv_batch_t<NORB, 16>(a, nitems, max_nU, s);
        case 24:
            return hv_batch_t<NORB, 24>(a, nitems, max_nU, s);
        case 32:
            return hv_batch_t<NORB, 32>(a, nitems, max_nU, s);
        default:
            return hipErrorInvalidValue;
    }
}

hipError_t launch_hv_batch(const HvBatchArgs& a, int nitems, int max_nU, int norb, int nslot, hipStream_t s) {
    if (nitems <= 0) return hipSuccess;
    switch (norb) {
        case 1:
            return hv_batch_nw<1>(a, nitems, max_nU, nslot, s);
        case 2:
            return hv_batch_nw<2>(a, nitems, max_nU, nslot, s);
        case 3:
            return hv_batch_nw<3>(a, nitems, max_nU, nslot, s);
        case 4:
            return hv_batch_nw<4>(a, nitems, max_nU, nslot, s);
        default:
            return hipErrorInvalidValue;
    }
}

// ------------------------------------------------------------------ segmented vector kernels

__device__ __forceinline__ bool eig_active(const EigArgs& e, int p, int pass) {
    if (e.done[p]) return false;
    if (pass == 1) return e.full[p] != 0;
    return pass != 2 || e.pass2[p] != 0;
}

// sum of part[row * ni + i0 + it] over it < n by one wave: lane l adds items l, l + 64, ... in order,
// then the lanes are combined in a fixed tree (same result for every launch)
__device__ __forceinline__ double wave_item_sum(const double* part, size_t row0, int n) {
    const int lane = threadIdx.x & 63;
    double s = 0.0;
    for (int it = lane; it < n; it += 64) s += part[row0 + it];
    return wave_sum(s);
}

// part[item] = sum of a*b over the item (pass = 0: every problem)
__global__ void __launch_bounds__(kEigThreads) eig_segdot_kernel(EigArgs e, const double* a, const double* b) {
    __shared__ double red[kEigThreads / 64];
    const int item = blockIdx.x;
    const int64_t base = e.item_start[item];
    const int len = e.item_len[item];
    double s = 0.0;
#pragma unroll
    for (int j = 0; j < kEigChunk / kEigThreads; ++j) {
        const int r = threadIdx.x + j * kEigThreads;
        if (r < len) s = fma(a[base + r], b[base + r], s);
    }
    s = block_sum<kEigThreads>(s, red);
    if (threadIdx.x == 0) e.part[item] = s;
}

// out[p * stride] = sum of part over p's items (op 1: square root); one wave per problem
__global__ void __launch_bounds__(64) eig_segsum_kernel(EigArgs e, double* out, int stride, int op) {
    const int p = blockIdx.x;
    const double s = wave_item_sum(e.part, e.pitem0[p], e.pnitem[p]);
    if (threadIdx.x == 0) out[(size_t)p * stride] = op == 1 ? sqrt(s) : s;
}

hipError_t launch_eig_segdot(const EigArgs& e, const double* a, const double* b, double* out, int stride, int op,
                             hipStream_t s) {
    hipLaunchKernelGGL(eig_segdot_kernel, dim3(e.ni), dim3(kEigThreads), 0, s, e, a, b);
    hipLaunchKernelGGL(eig_segsum_kernel, dim3(e.P), dim3(64), 0, s, e, out, stride, op);
    return hipGetLastError();
}

// part[i * ni + item] = <V_i, w> over the item for the VG basis vectors of group blockIdx.y
// (i0 <= i < k1); V[i] is the address of basis vector i
constexpr int kEigVG = 8;
__global__ void __launch_bounds__(kEigThreads) eig_multidot_kernel(EigArgs e, const double* const* V, const double* w,
                                                                   int i0, int k1, int pass) {
    __shared__ double red[kEigVG][kEigThreads / 64];
    const int item = blockIdx.x;
    const int p = e.item_prob[item];
    if (!eig_active(e, p, pass)) return;
    const int64_t base = e.item_start[item];
    const int len = e.item_len[item];
    constexpr int R = kEigChunk / kEigThreads;
    double wv[R];
#pragma unroll
    for (int j = 0; j < R; ++j) {
        const int r = threadIdx.x + j * kEigThreads;
        wv[j] = r < len ? w[base + r] : 0.0;
    }
    const int g0 = i0 + blockIdx.y * kEigVG;
    double acc[kEigVG];
#pragma unroll
    for (int q = 0; q < kEigVG; ++q) {
        const int i = g0 + q;
        double s = 0.0;
        if (i < k1) {
            const double* __restrict__ u = V[i] + base;
#pragma unroll
            for (int j = 0; j < R; ++j) {
                const int r = threadIdx.x + j * kEigThreads;
                if (r < len) s = fma(u[r], wv[j], s);
            }
        }
        acc[q] = wave_sum(s);
    }
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) {
#pragma unroll
        for (int q = 0; q < kEigVG; ++q) red[q][wid] = acc[q];
    }
    __syncthreads();
    if (threadIdx.x < kEigVG) {
        const int i = g0 + threadIdx.x;
        if (i < k1) {
            double s = 0.0;
            for (int wv2 = 0; wv2 < kEigThreads / 64; ++wv2) s += red[threadIdx.x][wv2];
            e.part[(size_t)i * e.ni + item] = s;
        }
    }
}

// h_i = sum of the item partials (one wave per vector i, i0 <= i < k1); c_i = h_i / nrm_i^2; the
// wave of vector k adds h_k / nrm_k to the step's alpha
__global__ void __launch_bounds__(kEigThreads) eig_reduce_h_kernel(EigArgs e, int i0, int k1, int pass) {
    const int p = blockIdx.x;
    if (!eig_active(e, p, pass)) return;
    const int i = i0 + blockIdx.y * (kEigThreads / 64) + (threadIdx.x >> 6);
    if (i >= k1) return;
    const double sum = wave_item_sum(e.part, (size_t)i * e.ni + e.pitem0[p], e.pnitem[p]);
    if ((threadIdx.x & 63) != 0) return;
    const double* nrm = e.nrm + (size_t)p * e.ldh;
    double* aux = e.aux + (size_t)p * 4;
    e.h[(size_t)p * e.ldh + i] = sum;
    if (i < k1) e.c[(size_t)p * e.ldh + i] = sum / (nrm[i] * nrm[i]);
    if (i == k1 - 1) aux[0] += sum / nrm[i];
}

// dst = src - sum_{i0 <= i < k1} c_i V_i over the item; part[item] = |dst|^2 over the item
__global__ void __launch_bounds__(kEigThreads) eig_project_kernel(EigArgs e, const double* const* V,
                                                                  const double* src, double* dst, int i0, int k1,
                                                                  int pass) {
    extern __shared__ double sc[];
    __shared__ double red[kEigThreads / 64];
    const int item = blockIdx.x;
    const int p = e.item_prob[item];
    if (!eig_active(e, p, pass)) return;
    for (int i = threadIdx.x; i < k1; i += kEigThreads) sc[i] = e.c[(size_t)p * e.ldh + i];
    __syncthreads();
    const int64_t base = e.item_start[item];
    const int len = e.item_len[item];
    constexpr int R = kEigChunk / kEigThreads;
    double acc[R];
#pragma unroll
    for (int j = 0; j < R; ++j) {
        const int r = threadIdx.x + j * kEigThreads;
        acc[j] = r < len ? src[base + r] : 0.0;
    }
    // (a full item: no bounds tests, eight basis vectors' loads in flight per step)
    if (len == kEigChunk) {
        int i = i0;
        for (; i + 8 <= k1; i += 8) {
            double u[8][R];
#pragma unroll
            for (int q = 0; q < 8; ++q)
#pragma unroll
                for (int j = 0; j < R; ++j) u[q][j] = V[i + q][base + threadIdx.x + j * kEigThreads];
#pragma unroll
            for (int q = 0; q < 8; ++q)
#pragma unroll
                for (int j = 0; j < R; ++j) acc[j] = fma(-sc[i + q], u[q][j], acc[j]);
        }
        for (; i < k1; ++i)
#pragma unroll
            for (int j = 0; j < R; ++j) acc[j] = fma(-sc[i], V[i][base + threadIdx.x + j * kEigThreads], acc[j]);
    } else {
        for (int i = i0; i < k1; ++i) {
            const double ci = sc[i];
            const double* __restrict__ u = V[i] + base;
#pragma unroll
            for (int j = 0; j < R; ++j) {
                const int r = threadIdx.x + j * kEigThreads;
                if (r < len) acc[j] = fma(-ci, u[r], acc[j]);
            }
        }
    }
    double ss = 0.0;
#pragma unroll
    for (int j = 0; j < R; ++j) {
        const int r = threadIdx.x + j * kEigThreads;
        if (r < len) {
            dst[base + r] = acc[j];
            ss = fma(acc[j], acc[j], ss);
        }
    }
    ss = block_sum<kEigThreads>(ss, red);
    if (threadIdx.x == 0) e.part[item] = ss;
}

// ------------------------------------------------------------------ FP64 MFMA contractions
//
// The whole-basis passes of full reorthogonalisation and the Ritz product on the matrix cores,
// v_mfma_f64_16x16x4_f64: D(16x16) += A(16x4) B(4x16) with A[m][k] in lane m + 16k, B[k][n] in
// lane n + 16k and D[i / 16 + 4j][i % 16] in element j of lane i's accumulator.  Against one w the
// Gram-Schmidt contractions are matrix-vector products, so one of the 16 columns carries data; the
// passes stay bound by HBM (8 bytes of basis per multiply-add) with the matrix cores well below
// their rate.  The Ritz product uses nev of the 16 columns.
typedef double mf4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ mf4 mfma16(double a, double b, mf4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// part[i * ni + item] = <V_i, w> over the item for the 16 basis vectors g0 <= i < g0 + 16 (i < k1)
// of group blockIdx.y.  Lane (m, q) = (lane % 16, lane / 16) holds elements 4q..4q+3 of each
// 16-element slice of vector g0 + m (two 16-byte loads); the lanes of column 0 hold the same
// elements of w.  Element j of the four goes to accumulator j: four independent MFMA chains.
__global__ void __launch_bounds__(kEigThreads) eig_multidot_mfma_kernel(EigArgs e, const double* const* V,
                                                                        const double* w, int i0, int k1, int pass) {
    __shared__ double red[kEigThreads / 64][16];
    const int item = blockIdx.x;
    const int p = e.item_prob[item];
    if (!eig_active(e, p, pass)) return;
    const int64_t base = e.item_start[item];
    const int len = e.item_len[item];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const int m = lane & 15, q = lane >> 4;
    const int g0 = i0 + blockIdx.y * 16;
    const bool have = g0 + m < k1;
    const double* __restrict__ u = have ? V[g0 + m] + base : w + base;  // (a valid address either way)
    const double* __restrict__ wb = w + base;
    mf4 acc[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[j] = mf4{0.0, 0.0, 0.0, 0.0};
    constexpr int kPerWave = kEigChunk / (kEigThreads / 64);  // 256 elements per wave
    const int r0 = wid * kPerWave + 4 * q;
    // every item starts on an even element, so 16-byte loads of a whole item are aligned
    if (len == kEigChunk && ((reinterpret_cast<uintptr_t>(u) | reinterpret_cast<uintptr_t>(wb)) & 15) == 0) {
        for (int r = r0; r < r0 + kPerWave; r += 16) {
            const double2 a0 = *reinterpret_cast<const double2*>(u + r);
            const double2 a1 = *reinterpret_cast<const double2*>(u + r + 2);
            const double2 b0 = *reinterpret_cast<const double2*>(wb + r);
            const double2 b1 = *reinterpret_cast<const double2*>(wb + r + 2);
            const double z = 0.0;
            acc[0] = mfma16(have ? a0.x : z, m == 0 ? b0.x : z, acc[0]);
            acc[1] = mfma16(have ? a0.y : z, m == 0 ? b0.y : z, acc[1]);
            acc[2] = mfma16(have ? a1.x : z, m == 0 ? b1.x : z, acc[2]);
            acc[3] = mfma16(have ? a1.y : z, m == 0 ? b1.y : z, acc[3]);
        }
    } else {
        for (int r = r0; r < r0 + kPerWave; r += 16) {
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const bool in = r + j < len;
                acc[j] = mfma16((have && in) ? u[r + j] : 0.0, (m == 0 && in) ? wb[r + j] : 0.0, acc[j]);
            }
        }
    }
    // column 0: lane 16 q holds D[q + 4 j][0] = <V_{g0 + q + 4 j}, w> over this wave's elements
    if (m == 0) {
#pragma unroll
        for (int j = 0; j < 4; ++j) red[wid][q + 4 * j] = (acc[0][j] + acc[1][j]) + (acc[2][j] + acc[3][j]);
    }
    __syncthreads();
    if (threadIdx.x < 16) {
        const int i = g0 + threadIdx.x;
        if (i < k1) {
            double sum = 0.0;
            for (int v = 0; v < kEigThreads / 64; ++v) sum += red[v][threadIdx.x];
            e.part[(size_t)i * e.ni + item] = sum;
        }
    }
}

// dst = src - sum_{i0 <= i < k1} c_i V_i over the item; part[item] = |dst|^2 over the item.
// Per 16-element slice: B[k][n] = V_{i + k}[slice + n] (lanes n + 16 k, 128 contiguous bytes per
// vector), A[0][k] = c_{i + k} (rows 1..15 zero), so lanes 0..15 end with D[0][n] in element 0.
// A wave works on four slices at once (four independent MFMA chains); the basis addresses of the
// pass sit in LDS.
__global__ void __launch_bounds__(kEigThreads) eig_project_mfma_kernel(EigArgs e, const double* const* V,
                                                                       const double* src, double* dst, int i0,
                                                                       int k1, int pass) {
    extern __shared__ double sc[];  // c_i, then the addresses of V_i (i < k1)
    __shared__ double red[kEigThreads / 64];
    const int item = blockIdx.x;
    const int p = e.item_prob[item];
    if (!eig_active(e, p, pass)) return;
    const double** vp = reinterpret_cast<const double**>(sc + k1);
    for (int i = threadIdx.x; i < k1; i += kEigThreads) {
        sc[i] = e.c[(size_t)p * e.ldh + i];
        vp[i] = V[i];
    }
    __syncthreads();
    const int64_t base = e.item_start[item];
    const int len = e.item_len[item];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const int n = lane & 15, k = lane >> 4;
    constexpr int kPerWave = kEigChunk / (kEigThreads / 64);
    double ss = 0.0;
    for (int s0 = wid * kPerWave; s0 < (wid + 1) * kPerWave; s0 += 64) {
        mf4 acc[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) acc[t] = mf4{0.0, 0.0, 0.0, 0.0};
        for (int i = i0; i < k1; i += 4) {
            const int ii = i + k;
            const bool have = ii < k1;
            const double* __restrict__ v = (have ? vp[ii] : src) + base + s0 + n;
            double b[4];
#pragma unroll
            for (int t = 0; t < 4; ++t) b[t] = (have && s0 + 16 * t + n < len) ? v[16 * t] : 0.0;
            const double a = (n == 0 && have) ? sc[ii] : 0.0;
#pragma unroll
            for (int t = 0; t < 4; ++t) acc[t] = mfma16(a, b[t], acc[t]);
        }
        if (lane < 16) {
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                const int r = s0 + 16 * t + n;
                if (r < len) {
                    const double v = src[base + r] - acc[t][0];
                    dst[base + r] = v;
                    ss = fma(v, v, ss);
                }
            }
        }
    }
    ss = block_sum<kEigThreads>(ss, red);
    if (threadIdx.x == 0) e.part[item] = ss;
}

__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
    return v;
}

// the step's alpha (accumulated over its passes) and beta = |V_{k+1}|, the exhaustion / cap test
__device__ void eig_close_step(const EigArgs& e, int p, int k, double wp2, double tiny) {
    double* aux = e.aux + (size_t)p * 4;
    double* alpha = e.alpha + (size_t)p * e.ldh;
    double* beta = e.beta + (size_t)p * e.ldh;
    alpha[k] = aux[0];
    aux[0] = 0.0;
    const double bk = sqrt(wp2);
    double tnorm = 0.0;
    for (int i = 0; i <= k; ++i) tnorm = fmax(tnorm, fabs(alpha[i]) + (i < k ? beta[i] : 0.0));
    beta[k] = bk;
    if (bk <= tiny * fmax(tnorm, 1e-300) || k + 1 >= e.mmax[p]) {
        e.done[p] = 1;
        e.mfinal[p] = k + 1;
    } else {
        e.nrm[(size_t)p * e.ldh + k + 1] = bk;
    }
}

// |w'|^2 from the item partials.
// Pass 0 (the three-term part): partial reorthogonalisation (H. D. Simon, "The Lanczos algorithm
// with partial reorthogonalization", Math. Comp. 42 (1984) 115-142).  With q_i = V_i / |V_i|, the
// overlaps omega_{k+1,i} ~ <q_{k+1}, q_i> follow from the recurrence's alpha, beta:
//   beta_k omega_{k+1,i} = beta_i omega_{k,i+1} + (alpha_i - alpha_k) omega_{k,i}
//                          + beta_{i-1} omega_{k,i-1} - beta_{k-1} omega_{k-1,i} (+- eps |T|),
// omega_{k+1,k} = eps |T| / beta_k, omega_{k+1,k+1} = 1.  When max_i<k |omega_{k+1,i}| exceeds
// sqrt(eps), V_{k+1} is projected against the whole basis (pass 1, then pass 2 where the DGKS test
// asks), and so is V_{k+2} at the next step; the estimates of a vector projected against the whole
// basis drop to eps.  Otherwise the step closes here, and the basis stays semi-orthogonal
// (|<q_i, q_j>| <= sqrt(eps)), which keeps the Ritz values exact to rounding.
// Pass 1 applies the DGKS test against pass 0's norm; pass 1 or 2 closes the step.
__global__ void __launch_bounds__(64) eig_finish_kernel(EigArgs e, int k, int pass, double eta2, double tiny) {
    const int p = blockIdx.x;
    if (!eig_active(e, p, pass)) return;
    const int lane = threadIdx.x;
    const double wp2 = __shfl(wave_item_sum(e.part, e.pitem0[p], e.pnitem[p]), 0, 64);
    double* aux = e.aux + (size_t)p * 4;
    const size_t ld = e.ldh;
    double* om = e.om + (size_t)p * 3 * ld;
    double* wn = om + (size_t)((k + 1) % 3) * ld;  // omega_{k+1}
    const double eps = 0x1p-52;
    if (pass == 0) {
        const double* alpha = e.alpha + (size_t)p * ld;
        const double* beta = e.beta + (size_t)p * ld;
        const double* wc = om + (size_t)(k % 3) * ld;        // omega_k
        const double* wp = om + (size_t)((k + 2) % 3) * ld;  // omega_{k-1}
        const double ak = aux[0], bk = sqrt(wp2);
        double tn = 0.0;  // |T| estimate: max over rows of |alpha| + neighbouring betas
        for (int i = lane; i < k; i += 64) tn = fmax(tn, fabs(alpha[i]) + beta[i] + (i > 0 ? beta[i - 1] : 0.0));
        tn = fmax(wave_max(tn), fabs(ak) + bk + (k > 0 ? beta[k - 1] : 0.0));
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
        mx = wave_max(mx);
        if (lane != 0) return;
        if (k == 0) om[0] = 1.0;  // omega_{0,0}
        wn[k] = bk > 0.0 ? rt / bk : 0.0;
        wn[k + 1] = 1.0;
        aux[1] = wp2;
        const bool need = bk > 0.0 && (e.full_mode || mx > sqrt(eps) || e.force[p]);
        if (need) {
            e.full[p] = 1;
            e.force[p] = (e.force[p] || e.full_mode) ? 0 : 1;
        } else {
            e.full[p] = 0;
            eig_close_step(e, p, k, wp2, tiny);
        }
        return;
    }
    if (lane != 0) return;
    aux[2] = wp2;
    // second pass: w' lost more than half its norm (DGKS), or always with full reorthogonalisation
    if (pass == 1 && (e.full_mode || wp2 < eta2 * aux[1])) {
        e.pass2[p] = 1;
        aux[1] = wp2;
        return;
    }
    e.pass2[p] = 0;
    for (int i = 0; i < k; ++i) wn[i] = eps;  // V_{k+1} is now orthogonal to the basis to rounding
    eig_close_step(e, p, k, wp2, tiny);
}

// One Gram-Schmidt pass of step k.  Pass 0: w (= H V_k) against V_{k-1}, V_k only, the three-term
// part, written to slot k + 1; pass 1 (only where pass 0's overlap estimates ask for it): slot
// k + 1 against the whole basis V_0..V_k, in place; pass 2 (only where pass 1's DGKS test asks
// for it): pass 1 again.
hipError_t launch_eig_pass(const EigArgs& e, const double* const* V, double* vk1, const double* w, int k, int pass,
                           double eta2, double tiny, hipStream_t s) {
    const int k1 = k + 1;
    const int i0 = pass == 0 ? (k > 0 ? k - 1 : 0) : 0;
    const double* src = pass == 0 ? w : vk1;
    double* dst = vk1;
    const int nv = k1 - i0;
    const int ng = (nv + kEigVG - 1) / kEigVG;
    if (e.full_mode && pass > 0) {  // CGS2 against the whole basis on the matrix cores
        hipLaunchKernelGGL(eig_multidot_mfma_kernel, dim3(e.ni, (nv + 15) / 16), dim3(kEigThreads), 0, s, e, V, src,
                           i0, k1, pass);
    } else {
        hipLaunchKernelGGL(eig_multidot_kernel, dim3(e.ni, ng), dim3(kEigThreads), 0, s, e, V, src, i0, k1, pass);
    }
    hipLaunchKernelGGL(eig_reduce_h_kernel, dim3(e.P, (nv + 3) / 4), dim3(kEigThreads), 0, s, e, i0, k1, pass);
    if (e.full_mode && pass > 0) {
        hipLaunchKernelGGL(eig_project_mfma_kernel, dim3(e.ni), dim3(kEigThreads), (size_t)k1 * 2 * sizeof(double), s,
                           e, V, src, dst, i0, k1, pass);
    } else {
        hipLaunchKernelGGL(eig_project_kernel, dim3(e.ni), dim3(kEigThreads), (size_t)k1 * sizeof(double), s, e, V,
                           src, dst, i0, k1, pass);
    }
    hipLaunchKernelGGL(eig_finish_kernel, dim3(e.P), dim3(64), 0, s, e, k, pass, eta2, tiny);
    return hipGetLastError();
}

// ------------------------------------------------------------------ many reference recurrences
//
// lanczos_iteration (ed_lanczos.F90:37-68) for several chains at once, packed like the eigensolver's
// problems: per chain p the scalars A[p][0..], B[p][0..] (B = squared norms of the unnormalised
// vectors u_j), a stop flag and the step count at breakdown.  Per step: u_j from w = H u_{j-1} and
// the two previous vectors (one launch for every chain), B[p][j] = |u_j|^2, w = H u_j, A[p][j] =
// <u_j, w> / B[p][j], breakdown when B[p][j] < thr for j >= 2 (a chain that stops is skipped).

// out = (w - A[j-1] u1) / sqrt(B[j-1]) + c3 u0 over the item, c3 = -sqrt(B[j-1] / B[j-2]) (j >= 2);
// part[item] = |out|^2 over the item.  The same operations per element as lz_update_kernel.
__global__ void __launch_bounds__(kEigThreads) lzb_update_kernel(EigArgs e, int j, const double* w, const double* u1,
                                                                 const double* u0, double* out, const double* A,
                                                                 const double* B, int ldh, const int32_t* stop) {
    __shared__ double red[kEigThreads / 64];
    const int item = blockIdx.x;
    const int p = e.item_prob[item];
    if (stop[p]) return;
    const double* Bp = B + (size_t)p * ldh;
    const double sb = sqrt(Bp[j - 1]);
    const double c1 = 1.0 / sb, a = A[(size_t)p * ldh + j - 1];
    const double c3 = (j >= 2) ? -(sb / sqrt(Bp[j - 2])) : 0.0;
    const int64_t base = e.item_start[item];
    const int len = e.item_len[item];
    double s = 0.0;
    for (int r = threadIdx.x; r < len; r += kEigThreads) {
        const int64_t i = base + r;
        double v = (w[i] - a * u1[i]) * c1;
        if (j >= 2) v += c3 * u0[i];
        out[i] = v;
        s += v * v;
    }
    s = block_sum<kEigThreads>(s, red);
    if (threadIdx.x == 0) e.part[item] = s;
}

// one wave per chain: which = 0: B[p][j] = sum of the item partials; which = 1: A[p][j] = sum / B[p][j]
// and the breakdown test (j >= 2 and B[p][j] < thr: stop, nsc = j)
__global__ void __launch_bounds__(64) lzb_finish_kernel(EigArgs e, int j, int which, double* A, double* B, int ldh,
                                                        int32_t* stop, int32_t* nsc, double thr) {
    const int p = blockIdx.x;
    if (stop[p]) return;
    const double sum = wave_item_sum(e.part, e.pitem0[p], e.pnitem[p]);
    if (threadIdx.x != 0) return;
    double* Bp = B + (size_t)p * ldh;
    if (which == 0) {
        Bp[j] = sum;
        return;
    }
    A[(size_t)p * ldh + j] = sum / Bp[j];
    if (j >= 2 && Bp[j] < thr) {
        stop[p] = 1;
        nsc[p] = j;
    }
}

hipError_t launch_lzb_update(const EigArgs& e, int j, const double* w, const double* u1, const double* u0,
                             double* out, const double* A, const double* B, int ldh, const int32_t* stop,
                             hipStream_t s) {
    hipLaunchKernelGGL(lzb_update_kernel, dim3(e.ni), dim3(kEigThreads), 0, s, e, j, w, u1, u0, out, A, B, ldh, stop);
    return hipGetLastError();
}

hipError_t launch_lzb_finish(const EigArgs& e, int j, int which, double* A, double* B, int ldh, int32_t* stop,
                             int32_t* nsc, double thr, hipStream_t s) {
    hipLaunchKernelGGL(lzb_finish_kernel, dim3(e.P), dim3(64), 0, s, e, j, which, A, B, ldh, stop, nsc, thr);
    return hipGetLastError();
}

// part[item] = <a, b> over the item, for chains that have not stopped
__global__ void __launch_bounds__(kEigThreads) lzb_dot_kernel(EigArgs e, const double* a, const double* b,
                                                              const int32_t* stop) {
    __shared__ double red[kEigThreads / 64];
    const int item = blockIdx.x;
    if (stop[e.item_prob[item]]) return;
    const int64_t base = e.item_start[item];
    const int len = e.item_len[item];
    double s = 0.0;
    for (int r = threadIdx.x; r < len; r += kEigThreads) s = fma(a[base + r], b[base + r], s);
    s = block_sum<kEigThreads>(s, red);
    if (threadIdx.x == 0) e.part[item] = s;
}

hipError_t launch_lzb_dot(const EigArgs& e, const double* a, const double* b, const int32_t* stop, hipStream_t s) {
    hipLaunchKernelGGL(lzb_dot_kernel, dim3(e.ni), dim3(kEigThreads), 0, s, e, a, b, stop);
    return hipGetLastError();
}

// ------------------------------------------------------------------ deflated operator

// c_i = -sigma_p <L_i, V_k> / |V_k| from the item partials of eig_multidot_kernel (locked vectors
// L_0..L_{nl-1}); one wave per problem
__global__ void __launch_bounds__(64) eig_defl_coef_kernel(EigArgs e, int k, int nl, const double* sigma) {
    const int p = blockIdx.x;
    if (e.done[p]) return;
    for (int i = 0; i < nl; ++i) {
        const double sum = wave_item_sum(e.part, (size_t)i * e.ni + e.pitem0[p], e.pnitem[p]);
        if (threadIdx.x == 0) e.c[(size_t)p * e.ldh + i] = -sigma[p] * (sum / e.nrm[(size_t)p * e.ldh + k]);
    }
}

hipError_t launch_eig_deflate(const EigArgs& e, const double* const* L, int nl, const double* sigma,
                              const double* vk, double* w, int k, hipStream_t s) {
    if (nl <= 0) return hipSuccess;
    const int ng = (nl + kEigVG - 1) / kEigVG;
    hipLaunchKernelGGL(eig_multidot_kernel, dim3(e.ni, ng), dim3(kEigThreads), 0, s, e, L, vk, 0, nl, 0);
    hipLaunchKernelGGL(eig_defl_coef_kernel, dim3(e.P), dim3(64), 0, s, e, k, nl, sigma);
    hipLaunchKernelGGL(eig_project_kernel, dim3(e.ni), dim3(kEigThreads), (size_t)nl * sizeof(double), s, e, L, w, w,
                       0, nl, 0);
    return hipGetLastError();
}

// ------------------------------------------------------------------ convergence on the device

// After step m - 1 (basis size m): the lowest nev Ritz pairs of T_m pass |beta_m z_{m,i}| <= tol
// |theta_i| or lie above e.above[p] (ritz_converged) -> cflag[p] = m.  One wave per problem.  The
// test runs on its own stream beside the steps that follow: it reads alpha, beta[0, m), which no
// later step writes, and it leaves done and mfinal to launch_eig_conv_apply on the steps' stream.
// A problem already stopped at a basis size <= m (exhaustion or cap, mfinal > 0) is not tested; a
// later step's stop (mfinal > m) does not keep the test from passing at m.
__global__ void __launch_bounds__(64) eig_conv_kernel(EigArgs e, int m, int nev, double tol, double* zs,
                                                      int ws_lds) {
    extern __shared__ double t3[];  // 3 ldh: T_m staged; then the workspace when it fits
    const int p = blockIdx.x;
    if (m < nev || e.cflag[p] != 0) return;
    const int mf = __atomic_load_n(e.mfinal + p, __ATOMIC_RELAXED);
    if (mf > 0 && mf <= m) return;
    const int ldh = e.ldh;
    double* ws = ws_lds ? t3 + 3 * ldh : zs + (size_t)p * nev * 3 * ldh;
    const bool conv = ritz_converged(e.alpha + (size_t)p * ldh, e.beta + (size_t)p * ldh, m, nev, tol, t3, t3 + ldh,
                                     t3 + 2 * ldh, ws, ldh, e.above ? e.above[p] : INFINITY);
    if (threadIdx.x == 0 && conv) e.cflag[p] = m;
}

__global__ void eig_conv_apply_kernel(EigArgs e) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= e.P) return;
    const int c = e.cflag[p];
    if (c > 0 && (!e.done[p] || e.mfinal[p] > c)) {
        e.done[p] = 1;
        e.mfinal[p] = c;
    }
}

hipError_t launch_eig_conv_apply(const EigArgs& e, hipStream_t s) {
    hipLaunchKernelGGL(eig_conv_apply_kernel, dim3((e.P + 255) / 256), dim3(256), 0, s, e);
    return hipGetLastError();
}

hipError_t launch_eig_conv(const EigArgs& e, int m, int nev, double tol, double* zs, hipStream_t s) {
    // the inverse-iteration workspace (at most 64 eigenvalues at a time) in LDS where it fits
    const size_t ws = (size_t)std::min(nev, 64) * 3 * e.ldh * sizeof(double);
    const size_t t3 = (size_t)3 * e.ldh * sizeof(double);
    const int in_lds = t3 + ws <= 96 * 1024;
    const size_t lds = t3 + (in_lds ? ws : 0);
    if (lds > 64 * 1024) {
        hipError_t err = hipFuncSetAttribute((const void*)eig_conv_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                             (int)lds);
        if (err != hipSuccess) return err;
    }
    hipLaunchKernelGGL(eig_conv_kernel, dim3(e.P), dim3(64), lds, s, e, m, nev, tol, zs, in_lds);
    return hipGetLastError();
}

// ------------------------------------------------------------------ Ritz vectors

// y = sum_{i < m_p} S_i V_i over the item, written to out[p] + j*dim_p; part[j*ni + item] = |y|^2
__global__ void __launch_bounds__(kEigThreads) eig_ritz_kernel(EigArgs e, const double* const* V, const double* S,
                                                               int lds, int nev, const int32_t* nf,
                                                               double* const* out) {
    extern __shared__ double ss[];
    __shared__ double red[kEigThreads / 64];
    const int item = blockIdx.x, j = blockIdx.y;
    const int p = e.item_prob[item];
    if (j >= nf[p]) return;
    const int m = e.mfinal[p];
    const double* Sp = S + ((size_t)p * nev + j) * lds;
    for (int i = threadIdx.x; i < m; i += kEigThreads) ss[i] = Sp[i];
    __syncthreads();
    const int64_t base = e.item_start[item];
    const int len = e.item_len[item];
    constexpr int R = kEigChunk / kEigThreads;
    double acc[R];
#pragma unroll
    for (int r = 0; r < R; ++r) acc[r] = 0.0;
    for (int i = 0; i < m; ++i) {
        const double si = ss[i];
        const double* __restrict__ u = V[i] + base;
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int t = threadIdx.x + r * kEigThreads;
            if (t < len) acc[r] = fma(si, u[t], acc[r]);
        }
    }
    double* y = out[p] + (size_t)j * e.pdim[p] + (base - e.poff[p]);
    double s2 = 0.0;
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int t = threadIdx.x + r * kEigThreads;
        if (t < len) {
            y[t] = acc[r];
            s2 = fma(acc[r], acc[r], s2);
        }
    }
    s2 = block_sum<kEigThreads>(s2, red);
    if (threadIdx.x == 0) e.part[(size_t)j * e.ni + item] = s2;
}

// Ritz vectors on the matrix cores: out[p] + j dim_p = sum_{i < m_p} S_{j,i} V_i over the item for
// the 16 Ritz vectors j = 16 blockIdx.y + n (n < 16, j < nf[p]); part[j * ni + item] = |y_j|^2.
// Per 16-element slice and 4 basis vectors: A[m][k] = V_{i+k}[slice + m], B[k][n] = S_{j,i+k}.
__global__ void __launch_bounds__(kEigThreads) eig_ritz_mfma_kernel(EigArgs e, const double* const* V,
                                                                    const double* S, int lds, int nev,
                                                                    const int32_t* nf, double* const* out) {
    extern __shared__ double ss[];  // [16][m]
    __shared__ double red[kEigThreads / 64][16];
    const int item = blockIdx.x, j0 = blockIdx.y * 16;
    const int p = e.item_prob[item];
    const int nj = min(16, nf[p] - j0);
    if (nj <= 0) return;
    const int m = e.mfinal[p];
    for (int t = threadIdx.x; t < 16 * m; t += kEigThreads) {
        const int n = t / m, i = t % m;
        ss[t] = n < nj ? S[((size_t)p * nev + j0 + n) * lds + i] : 0.0;
    }
    __syncthreads();
    const int64_t base = e.item_start[item];
    const int len = e.item_len[item];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const int lm = lane & 15, lk = lane >> 4;  // A row / B column, and the k index
    constexpr int kPerWave = kEigChunk / (kEigThreads / 64);
    double s2 = 0.0;  // |y_n|^2 over this lane's elements, n = lane % 16
    const int64_t ybase = base - e.poff[p];
    double* y = lm < nj ? out[p] + (size_t)(j0 + lm) * e.pdim[p] + ybase : nullptr;
    for (int s0 = wid * kPerWave; s0 < (wid + 1) * kPerWave; s0 += 64) {  // four slices, four MFMA chains
        mf4 acc[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) acc[t] = mf4{0.0, 0.0, 0.0, 0.0};
        for (int i = 0; i < m; i += 4) {
            const int ii = i + lk;
            const bool have = ii < m;
            const double* __restrict__ v = V[have ? ii : 0] + base + s0 + lm;
            double a[4];
#pragma unroll
            for (int t = 0; t < 4; ++t) a[t] = (have && s0 + 16 * t + lm < len) ? v[16 * t] : 0.0;
            const double b = have ? ss[lm * m + ii] : 0.0;
#pragma unroll
            for (int t = 0; t < 4; ++t) acc[t] = mfma16(a[t], b, acc[t]);
        }
        if (y) {
#pragma unroll
            for (int t = 0; t < 4; ++t)
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const int rr = s0 + 16 * t + lk + 4 * j;
                    if (rr < len) {
                        y[rr] = acc[t][j];
                        s2 = fma(acc[t][j], acc[t][j], s2);
                    }
                }
        }
    }
    // lanes n, n + 16, n + 32, n + 48 hold parts of |y_n|^2
    s2 += __shfl_xor(s2, 16, 64);
    s2 += __shfl_xor(s2, 32, 64);
    if (lane < 16) red[wid][lane] = s2;
    __syncthreads();
    if (threadIdx.x < nj) {
        double t = 0.0;
        for (int v = 0; v < kEigThreads / 64; ++v) t += red[v][threadIdx.x];
        e.part[(size_t)(j0 + threadIdx.x) * e.ni + item] = t;
    }
}

__global__ void __launch_bounds__(64) eig_ritz_norm_kernel(EigArgs e, int nev, const int32_t* nf, double* n2) {
    const int p = blockIdx.x, j = blockIdx.y;
    if (j >= nf[p]) return;
    const double s = wave_item_sum(e.part, (size_t)j * e.ni + e.pitem0[p], e.pnitem[p]);
    if (threadIdx.x == 0) n2[(size_t)p * nev + j] = s;
}

__global__ void __launch_bounds__(kEigThreads) eig_ritz_scale_kernel(EigArgs e, int nev, const int32_t* nf,
                                                                     double* const* out, const double* n2) {
    const int item = blockIdx.x, j = blockIdx.y;
    const int p = e.item_prob[item];
    if (j >= nf[p]) return;
    const double f = 1.0 / sqrt(n2[(size_t)p * nev + j]);
    const int64_t base = e.item_start[item];
    const int len = e.item_len[item];
    double* y = out[p] + (size_t)j * e.pdim[p] + (base - e.poff[p]);
    for (int t = threadIdx.x; t < len; t += kEigThreads) y[t] = y[t] * f;
}

hipError_t launch_eig_ritz(const EigArgs& e, const double* const* V, const double* S, int lds, int nev,
                           const int32_t* nf, double* const* out, double* n2, hipStream_t s) {
    if (e.full_mode) {
        const size_t shm = (size_t)16 * lds * sizeof(double);
        if (shm > 64 * 1024) {
            hipError_t err = hipFuncSetAttribute((const void*)eig_ritz_mfma_kernel,
                                                 hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm);
            if (err != hipSuccess) return err;
        }
        hipLaunchKernelGGL(eig_ritz_mfma_kernel, dim3(e.ni, (nev + 15) / 16), dim3(kEigThreads), shm, s, e, V, S, lds,
                           nev, nf, out);
    } else {
        hipLaunchKernelGGL(eig_ritz_kernel, dim3(e.ni, nev), dim3(kEigThreads), (size_t)lds * sizeof(double), s, e, V,
                           S, lds, nev, nf, out);
    }
    hipLaunchKernelGGL(eig_ritz_norm_kernel, dim3(e.P, nev), dim3(64), 0, s, e, nev, nf, n2);
    hipLaunchKernelGGL(eig_ritz_scale_kernel, dim3(e.ni, nev), dim3(kEigThreads), 0, s, e, nev, nf, out, n2);
    return hipGetLastError();
}

}  // namespace moed
