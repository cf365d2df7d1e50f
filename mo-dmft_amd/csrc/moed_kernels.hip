// HIP kernels of libmoed for gfx950 (MI355X, CDNA4, wave64) and their launchers.
//
// Layout of a sector vector in HBM: x[iu + nU*id], up string rank iu fastest
// (ed_basis.F90:120-123).  A "row" is the nU contiguous doubles of one down string.
#include <hip/hip_runtime.h>

#include <cmath>

#include "moed_cplx.h"
#include "moed_device.h"
#include "moed_internal.h"

namespace moed {

// ---------------------------------------------------------------- apply_c

// y(out) = 0.0 + sgn * x(src) per target element (ed_operators.F90:44-58: vec_out(i) = vec_out(i) +
// vec_all(j)*sgn with vec_out zeroed), y = 0 where the operator annihilates the target state.
__global__ void apply_c_kernel(ApplyCArgs a) {
    const int iu = blockIdx.x * blockDim.x + threadIdx.x;
    const int id = blockIdx.y;
    if (iu >= a.nU_out) return;
    double v = 0.0;
    if (a.spin == 0) {
        const int32_t m = a.map[iu];
        if (m >= 0) {
            const double xv = a.x[(size_t)(m & 0x3fffffff) + (size_t)a.nU_in * id];
            v = 0.0 + ((m >> 30) & 1 ? -xv : xv);
        }
    } else {
        const int32_t m = a.map[id];
        if (m >= 0) {
            const double xv = a.x[(size_t)iu + (size_t)a.nU_in * (m & 0x3fffffff)];
            v = 0.0 + ((m >> 30) & 1 ? -xv : xv);
        }
    }
    a.y[(size_t)iu + (size_t)a.nU_out * id] = v;
}

hipError_t launch_apply_c(const ApplyCArgs& a, hipStream_t s) {
    dim3 grid((a.nU_out + 255) / 256, a.nD_out);
    hipLaunchKernelGGL(apply_c_kernel, grid, dim3(256), 0, s, a);
    return hipGetLastError();
}

// ---------------------------------------------------------------- fills

__global__ void fill_start_kernel(uint64_t seed, int64_t n, double* x) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        uint64_t z = (seed ^ (uint64_t)i) + 0x9E3779B97F4A7C15ull;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z = z ^ (z >> 31);
        const double u = (double)(z >> 11) * (1.0 / 9007199254740992.0);
        x[i] = 2.0 * u - 1.0;
    }
}

hipError_t launch_fill_start(uint64_t seed, int64_t n, double* x, hipStream_t s) {
    const int64_t nb = std::min<int64_t>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(fill_start_kernel, dim3((unsigned)std::max<int64_t>(nb, 1)), dim3(256), 0, s, seed, n, x);
    return hipGetLastError();
}

__global__ void fill_kernel(double v, int64_t n, double* x) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        x[i] = v;
}

hipError_t launch_fill(double v, int64_t n, double* x, hipStream_t s) {
    const int64_t nb = std::min<int64_t>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(fill_kernel, dim3((unsigned)std::max<int64_t>(nb, 1)), dim3(256), 0, s, v, n, x);
    return hipGetLastError();
}

__global__ void pad_rows_kernel(const double* __restrict__ src, int64_t nU, int64_t n, double* __restrict__ dst,
                                int64_t ldx) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int64_t id = i / nU;
    dst[id * ldx + (i - id * nU)] = src[i];
}

hipError_t launch_pad_rows(const double* src, int64_t nU, int64_t nD, double* dst, int64_t ldx, hipStream_t s) {
    const int64_t n = nU * nD;
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(pad_rows_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, src, nU, n, dst, ldx);
    return hipGetLastError();
}

__global__ void scale_copy_kernel(const double* x, double alpha, int64_t n, double* y) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        y[i] = alpha * x[i];
}

hipError_t launch_scale_copy(const double* x, double alpha, int64_t n, double* y, hipStream_t s) {
    const int64_t nb = std::min<int64_t>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(scale_copy_kernel, dim3((unsigned)std::max<int64_t>(nb, 1)), dim3(256), 0, s, x, alpha, n,
                       y);
    return hipGetLastError();
}

// ---------------------------------------------------------------- reductions

__global__ void __launch_bounds__(256) dot_partial_kernel(const double* a, const double* b, int64_t n, double* part) {
    __shared__ double red[4];
    double s = 0.0;
    for (int64_t i = blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) s += a[i] * b[i];
    s = block_sum<256>(s, red);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}

hipError_t launch_dot_partial(const double* a, const double* b, int64_t n, double* part, hipStream_t s) {
    hipLaunchKernelGGL(dot_partial_kernel, dim3(kRedBlocks), dim3(256), 0, s, a, b, n, part);
    return hipGetLastError();
}

__global__ void __launch_bounds__(1024) reduce_final_kernel(const double* part, int npart, double* out, int slot,
                                                            int op, const double* divp) {
    __shared__ double red[16];
    double s = 0.0;
    for (int i = threadIdx.x; i < npart; i += 1024) s += part[i];
    s = block_sum<1024>(s, red);
    if (threadIdx.x == 0) {
        if (op == 1) s = s / divp[0];
        if (op == 2) s = sqrt(s);
        out[slot] = s;
    }
}

hipError_t launch_reduce_final(const double* part, int npart, double* out, int slot, int op, const double* divp,
                               hipStream_t s) {
    hipLaunchKernelGGL(reduce_final_kernel, dim3(1), dim3(1024), 0, s, part, npart, out, slot, op, divp);
    return hipGetLastError();
}

// ---------------------------------------------------------------- Lanczos (reference recurrence)

__global__ void __launch_bounds__(256) lz_update_kernel(int j, const double* w, const double* u1, const double* u0,
                                                        double* out, int64_t n, LzScal sc, double* part) {
    __shared__ double red[4];
    if (*sc.stop) return;
    const double sb = sqrt(sc.B[j - 1]);
    const double c1 = 1.0 / sb, a = sc.A[j - 1];
    const double c3 = (j >= 2) ? -(sb / sqrt(sc.B[j - 2])) : 0.0;
    // two elements per thread and iteration as 16-byte accesses (any 8-byte alignment), the last
    // element of an odd n by thread 0 of block 0; each element as (w - a u1) c1 + c3 u0
    typedef double d2u __attribute__((ext_vector_type(2), aligned(8)));
    double s = 0.0;
    const int64_t np = n >> 1;
    for (int64_t p = blockIdx.x * 256 + threadIdx.x; p < np; p += (int64_t)gridDim.x * 256) {
        const d2u wv = reinterpret_cast<const d2u*>(w)[p], uv = reinterpret_cast<const d2u*>(u1)[p];
        double v0 = (wv.x - a * uv.x) * c1, v1 = (wv.y - a * uv.y) * c1;
        if (j >= 2) {
            const d2u zv = reinterpret_cast<const d2u*>(u0)[p];
            v0 += c3 * zv.x;
            v1 += c3 * zv.y;
        }
        const d2u o = {v0, v1};
        reinterpret_cast<d2u*>(out)[p] = o;
        s += v0 * v0;
        s += v1 * v1;
    }
    if ((n & 1) && blockIdx.x == 0 && threadIdx.x == 0) {
        const int64_t i = n - 1;
        double v = (w[i] - a * u1[i]) * c1;
        if (j >= 2) v += c3 * u0[i];
        out[i] = v;
        s += v * v;
    }
    s = block_sum<256>(s, red);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}

hipError_t launch_lz_update(int j, const double* w, const double* u1, const double* u0, double* out, int64_t n,
                            LzScal sc, double* part, hipStream_t s) {
    hipLaunchKernelGGL(lz_update_kernel, dim3(kRedBlocks), dim3(256), 0, s, j, w, u1, u0, out, n, sc, part);
    return hipGetLastError();
}

__global__ void __launch_bounds__(1024) lz_finish_b_kernel(int j, const double* part, int npart, LzScal sc) {
    __shared__ double red[16];
    if (*sc.stop) return;
    double s = 0.0;
    for (int i = threadIdx.x; i < npart; i += 1024) s += part[i];
    s = block_sum<1024>(s, red);
    if (threadIdx.x == 0) sc.B[j] = s;
}

__global__ void __launch_bounds__(1024) lz_finish_a_kernel(int j, const double* part, int npart, LzScal sc,
                                                           double thr) {
    __shared__ double red[16];
    if (*sc.stop) return;
    double s = 0.0;
    for (int i = threadIdx.x; i < npart; i += 1024) s += part[i];
    s = block_sum<1024>(s, red);
    if (threadIdx.x == 0) {
        sc.A[j] = s / sc.B[j];
        if (j >= 2 && sc.B[j] < thr) {
            *sc.stop = 1;
            *sc.nsc = j;
        }
    }
}

hipError_t launch_lz_finish_b(int j, const double* part, int npart, LzScal sc, hipStream_t s) {
    hipLaunchKernelGGL(lz_finish_b_kernel, dim3(1), dim3(1024), 0, s, j, part, npart, sc);
    return hipGetLastError();
}
hipError_t launch_lz_finish_a(int j, const double* part, int npart, LzScal sc, double thr, hipStream_t s) {
    hipLaunchKernelGGL(lz_finish_a_kernel, dim3(1), dim3(1024), 0, s, j, part, npart, sc, thr);
    return hipGetLastError();
}

// ---------------------------------------------------------------- continued fraction

// green_diag's continued fraction (ed_green.F90:160-168, 176-184) at one frequency per thread, with
// z = cmplx(+-E, w) rounded as the reference's default-kind cmplx() when compat is set, and
// G += weight * g in the reference's complex arithmetic.
// g(z) of green_diag's recurrence over A[0..n], B[0..n] at frequency w (sign +1: z - a_j, -1: z + a_j)
__device__ __forceinline__ cxd cf_eval(const double* A, const double* B, int n, int sign, double E, double w,
                                       int compat) {
    const double Es = sign > 0 ? E : -E;
    const cxd z = {compat ? round_f32(Es) : Es, compat ? round_f32(w) : w};
    cxd g;
    if (sign > 0) {
        g = cx_rdiv(B[n], {z.re - A[n], z.im});
        for (int j = n - 1; j >= 0; --j) g = cx_rdiv(B[j], {(z.re - A[j]) - g.re, z.im - g.im});
    } else {
        g = cx_rdiv(B[n], {z.re + A[n], z.im});
        for (int j = n - 1; j >= 0; --j) g = cx_rdiv(B[j], {(z.re + A[j]) - g.re, z.im - g.im});
    }
    return g;
}

__global__ void cf_kernel(CfArgs a) {
    extern __shared__ double sab[];  // a[0..n], b[0..n]
    for (int k = threadIdx.x; k <= a.n; k += blockDim.x) {
        sab[k] = a.a[k];
        sab[a.n + 1 + k] = a.b[k];
    }
    __syncthreads();
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.nw) return;
    const cxd g = cf_eval(sab, sab + a.n + 1, a.n, a.sign, a.E, a.omega[i], a.compat);
    double* G = a.G + (size_t)2 * i * a.stride;
    const cxd wg = cx_mul({a.weight, 0.0}, g);
    G[0] = G[0] + wg.re;
    G[1] = G[1] + wg.im;
}

// one thread per (frequency, orbital); the chains of the orbital in chain order, each chain's
// coefficients staged in LDS by the whole block in turn
__global__ void __launch_bounds__(256) cf_batch_kernel(CfBatchArgs a) {
    extern __shared__ double sab[];  // A[0..ldh), B[0..ldh) of the current chain
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int o = blockIdx.y;
    double* G = a.G + (size_t)2 * ((size_t)o + (size_t)a.norb * (i < a.nw ? i : 0));
    double gre = i < a.nw ? G[0] : 0.0, gim = i < a.nw ? G[1] : 0.0;
    for (int c = 0; c < a.nchain; ++c) {
        if (a.iorb[c] != o || !a.use[c]) continue;  // uniform over the block
        const int n = a.nsc[c];
        __syncthreads();
        for (int k = threadIdx.x; k <= n; k += blockDim.x) {
            sab[k] = a.A[(size_t)c * a.ldh + k];
            sab[a.ldh + k] = k == 0 ? a.b0[c] : a.B[(size_t)c * a.ldh + k];
        }
        __syncthreads();
        if (i < a.nw) {
            const cxd g = cf_eval(sab, sab + a.ldh, n, a.sign[c], a.E[c], a.omega[i], a.compat);
            const cxd wg = cx_mul({a.weight[c], 0.0}, g);
            gre = gre + wg.re;
            gim = gim + wg.im;
        }
    }
    if (i < a.nw) {
        G[0] = gre;
        G[1] = gim;
    }
}

hipError_t launch_cf_batch(const CfBatchArgs& a, hipStream_t s) {
    if (a.nchain <= 0) return hipSuccess;
    const size_t lds = (size_t)2 * a.ldh * sizeof(double);
    if (lds > 64 * 1024) return hipErrorInvalidValue;
    hipLaunchKernelGGL(cf_batch_kernel, dim3((a.nw + 255) / 256, a.norb), dim3(256), lds, s, a);
    return hipGetLastError();
}

hipError_t launch_cf(const CfArgs& a, hipStream_t s) {
    hipLaunchKernelGGL(cf_kernel, dim3((a.nw + 255) / 256), dim3(256), (size_t)2 * (a.n + 1) * sizeof(double), s, a);
    return hipGetLastError();
}

// dos_diag (dos.f90:87-103): z+ = cmplx(ev + w, small), z- = cmplx(w - ev, small) (REAL*4 parts in
// compat mode), the two recurrences of green_diag, then -aimag(G+ + G-)/pi
__global__ void dos_kernel(DosArgs a) {
#pragma clang fp contract(off)
    extern __shared__ double sab[];  // ap, bp, an, bn
    const int np = a.nsp + 1, nn = a.nsn + 1;
    for (int k = threadIdx.x; k < np; k += blockDim.x) {
        sab[k] = a.ap[k];
        sab[np + k] = a.bp[k];
    }
    for (int k = threadIdx.x; k < nn; k += blockDim.x) {
        sab[2 * np + k] = a.an[k];
        sab[2 * np + nn + k] = a.bn[k];
    }
    __syncthreads();
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.nw) return;
    const double* AP = sab;
    const double* BP = sab + np;
    const double* AN = sab + 2 * np;
    const double* BN = sab + 2 * np + nn;
    const double w = a.omega[i];
    const double zi = a.compat ? round_f32(a.small) : a.small;
    double zr = w + a.ev;
    zr = a.compat ? round_f32(zr) : zr;
    cxd gp = cx_rdiv(BP[a.nsp], {zr - AP[a.nsp], zi});
    for (int j = a.nsp - 1; j >= 0; --j) gp = cx_rdiv(BP[j], {(zr - AP[j]) - gp.re, zi - gp.im});
    zr = w - a.ev;
    zr = a.compat ? round_f32(zr) : zr;
    cxd gn = cx_rdiv(BN[a.nsn], {zr + AN[a.nsn], zi});
    for (int j = a.nsn - 1; j >= 0; --j) gn = cx_rdiv(BN[j], {(zr + AN[j]) - gn.re, zi - gn.im});
    const double pi = 0x1.921fb54442d18p+1;  // 4 atan(1) (ed_config.f90:7)
    const double v = -(gp.im + gn.im) / pi;
    double* o = a.out + (size_t)i * a.stride;
    *o = a.accumulate ? *o + a.weight * v : v;
}

hipError_t launch_dos(const DosArgs& a, hipStream_t s) {
    const size_t lds = (size_t)2 * (a.nsp + 1 + a.nsn + 1) * sizeof(double);
    hipLaunchKernelGGL(dos_kernel, dim3((a.nw + 255) / 256), dim3(256), lds, s, a);
    return hipGetLastError();
}

// Adjoint of the continued fraction.  With D_j = z -+ a_j - t_{j+1}, t_j = b_j / D_j, g = t_0:
// dg/dt_j = P_j = prod_{k<j} t_k / D_k, dg/db_j = P_j / D_j, dg/da_j = +-P_j t_j / D_j.
// For real theta: dL/dtheta = Re(conj(gbar) dg/dtheta), gbar = dL/dRe g + i dL/dIm g.
// contrib[(j)*nw + i] per frequency, then summed over frequencies in index order.
__global__ void cf_adjoint_kernel(CfArgs a, const double* gbar, double* scratch, double* ca, double* cb) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.nw) return;
    const int n = a.n, nw = a.nw;
    const double E = a.sign > 0 ? a.E : -a.E;
    const cxd z = {a.compat ? round_f32(E) : E, a.compat ? round_f32(a.omega[i]) : a.omega[i]};
    const double sgn = a.sign > 0 ? -1.0 : 1.0;  // D_j = z + sgn*a_j - t_{j+1}
    // scratch holds D_j and t_j per frequency: [(j*nw + i)*4]
    cxd t = {0.0, 0.0};
    for (int j = n; j >= 0; --j) {
        const cxd D = {(z.re + sgn * a.a[j]) - t.re, z.im - t.im};
        t = cx_rdiv(a.b[j], D);
        double* sp = scratch + ((size_t)j * nw + i) * 4;
        sp[0] = D.re;
        sp[1] = D.im;
        sp[2] = t.re;
        sp[3] = t.im;
    }
    const cxd gb = {gbar[2 * i], gbar[2 * i + 1]};
    cxd P = {1.0, 0.0};
    for (int j = 0; j <= n; ++j) {
        const double* sp = scratch + ((size_t)j * nw + i) * 4;
        const cxd D = {sp[0], sp[1]}, tj = {sp[2], sp[3]};
        const cxd dgb = cx_div(P, D);                          // dg/db_j
        const cxd dga = cx_mul({-sgn, 0.0}, cx_div(cx_mul(P, tj), D));  // dg/da_j
        ca[(size_t)j * nw + i] = gb.re * dga.re + gb.im * dga.im;
        cb[(size_t)j * nw + i] = gb.re * dgb.re + gb.im * dgb.im;
        P = cx_div(cx_mul(P, tj), D);
    }
}

__global__ void __launch_bounds__(256) sum_rows_kernel(const double* c, int nw, double* out) {
    __shared__ double red[4];
    double s = 0.0;
    for (int i = threadIdx.x; i < nw; i += 256) s += c[(size_t)blockIdx.x * nw + i];
    s = block_sum<256>(s, red);
    if (threadIdx.x == 0) out[blockIdx.x] = s;
}

hipError_t launch_cf_adjoint(const CfArgs& a, const double* gbar, double* scratch, double* ga, double* gb,
                             hipStream_t s) {
    const size_t nw = a.nw, n1 = a.n + 1;
    double* ca = scratch + n1 * nw * 4;
    double* cbp = ca + n1 * nw;
    hipLaunchKernelGGL(cf_adjoint_kernel, dim3((a.nw + 255) / 256), dim3(256), 0, s, a, gbar, scratch, ca, cbp);
    hipLaunchKernelGGL(sum_rows_kernel, dim3((unsigned)n1), dim3(256), 0, s, ca, a.nw, ga);
    hipLaunchKernelGGL(sum_rows_kernel, dim3((unsigned)n1), dim3(256), 0, s, cbp, a.nw, gb);
    return hipGetLastError();
}

// ---------------------------------------------------------------- bath fit (FUNC + dfunc)

// x packing of ev_to_x (func.f90:165-191): x[0:nbath] bath energies, x[nbath + o*nbath + k] = V(o, k),
// x[nx-norb+o] = ef(o).
__device__ __forceinline__ cxd fit_delta(const double* x, int nx, int norb, int nbath, int o, cxd z) {
    // Delta_o(i w) = sum_k V(o,k)^2 / (z - e_k) + ef_o  (func.f90:19-29), in the reference's order
    double re = 0.0, im = 0.0;
    for (int k = 0; k < nbath; ++k) {
        const double v = x[nbath + o * nbath + k];
        const cxd q = cx_rdiv(v * v, {z.re - x[k], z.im});
        re = re + q.re;
        im = im + q.im;
    }
    return {re + x[nx - norb + o], im};
}

// r_o(w) = Delta_o(i w) - Dev_o(w) for one parameter vector, one thread per frequency; with a zero
// target this is calc_dev's Delta (ed_green.F90:260-273)
__global__ void fit_residual_kernel(FitArgs a) {
    const int iw = blockIdx.x * blockDim.x + threadIdx.x;
    if (iw >= a.nw) return;
    const cxd z = {0.0, a.compat ? round_f32(a.omega[iw]) : a.omega[iw]};
    for (int o = 0; o < a.norb; ++o) {
        const cxd d = fit_delta(a.x, a.nx, a.norb, a.nbath, o, z);
        const size_t io = (size_t)o + (size_t)a.norb * iw;
        a.work[2 * io] = d.re - a.dev[2 * io];
        a.work[2 * io + 1] = d.im - a.dev[2 * io + 1];
    }
}

hipError_t launch_fit_residual(const FitArgs& a, hipStream_t s) {
    hipLaunchKernelGGL(fit_residual_kernel, dim3((a.nw + 127) / 128), dim3(128), 0, s, a);
    return hipGetLastError();
}

// ---------------------------------------------------------------- lattice step

// ed_delta_new (ed_green.F90:187-215): one thread per (orbital blockIdx.y, frequency).  The k-sum
// runs over q in order; 1/Gr + D_ev + rMu is the same for every q and is formed once.  Gksum /
// float(Nq) is the complex division by (Nq, 0), as the reference build evaluates it.
__global__ void __launch_bounds__(256) delta_new_kernel(DeltaArgs a) {
#pragma clang fp contract(off)
    const int iw = blockIdx.x * blockDim.x + threadIdx.x;
    const int o = blockIdx.y;
    if (iw >= a.nw) return;
    const size_t e = (size_t)iw * a.norb + o;
    const cxd G = {a.gr[2 * e], a.gr[2 * e + 1]};
    const cxd D = {a.dev[2 * e], a.dev[2 * e + 1]};
    const cxd g1 = cx_rdiv(1.0, G);
    const cxd base = {(g1.re + D.re) + a.rmu, (g1.im + D.im) + 0.0};
    const cptr<double> hk = as_const(a.hk);
    cxd s = {0.0, 0.0};
    for (int q = 0; q < a.nq; ++q) {
        const cxd t = {base.re - hk[(size_t)q * a.norb + o], base.im - 0.0};
        const cxd r = cx_rdiv(1.0, t);
        s = {s.re + r.re, s.im + r.im};
    }
    s = cx_div(s, {(double)(float)a.nq, 0.0});
    if (a.gksum) {
        a.gksum[2 * e] = s.re;
        a.gksum[2 * e + 1] = s.im;
    }
    const cxd gi = cx_rdiv(1.0, s);
    const cxd dn = {(g1.re - gi.re) + D.re, (g1.im - gi.im) + D.im};
    a.dev[2 * e] = dn.re;
    a.dev[2 * e + 1] = dn.im;
}

hipError_t launch_delta_new(const DeltaArgs& a, hipStream_t s) {
    if (a.nw <= 0 || a.norb <= 0) return hipSuccess;
    hipLaunchKernelGGL(delta_new_kernel, dim3((a.nw + 255) / 256, a.norb), dim3(256), 0, s, a);
    return hipGetLastError();
}

}  // namespace moed
