// Bath fit on the device: chi^2 (FUNC) and its gradient (dfunc) of func.f90:1-142, and the whole
// conjugate-gradient minimiser of frprmn.f90 (frprmn, linmin, mnbrak, brent, f1dim) for many
// independent fits in one launch.
//
// One workgroup owns one fit for its whole life: the CG vectors (p, xi, g, h, pcom, xicom) sit in
// LDS, every thread keeps the same copy of the line-search scalars, and each function evaluation
// is a block-wide pass over the Matsubara frequencies.  Nothing returns to the host between
// evaluations.  A launch runs at most `budget` CG iterations per fit and leaves the state at an
// iteration boundary in HBM; the host relaunches until every fit has finished, so one launch stays
// short and the result does not depend on the budget.
//
// Sums over frequencies come in two orders:
//  * reference order (default): every sum gives the result of the reference's sequential loop, so
//    chi^2, the gradient and the CG trajectory are bit-identical to the reference build.  The
//    gradient's sums run in one lane per sum over terms staged in LDS; chi^2's sum, the one every
//    line-search step waits for, is replayed by the whole block (moed_osum.h);
//  * tree order (MOED_F_FIT_TREE): per-thread partial sums over a fixed frequency stride, then a
//    fixed wave/block tree.  Deterministic and independent of the batch, but rounded differently
//    from the reference; far less latency per evaluation.
// Couplings V(o,k) that are exactly zero are skipped: each of their terms is a signed zero, and a
// running sum that starts at +0.0 is unchanged by adding signed zeros (inputs are finite).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <mutex>
#include <string>
#include <vector>

#include "moed_cplx.h"
#include "moed_device.h"
#include "moed_fit.h"
#include "moed_internal.h"
#include "moed_osum.h"

namespace moed {
namespace {

constexpr int kFitThreads = 1024;           // 16 waves: four per SIMD hide the division latency
constexpr int kFitNx = 64;                    // >= nxsize; the reference's nmax is 60 (frprmn.f90:12)
constexpr int kFitMaxBath = kMaxSite - 1;
constexpr int kFitCols = 2 * kFitMaxBath + 1;  // gradient columns of one orbital pass: e_k, V(o,k), ef_o
constexpr int kSumSeq = 512;                   // reference-order FUNC: leading terms added by one lane
constexpr int kFitPairF = 2 * kFitThreads;     // reference-order FUNC: frequencies per chunk, two per thread
constexpr int kFitChunkG = 256;                // frequencies per staged gradient chunk
constexpr int kGrp = 4;                        // tree order: frequencies per thread and group
constexpr int kItmax = 20000, kBrentItmax = 100;  // frprmn.f90:12, :126
constexpr int kStateD = 3 * kFitNx + 2;       // per fit: xi, g, h, fp, fret
constexpr int kStateI = 4;                    // per fit: phase, its, rc, nfev

enum { kPhaseStart = 0, kPhaseRun = 1, kPhaseDone = 2 };

struct FitProb {
    int32_t nx, norb, nbath, nw, compat, tree;
    const double* omega;
    const double* dev;  // interleaved complex (norb, nw) column-major
};

struct FitCgArgs {
    FitProb q;
    int32_t nfit, budget;
    double ftol;
    double* x;       // [nfit][nx]: start point in, result out
    double* st;      // [nfit][kStateD]
    int32_t* ist;    // [nfit][kStateI]
    int32_t* active; // fits still running after this launch
};

struct FitEvalArgs {
    FitProb q;
    const double* x;  // [nfit][nx]
    double* f;        // [nfit] or null
    double* g;        // [nfit][nx] or null
};

struct FitSmem {
    double p[kFitNx], xi[kFitNx], g[kFitNx], h[kFitNx], pcom[kFitNx], xicom[kFitNx], xt[kFitNx];
    double gacc[kFitNx];
    double bcast[2];
    double red[kFitThreads / 64];
    double gred[kFitThreads / 64][kFitCols];      // per-wave column sums (tree gradient)
    // tree order: the nonzero couplings of each orbital at the point xt (V^2, V, e_k, k)
    double nzv2[kMaxOrb][kFitMaxBath], nzv[kMaxOrb][kFitMaxBath], nzek[kMaxOrb][kFitMaxBath];
    int32_t nzk[kMaxOrb][kFitMaxBath], nnz[kMaxOrb];
    union alignas(16) {  // 16-byte aligned: chain_sum reads the staged terms in pairs
        struct {  // reference-order FUNC: the terms of kFitPairF frequencies
            double T[kFitPairF * kMaxOrb];
        } os;
        double cols[kFitCols * kFitChunkG];      // gradient terms of one chunk (reference order)
        struct {  // tree order, Nw <= kFitThreads: thread t's frequency iw = t, set once per launch
            double zi[kFitThreads], rw[kFitThreads];                  // Im z, 1/w_n
            double dre[kMaxOrb][kFitThreads], dim[kMaxOrb][kFitThreads];  // the target Dev_o(w)
        } tc;
        struct {  // tree order, kFitThreads < Nw <= kGrp kFitThreads: iw = t + j kFitThreads
            double zi[kGrp * kFitThreads], rw[kGrp * kFitThreads];  // Im z, 1/w_n (0 past the end)
        } tw;
    } u;
    // tree order, per line search: the couplings that can be nonzero on the line pcom + t xicom
    // (V(o,k) with pcom or xicom nonzero; any other V is exactly 0 at every finite t)
    int32_t lk[kMaxOrb][16], lnn[kMaxOrb];
    // tree order, per wave: the couplings of the point being evaluated, formed by the wave itself
    struct {
        double ek[kMaxOrb][16], v2[kMaxOrb][16], ef[kMaxOrb];
    } wp[kFitThreads / 64];
    double fred[2][kFitThreads / 64];  // tree-order FUNC: the waves' partial sums (two buffers)
    // reference-order FUNC (ordered_sum): the waves' count totals and first ties (-1: a negative or
    // NaN term), and the running sum and position handed back
    double os_s, os_qk[kFitThreads / 64];
    int32_t os_pos, os_qf[kFitThreads / 64];
    int32_t tc_mode;                   // 1: u.tc, 2: u.tw holds the frequencies of this launch
    int32_t cmd, cbuf;                 // the minimiser wave's posted command and its arguments
    double cxv;
};

// the fit's LDS, shared by every device function of this file (one workgroup = one fit)
__shared__ FitSmem S;

template <class T>
using gptr = const __attribute__((address_space(1))) T*;

__device__ __forceinline__ const double* uniform_ptr(const double* p) {
    const uint64_t v = (uint64_t)p;
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v), hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
    return (const double*)(((uint64_t)hi << 32) | lo);
}

// the problem with every field wave-uniform (scalar registers) inside out-of-line functions
struct FitView {
    int nx, norb, nbath, nw, compat;
    gptr<double> omega;
    gptr<double> dev;
};
__device__ __forceinline__ FitView view(const FitProb& q) {
    return FitView{uniform_i32(q.nx), uniform_i32(q.norb), uniform_i32(q.nbath), uniform_i32(q.nw),
                   uniform_i32(q.compat), (gptr<double>)uniform_ptr(q.omega), (gptr<double>)uniform_ptr(q.dev)};
}

__device__ __forceinline__ double zim(const FitView& q, double w) { return q.compat ? round_f32(w) : w; }

// Delta_o(i w) - Dev_o(w) at the point xt (func.f90:19-29): the k sum in the reference's order
__device__ __forceinline__ cxd residual(const FitView& q, const double* xt, int o, int iw, double zi) {
#pragma clang fp contract(off)
    double re = 0.0, im = 0.0;
    for (int k = 0; k < q.nbath; ++k) {
        const double v = xt[q.nbath + o * q.nbath + k];
        if (v == 0.0) continue;
        const cxd t = cx_rdiv(v * v, {0.0 - xt[k], zi});
        re = re + t.re;
        im = im + t.im;
    }
    const size_t io = (size_t)o + (size_t)q.norb * iw;
    return {(re + xt[q.nx - q.norb + o]) - q.dev[2 * io], im - q.dev[2 * io + 1]};
}

// s + buf[0] + buf[1] + ... + buf[n-1], one rounding per add in this order.  buf (LDS) is 16-byte
// aligned.  The terms arrive as pairs through a ring of kChainPairs 16-byte loads, refilled
// kChainGroup loads at a time right after their pairs are added: an add waits only for loads
// issued a whole ring earlier.  Measured on gfx950 with one lane active: a dependent FP64 add takes
// 8.4 cycles and each LDS load instruction holds its wave ~8 cycles, so an LDS-fed chain costs
// ~12.6 cycles per add with loads in groups of four (15.3 with single loads between adds).
constexpr int kChainPairs = 16, kChainGroup = 4;
__device__ __forceinline__ double chain_sum(double s, const double* buf, int n) {
#pragma clang fp contract(off)
    const double2* b2 = reinterpret_cast<const double2*>(buf);
    const int n2 = n >> 1;
    int j = 0;  // pairs loaded
    if (n2 >= 2 * kChainPairs) {
        double2 r[kChainPairs];
#pragma unroll
        for (int t = 0; t < kChainPairs; ++t) r[t] = b2[t];
        __builtin_amdgcn_sched_barrier(0);
        for (j = kChainPairs; j + kChainPairs <= n2; j += kChainPairs) {
#pragma unroll
            for (int g = 0; g < kChainPairs; g += kChainGroup) {
#pragma unroll
                for (int t = g; t < g + kChainGroup; ++t) {
                    s = s + r[t].x;
                    s = s + r[t].y;
                }
                __builtin_amdgcn_sched_barrier(0);  // keep the ring's order
#pragma unroll
                for (int t = g; t < g + kChainGroup; ++t) r[t] = b2[j + t];
                __builtin_amdgcn_sched_barrier(0);
            }
        }
#pragma unroll
        for (int t = 0; t < kChainPairs; ++t) {
            s = s + r[t].x;
            s = s + r[t].y;
        }
    }
    for (; j < n2; ++j) {
        const double2 v = b2[j];
        s = s + v.x;
        s = s + v.y;
    }
    if (n & 1) s = s + buf[n - 1];
    return s;
}

// Tree order restates the terms with fewer divisions (each term then differs from the reference's
// by a few ulps): 1/(i w - e_k) = (-e_k - i w) / (e_k^2 + w^2) with one division, 1/w_n once per
// frequency, |r|^2 instead of hypot(r)^2.
// 1/d for d > 0: v_rcp_f64 (about 2^-26 relative) and two Newton steps (the error squares each
// step: rounding level, within an ulp of the division)
__device__ __forceinline__ double fast_rcp(double d) {
    double r = __builtin_amdgcn_rcp(d);
    r = fma(fma(-d, r, 1.0), r, r);
    r = fma(fma(-d, r, 1.0), r, r);
    return r;
}

__device__ __forceinline__ cxd inv_den(double ek, double zi) {
#pragma clang fp contract(off)
    const double c = 0.0 - ek, inv = fast_rcp(c * c + zi * zi);
    return {c * inv, -(zi * inv)};
}

// the nonzero couplings of every orbital at xt, in k order; ends with a barrier.  Wave 0 holds
// (orbital o, bath k) in lane 16 o + k; a ballot of the nonzero ones gives each its place in its
// orbital's list (the number of nonzero k' < k).
static_assert(kMaxOrb * 16 <= 64 && kFitMaxBath <= 16, "one wave holds every (orbital, bath) pair");
__device__ __forceinline__ void nonzero_couplings(const FitView& q) {
    const int tid = threadIdx.x;
    if (tid < 64) {
        const int o = tid >> 4, k = tid & 15;
        const bool in = o < q.norb && k < q.nbath;
        const double v = in ? S.xt[q.nbath + o * q.nbath + k] : 0.0;
        const bool nz = in && v != 0.0;
        const uint64_t mask = __ballot(nz);
        const uint64_t seg = mask & (0xffffull << (16 * o));
        if (nz) {
            const int n = __popcll(seg & ((1ull << tid) - 1ull));
            S.nzv2[o][n] = v * v;
            S.nzv[o][n] = v;
            S.nzek[o][n] = S.xt[k];
            S.nzk[o][n] = k;
        }
        if (k == 0 && o < q.norb) S.nnz[o] = __popcll(seg);
    }
    __syncthreads();
}

// A thread's group of kGrp frequencies iw = base + tid + j kFitThreads.  The k loop runs outside
// the frequency loop, so V(o,k) and e_k are read and tested once per k and the kGrp divisions of
// one k are independent of each other.  Slots j >= jn (block-uniform) lie past the last frequency
// for every thread: the group helpers are instantiated per slot count JN = jn and skip them.  Their
// terms are exactly +0.0 (zero residual, zero weight), so skipping them leaves every sum
// bit-identical while a short frequency axis (Nw < kGrp kFitThreads) costs only the slots it fills.
struct FreqGroup {
    double zi[kGrp], rw[kGrp];  // Im z, 1/w_n (0 for a frequency past the end)
    int iw[kGrp];
    int jn;                     // slots holding any frequency of the block
};

__device__ __forceinline__ FreqGroup freq_group(const FitView& q, int base) {
    FreqGroup g;
    g.jn = min(kGrp, (q.nw - base + kFitThreads - 1) / kFitThreads);
#pragma unroll
    for (int j = 0; j < kGrp; ++j) {
        const int iw = base + (int)threadIdx.x + j * kFitThreads;
        const bool ok = iw < q.nw;
        const double w = ok ? q.omega[iw] : 1.0;
        g.iw[j] = ok ? iw : -1;
        g.zi[j] = zim(q, w);
        g.rw[j] = ok ? 1.0 / w : 0.0;
    }
    return g;
}

// r_o(w) = Delta_o(i w) - Dev_o(w) for the group (zero past the end)
struct GroupResidual {
    double re[kGrp], im[kGrp];
};
template <int JN>
__device__ __forceinline__ GroupResidual residual_group(const FitView& q, const double* xt, int o,
                                                        const FreqGroup& g) {
    GroupResidual r;
    // the target's loads go out first, so that their latency overlaps the sums over the couplings
    double dre[JN], dim[JN];
#pragma unroll
    for (int j = 0; j < JN; ++j) {
        const size_t io = (size_t)o + (size_t)q.norb * (g.iw[j] >= 0 ? g.iw[j] : 0);
        dre[j] = q.dev[2 * io];
        dim[j] = q.dev[2 * io + 1];
    }
#pragma unroll
    for (int j = 0; j < kGrp; ++j) r.re[j] = r.im[j] = 0.0;
    const int nn = uniform_i32(S.nnz[o]);
    for (int n = 0; n < nn; ++n) {
        const double ek = S.nzek[o][n], v2 = S.nzv2[o][n];
#pragma unroll
        for (int j = 0; j < JN; ++j) {
            const cxd d = inv_den(ek, g.zi[j]);
            r.re[j] = fma(v2, d.re, r.re[j]);
            r.im[j] = fma(v2, d.im, r.im[j]);
        }
    }
    const double ef = xt[q.nx - q.norb + o];
#pragma unroll
    for (int j = 0; j < JN; ++j) {
        const bool ok = g.iw[j] >= 0;
        r.re[j] = ok ? (r.re[j] + ef) - dre[j] : 0.0;
        r.im[j] = ok ? r.im[j] - dim[j] : 0.0;
    }
    return r;
}

// FUNC (func.f90:30-42): sum over (w, o) of |r|^2 / w_n, divided by float(Norb*Norb*Nw).
// Reference order: eval_func_ref.  Tree order: per-thread partials, then block_sum.
template <int JN>
__device__ __forceinline__ double func_group(const FitView& q, const FreqGroup& g, double s) {
#pragma clang fp contract(off)
    for (int o = 0; o < q.norb; ++o) {
        const GroupResidual r = residual_group<JN>(q, S.xt, o, g);
#pragma unroll
        for (int j = 0; j < JN; ++j) s = fma(fma(r.re[j], r.re[j], r.im[j] * r.im[j]), g.rw[j], s);
    }
    return s;
}

__device__ __noinline__ double eval_func_tree(const FitProb qa) {
#pragma clang fp contract(off)
    const FitView q = view(qa);
    const int tid = threadIdx.x;
    double s = 0.0;
    nonzero_couplings(q);
    {
        for (int base = 0; base < q.nw; base += kGrp * kFitThreads) {
            const FreqGroup g = freq_group(q, base);
            switch (g.jn) {
                case 4: s = func_group<4>(q, g, s); break;
                case 3: s = func_group<3>(q, g, s); break;
                case 2: s = func_group<2>(q, g, s); break;
                default: s = func_group<1>(q, g, s); break;
            }
        }
        s = block_sum<kFitThreads>(s, S.red);
    }
    if (tid == 0) S.bcast[0] = s / (double)(float)(q.norb * q.norb * q.nw);
    __syncthreads();
    return S.bcast[0];
}

// Lane 0's sequential adds for ordered_sum, leaving the running sum and the next position in S: the
// leading nseq terms through chain_sum's load ring (a sum that starts small crosses a binade at about
// every doubling of its length, and one lane adds a few hundred terms in the time of a pass), then
// one by one while the sum is zero, subnormal or not finite, or (all) to the end.
__device__ __noinline__ void ordered_sum_seq(double s, int pos, int n, int nseq, bool all) {
#pragma clang fp contract(off)
    if (pos == 0 && nseq > 0) {
        pos = min(nseq, n);
        s = chain_sum(s, S.u.os.T, pos);
    }
    while (pos < n && (all || pos < nseq || !sum_normal(s))) s = s + S.u.os.T[pos++];
    S.os_s = s;
    S.os_pos = pos;
}

// s + T[0] + ... + T[n-1] in this order with one rounding per add (moed_osum.h), formed by the whole
// block: the n <= kFitThreads no terms (no <= 2 kMaxOrb) sit in S.u.os.T, thread j's at j no ..
// j no + no - 1.  Every thread calls it after the barrier that follows the staging, with the same
// s; the first nseq terms go through lane 0 (ordered_sum_seq).  Returns the sum in every thread;
// the terms may be restaged on return.  One pass from position pos: every term's count in s's
// binade (sum_count), a wave scan of the threads' totals (doubles: integers, exact below 2^53), the
// waves' totals, first ties and negative/NaN flags through LDS.  If the whole rest stays in the
// binade with no tie, every thread takes the sum from the total; otherwise the owner of the first
// term that leaves the binade (or of the first tie, if earlier) adds it for real, and the next pass
// starts after it.
__device__ __forceinline__ double ordered_sum(double s, int n, int no, int nseq) {
#pragma clang fp contract(off)
    constexpr int NW = kFitThreads / 64;
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const double lim = (double)kSumLim;
    int pos = 0;
    while (pos < n) {
        if (!sum_normal(s) || pos < nseq) {
            __syncthreads();  // every thread has read the previous step's os_s and os_pos
            if (tid == 0) ordered_sum_seq(s, pos, n, nseq, false);
        } else {
            const int e0 = ilogb(s);
            double kt = 0.0;
            int ft = n;
            bool bad = false;
            for (int o = 0; o < no; ++o) {
                const int i = tid * no + o;
                if (i >= pos && i < n) {
                    const double t = S.u.os.T[i];
                    bad = bad || !(t >= 0.0);
                    bool tie;
                    kt = kt + sum_count(t, e0, tie);
                    if (tie) ft = min(ft, i);
                }
            }
            double incl = kt;
#pragma unroll
            for (int d = 1; d < 64; d <<= 1) {
                const double v = __shfl_up(incl, d, 64);
                if (lane >= d) incl = incl + v;
            }
            double excl = __shfl_up(incl, 1, 64);
            if (lane == 0) excl = 0.0;
#pragma unroll
            for (int d = 32; d > 0; d >>= 1) ft = min(ft, __shfl_xor(ft, d, 64));
            const uint64_t bb = __ballot(bad);
            if (lane == 63) S.os_qk[wid] = incl;
            if (lane == 0) S.os_qf[wid] = bb != 0 ? -1 : ft;
            __syncthreads();
            double woff = 0.0, tot = 0.0;
            int ftie = n;
            bool anybad = false;
#pragma unroll
            for (int w = 0; w < NW; ++w) {
                const double t = S.os_qk[w];
                if (w < wid) woff = woff + t;
                tot = tot + t;
                const int f = S.os_qf[w];
                anybad = anybad || f < 0;
                ftie = min(ftie, f < 0 ? n : f);
            }
            const double M = ldexp(s, 52 - e0), thr = lim - M;
            if (!anybad && ftie >= n && tot <= thr) {
                s = ldexp(M + tot, e0 - 52);
                pos = n;
                continue;  // block-uniform: every thread read the same totals
            }
            if (anybad) {
                if (tid == 0) ordered_sum_seq(s, pos, n, nseq, true);
            } else {
                // The term to add for real is the first past the binade's limit or the first tie,
                // whichever comes first; exactly one thread owns it and knows it: the thread whose
                // prefixes straddle the limit (start <= thr < end; exact where it matters: a prefix
                // at or below thr < 2^53 is an exact integer, and one past it stays past it when
                // rounded) if its crossing comes before the first tie, else the tie's owner, whose
                // prefix before the tie is then still within the limit.
                const double start = woff + excl, end = woff + incl;
                int is = -1;
                if (start <= thr && end > thr) {
                    double c = start;
                    for (int o = 0; o < no; ++o) {
                        const int i = tid * no + o;
                        if (i >= pos && i < n) {
                            bool tie;
                            c = c + sum_count(S.u.os.T[i], e0, tie);
                            if (c > thr) {
                                if (i < ftie) is = i;
                                break;
                            }
                        }
                    }
                }
                if (ftie < n && tid == ftie / no) {
                    double c = start;
                    for (int o = 0; o < no; ++o) {
                        const int i = tid * no + o;
                        if (i >= pos && i < ftie) {
                            bool tie;
                            c = c + sum_count(S.u.os.T[i], e0, tie);
                        }
                    }
                    if (c <= thr) is = ftie;
                }
                if (is >= 0) {
                    double c = start;
                    for (int o = 0; o < no; ++o) {
                        const int i = tid * no + o;
                        if (i >= pos && i < is) {
                            bool tie;
                            c = c + sum_count(S.u.os.T[i], e0, tie);
                        }
                    }
                    S.os_s = ldexp(M + c, e0 - 52) + S.u.os.T[is];
                    S.os_pos = is + 1;
                }
            }
        }
        __syncthreads();
        s = S.os_s;
        pos = uniform_i32(S.os_pos);
    }
    return s;
}

// residual() over the lists nonzero_couplings built: the divisions of four couplings issued
// together, the adds in k order; the same operations as residual(), bit-identical
__device__ __forceinline__ cxd residual_listed(const FitView& q, int o, int iw, double zi) {
#pragma clang fp contract(off)
    double re = 0.0, im = 0.0;
    const int nn = uniform_i32(S.nnz[o]);
    for (int n0 = 0; n0 < nn; n0 += 4) {
        cxd t[4];
#pragma unroll
        for (int u = 0; u < 4; ++u)
            if (n0 + u < nn) t[u] = cx_rdiv(S.nzv2[o][n0 + u], {0.0 - S.nzek[o][n0 + u], zi});
#pragma unroll
        for (int u = 0; u < 4; ++u)
            if (n0 + u < nn) {
                re = re + t[u].re;
                im = im + t[u].im;
            }
    }
    const size_t io = (size_t)o + (size_t)q.norb * iw;
    return {(re + S.xt[q.nx - q.norb + o]) - q.dev[2 * io], im - q.dev[2 * io + 1]};
}

// FUNC in reference order (func.f90:30-42): the terms |r|^2 / w_n in the reference's (w, o) order,
// kFitPairF frequencies at a time, each thread two neighbouring ones (their divisions independent of
// each other, so they overlap), each chunk's added to the running sum by ordered_sum.
__device__ __noinline__ double eval_func_ref(const FitProb qa) {
#pragma clang fp contract(off)
    const FitView q = view(qa);
    const int tid = threadIdx.x;
    double s = 0.0;
    nonzero_couplings(q);
    for (int base = 0; base < q.nw; base += kFitPairF) {
        const int iwa = base + 2 * tid, iwb = iwa + 1;
        if (iwa < q.nw) {
            double* st = S.u.os.T + 2 * tid * q.norb;
            const bool hb = iwb < q.nw;
            const double wa = q.omega[iwa], wb = hb ? q.omega[iwb] : 1.0;
            const double za = zim(q, wa), zb = zim(q, wb);
            for (int o = 0; o < q.norb; ++o) {
                const cxd ra = residual_listed(q, o, iwa, za);
                const cxd rb = residual_listed(q, o, hb ? iwb : iwa, zb);
                const double ha = hypot_corrected(ra.re, ra.im), hbv = hypot_corrected(rb.re, rb.im);
                st[o] = ha * ha / wa;
                if (hb) st[q.norb + o] = hbv * hbv / wb;
            }
        }
        __syncthreads();
        s = ordered_sum(s, min(kFitPairF, q.nw - base) * q.norb, 2 * q.norb, base == 0 ? kSumSeq : 0);
    }
    if (tid == 0) S.bcast[0] = s / (double)(float)(q.norb * q.norb * q.nw);
    __syncthreads();
    return S.bcast[0];
}

// Tree order with the frequencies cached in LDS, filled once per launch; ends with a barrier.
// Mode 1 (Nw <= kFitThreads): Im z, 1/w_n and the target of thread t's frequency iw = t.
// Mode 2 (Nw <= kGrp kFitThreads): Im z and 1/w_n of every slot iw = t + j kFitThreads, with
// freq_group's values for a slot past the end (Im z of w = 1, weight 0); the target stays in HBM.
__device__ void tree_cache_fill(const FitProb qa) {
    const FitView q = view(qa);
    const int t = threadIdx.x;
    const bool tree = uniform_i32(qa.tree);
    const int mode = !tree ? 0 : q.nw <= kFitThreads ? 1 : q.nw <= kGrp * kFitThreads ? 2 : 0;
    if (mode == 1) {
        const bool ok = t < q.nw;
        const double w = ok ? q.omega[t] : 1.0;
        S.u.tc.zi[t] = zim(q, w);
        S.u.tc.rw[t] = ok ? 1.0 / w : 0.0;
        for (int o = 0; o < q.norb; ++o) {
            const size_t io = (size_t)o + (size_t)q.norb * (ok ? t : 0);
            S.u.tc.dre[o][t] = q.dev[2 * io];
            S.u.tc.dim[o][t] = q.dev[2 * io + 1];
        }
    } else if (mode == 2) {
        for (int j = 0; j < kGrp; ++j) {
            const int iw = t + j * kFitThreads;
            const bool ok = iw < q.nw;
            const double w = ok ? q.omega[iw] : 1.0;
            S.u.tw.zi[iw] = zim(q, w);
            S.u.tw.rw[iw] = ok ? 1.0 / w : 0.0;
        }
    }
    if (t == 0) S.tc_mode = mode;
    __syncthreads();
}

// the nonzero-coupling lists of nonzero_couplings for wave 0 only, without the barrier: the caller
// has wave 0 write xt and then call this (LDS operations of one wave complete in order)
__device__ __forceinline__ void nonzero_couplings_wave0(const FitView& q) {
    const int tid = threadIdx.x;
    const int o = tid >> 4, k = tid & 15;
    const bool in = o < q.norb && k < q.nbath;
    const double v = in ? S.xt[q.nbath + o * q.nbath + k] : 0.0;
    const bool nz = in && v != 0.0;
    const uint64_t mask = __ballot(nz);
    const uint64_t seg = mask & (0xffffull << (16 * o));
    if (nz) {
        const int n = __popcll(seg & ((1ull << tid) - 1ull));
        S.nzv2[o][n] = v * v;
        S.nzv[o][n] = v;
        S.nzek[o][n] = S.xt[k];
        S.nzk[o][n] = k;
    }
    if (k == 0 && o < q.norb) S.nnz[o] = __popcll(seg);
}

// FUNC in tree order from the cached frequencies (mode 1) and the coupling lists already built
// for xt; the same operations as eval_func_tree with one frequency per thread, one barrier: every
// thread adds the waves' partials in index order (block_sum's order).  Callers are preceded by a
// barrier, so no wave still reads the partials buffer of an earlier evaluation.
__device__ __forceinline__ double eval_func_tree_cached(const FitProb& qa) {
#pragma clang fp contract(off)
    const FitView q = view(qa);
    const int tid = threadIdx.x;
    double s = 0.0;
    if (tid < q.nw) {
        const double zi = S.u.tc.zi[tid], rw = S.u.tc.rw[tid];
        for (int o = 0; o < q.norb; ++o) {
            double re = 0.0, im = 0.0;
            const int nn = uniform_i32(S.nnz[o]);
            for (int n = 0; n < nn; ++n) {
                const cxd d = inv_den(S.nzek[o][n], zi);
                const double v2 = S.nzv2[o][n];
                re = fma(v2, d.re, re);
                im = fma(v2, d.im, im);
            }
            re = (re + S.xt[q.nx - q.norb + o]) - S.u.tc.dre[o][tid];
            im = im - S.u.tc.dim[o][tid];
            s = fma(fma(re, re, im * im), rw, s);
        }
    }
    s = wave_sum0(s);
    if ((tid & 63) == 0) S.fred[0][tid >> 6] = s;
    __syncthreads();
    double tot = 0.0;
    for (int w = 0; w < kFitThreads / 64; ++w) tot += S.fred[0][w];
    return tot / (double)(float)(q.norb * q.norb * q.nw);
}

__device__ __forceinline__ double eval_func(const FitProb& q) {
    if (uniform_i32(S.tc_mode) == 1) {
        if (threadIdx.x < 64) nonzero_couplings_wave0(view(q));
        __syncthreads();
        return eval_func_tree_cached(q);
    }
    return uniform_i32(q.tree) ? eval_func_tree(q) : eval_func_ref(q);
}

// The couplings that can be nonzero on the line pcom + t xicom (in k order, as nonzero_couplings
// orders them).  A coupling on the list that is exactly zero
// at some t adds fma(0, d, r) = r to each sum (r starts at +0.0 and d is finite), so evaluations
// over this list equal evaluations over the point's own nonzero list bit for bit.
// Run by wave 0 alone (the minimiser wave); the workers read the lists after the next command.
__device__ __forceinline__ void line_couplings_wave0(const FitView& q) {
    const int tid = threadIdx.x;
    const int o = tid >> 4, k = tid & 15;
    const bool in = o < q.norb && k < q.nbath;
    const int j = in ? q.nbath + o * q.nbath + k : 0;
    const bool nz = in && (S.pcom[j] != 0.0 || S.xicom[j] != 0.0);
    const uint64_t mask = __ballot(nz);
    const uint64_t seg = mask & (0xffffull << (16 * o));
    if (nz) S.lk[o][__popcll(seg & ((1ull << tid) - 1ull))] = k;
    if (k == 0 && o < q.norb) S.lnn[o] = __popcll(seg);
}

// mode 2: the terms of thread t's JN frequency slots, func_group's operations and order
template <int JN>
__device__ __forceinline__ double func_point(const FitView& q, const double (*ek)[16], const double (*v2)[16],
                                             const double* ef) {
#pragma clang fp contract(off)
    const int tid = threadIdx.x;
    double zi[JN], rw[JN];
#pragma unroll
    for (int j = 0; j < JN; ++j) {
        zi[j] = S.u.tw.zi[tid + j * kFitThreads];
        rw[j] = S.u.tw.rw[tid + j * kFitThreads];
    }
    double s = 0.0;
    for (int o = 0; o < q.norb; ++o) {
        double dre[JN], dim[JN], re[JN], im[JN];
#pragma unroll
        for (int j = 0; j < JN; ++j) {
            const int iw = tid + j * kFitThreads;
            const size_t io = (size_t)o + (size_t)q.norb * (iw < q.nw ? iw : 0);
            dre[j] = q.dev[2 * io];
            dim[j] = q.dev[2 * io + 1];
            re[j] = im[j] = 0.0;
        }
        const int nn = uniform_i32(S.lnn[o]);
        for (int n = 0; n < nn; ++n) {
            const double e = ek[o][n], w2 = v2[o][n];
#pragma unroll
            for (int j = 0; j < JN; ++j) {
                const cxd d = inv_den(e, zi[j]);
                re[j] = fma(w2, d.re, re[j]);
                im[j] = fma(w2, d.im, im[j]);
            }
        }
        const double f = ef[o];
#pragma unroll
        for (int j = 0; j < JN; ++j) {
            const bool ok = tid + j * kFitThreads < q.nw;
            const double r1 = ok ? (re[j] + f) - dre[j] : 0.0, r2 = ok ? im[j] - dim[j] : 0.0;
            s = fma(fma(r1, r1, r2 * r2), rw[j], s);
        }
    }
    return s;
}

// f1dim's FUNC in tree order with cached frequencies (mode 2) at pcom + xv xicom.  Every wave forms the
// point's couplings itself in its own LDS slot (lane 16 o + n: coupling n of orbital o; lane
// 16 o + 15: ef_o), with f1dim's arithmetic, so no block barrier precedes the sums; the one barrier
// is the waves' partials, in buffer `buf`.  Consecutive calls alternate buffers: a wave can reach
// the next evaluation's write only after every wave has left this one's barrier, and it can reach
// the write after that only after every wave has read this one's partials.  Equal bit for bit to
// writing xt and calling eval_func.
__device__ __forceinline__ void form_point(const FitView& q, double xv) {
#pragma clang fp contract(off)
    const int lane = threadIdx.x & 63;
    auto& P = S.wp[threadIdx.x >> 6];
    const int o = lane >> 4, n = lane & 15;
    if (o < q.norb) {
        if (n < S.lnn[o]) {
            const int k = S.lk[o][n], j = q.nbath + o * q.nbath + k;
            const double v = S.pcom[j] + xv * S.xicom[j];
            P.ek[o][n] = S.pcom[k] + xv * S.xicom[k];
            P.v2[o][n] = v * v;
        } else if (n == 15) {
            const int j = q.nx - q.norb + o;
            P.ef[o] = S.pcom[j] + xv * S.xicom[j];
        }
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// the waves' partials of one evaluation, summed in block_sum's order by every thread
__device__ __forceinline__ double point_total(const FitView& q, double s, int buf) {
#pragma clang fp contract(off)
    s = wave_sum0(s);
    if ((threadIdx.x & 63) == 0) S.fred[buf][threadIdx.x >> 6] = s;
    __syncthreads();
    double tot = 0.0;
    for (int w = 0; w < kFitThreads / 64; ++w) tot += S.fred[buf][w];
    return tot / (double)(float)(q.norb * q.norb * q.nw);
}

// mode 2: up to kGrp frequencies per thread (eval_func_tree's operations); out of line, so the
// line search keeps one copy of it
__device__ __noinline__ double f1dim_tree_wide(const FitProb qa, double xv, int buf) {
    const FitView q = view(qa);
    form_point(q, xv);
    const auto& P = S.wp[threadIdx.x >> 6];
    double s;
    switch ((q.nw + kFitThreads - 1) / kFitThreads) {
        case 4: s = func_point<4>(q, P.ek, P.v2, P.ef); break;
        case 3: s = func_point<3>(q, P.ek, P.v2, P.ef); break;
        default: s = func_point<2>(q, P.ek, P.v2, P.ef); break;
    }
    return point_total(q, s, buf);
}

// gradient terms of one (w, o) (func.f90:89-124): column k = d/de_k, nbath + k = d/dV(o,k),
// 2 nbath = d/def_o.  The k loop is unrolled to kFitMaxBath so that a caller may keep its
// per-column partial sums in registers.
template <class Put>
__device__ __forceinline__ void grad_terms(const FitView& q, const double* xt, int o, int iw, Put put) {
#pragma clang fp contract(off)
    const double w = q.omega[iw], zi = zim(q, w);
    const cxd r = residual(q, xt, o, iw, zi);
    const cxd cr = {r.re, -r.im};
    const int nb = q.nbath;
#pragma unroll
    for (int k = 0; k < kFitMaxBath; ++k) {
        if (k < nb) {
            const double v = xt[nb + o * nb + k];
            if (v != 0.0) {
                const cxd den = {0.0 - xt[k], zi};
                const cxd t = cx_div(cx_div(cx_mul(cx_mul(cr, {v, 0.0}), {v, 0.0}), den), den);
                put(k, 2.0 * t.re / w);
                const cxd u = cx_div(cx_mul(cr, {v, 0.0}), den);
                const double tre = u.re + u.re;
                put(kFitMaxBath + k, 2.0 * tre / w);
            }
        }
    }
    put(2 * kFitMaxBath, 2.0 * r.re / w);
}

// one (orbital, frequency group) of the tree-order dfunc: per-wave column sums into S.gred
template <int JN>
__device__ __forceinline__ void grad_group(const FitView& q, int o, const FreqGroup& g, int lane, int wid) {
#pragma clang fp contract(off)
    const GroupResidual r = residual_group<JN>(q, S.xt, o, g);
    const int nn = uniform_i32(S.nnz[o]);
    for (int n = 0; n < nn; ++n) {
        const int k = S.nzk[o][n];
        const double ek = S.nzek[o][n], v2 = S.nzv2[o][n], v = S.nzv[o][n];
        double pde = 0.0, pdv = 0.0;
#pragma unroll
        for (int j = 0; j < JN; ++j) {
            const cxd d = inv_den(ek, g.zi[j]);
            const double ure = fma(r.re[j], d.re, r.im[j] * d.im);
            const double uim = fma(r.re[j], d.im, -(r.im[j] * d.re));
            const double rw2 = 2.0 * g.rw[j];
            pde = fma(v2 * fma(ure, d.re, -(uim * d.im)), rw2, pde);
            pdv = fma(2.0 * v * ure, rw2, pdv);
        }
        pde = wave_sum0(pde);
        pdv = wave_sum0(pdv);
        if (lane == 0) {
            S.gred[wid][k] = pde;
            S.gred[wid][kFitMaxBath + k] = pdv;
        }
    }
    double pdef = 0.0;
#pragma unroll
    for (int j = 0; j < JN; ++j) pdef = fma(r.re[j], 2.0 * g.rw[j], pdef);
    pdef = wave_sum0(pdef);
    if (lane == 0) S.gred[wid][2 * kFitMaxBath] = pdef;
}

// dfunc (func.f90:47-142) at xt into out[0:nx] (LDS).  Reference order: d/de_k runs over o, then w
// (one chain through all orbital passes); d/dV(o,k) and d/def_o over w.
__device__ __noinline__ void eval_grad_tree(const FitProb qa, double* out) {
#pragma clang fp contract(off)
    const FitView q = view(qa);
    const int tid = threadIdx.x, nb = q.nbath, lane = tid & 63, wid = tid >> 6;
    const double scale1 = (double)(float)(q.norb * q.norb), scale2 = (double)(float)q.nw;
    for (int c = tid; c < q.nx; c += kFitThreads) S.gacc[c] = 0.0;
    nonzero_couplings(q);
    // with d = 1/(i w - e_k): d/de_k 2 Re(conj(r) V^2 d^2)/w, d/dV 4 Re(conj(r) V d)/w, d/def 2 Re(r)/w.
    // Per (orbital, frequency group): each column is summed over the thread's frequencies, over the
    // wave, then over the waves in index order, and added to the running total.
    for (int o = 0; o < q.norb; ++o) {
        for (int base = 0; base < q.nw; base += kGrp * kFitThreads) {
            const FreqGroup g = freq_group(q, base);
            switch (g.jn) {
                case 4: grad_group<4>(q, o, g, lane, wid); break;
                case 3: grad_group<3>(q, o, g, lane, wid); break;
                case 2: grad_group<2>(q, o, g, lane, wid); break;
                default: grad_group<1>(q, o, g, lane, wid); break;
            }
            __syncthreads();
            if (tid < kFitCols) {
                const int c = tid, k = c < kFitMaxBath ? c : c - kFitMaxBath;
                if (c == 2 * kFitMaxBath || (k < nb && S.xt[nb + o * nb + k] != 0.0)) {
                    double v = 0.0;
                    for (int w = 0; w < kFitThreads / 64; ++w) v = v + S.gred[w][c];
                    const int comp = c < kFitMaxBath ? k : (c < 2 * kFitMaxBath ? nb + o * nb + k : q.nx - q.norb + o);
                    S.gacc[comp] = S.gacc[comp] + v;
                }
            }
            __syncthreads();
        }
    }
    for (int c = tid; c < q.nx; c += kFitThreads) out[c] = S.gacc[c] / scale1 / scale2;
    __syncthreads();
}

__device__ __noinline__ void eval_grad_ref(const FitProb qa, double* out) {
#pragma clang fp contract(off)
    const FitView q = view(qa);
    const int tid = threadIdx.x, nb = q.nbath;
    const double scale1 = (double)(float)(q.norb * q.norb), scale2 = (double)(float)q.nw;
    {
        double s = 0.0;  // lane c < nx of wave 0: running sum of component c
        for (int o = 0; o < q.norb; ++o) {
            // staged column of this lane's component in this pass (-1: none), and whether it has terms
            int col = -1;
            if (tid < nb)
                col = tid;
            else if (tid >= nb + o * nb && tid < nb + (o + 1) * nb)
                col = kFitMaxBath + (tid - nb - o * nb);
            else if (tid == q.nx - q.norb + o)
                col = 2 * kFitMaxBath;
            const int k = col < kFitMaxBath ? col : col - kFitMaxBath;
            const bool on = col >= 0 && (col == 2 * kFitMaxBath || S.xt[nb + o * nb + k] != 0.0);
            for (int base = 0; base < q.nw; base += kFitChunkG) {
                const int iw = base + tid;
                if (tid < kFitChunkG && iw < q.nw)
                    grad_terms(q, S.xt, o, iw, [&](int c, double t) { S.u.cols[c * kFitChunkG + tid] = t; });
                __syncthreads();
                if (on) s = chain_sum(s, S.u.cols + col * kFitChunkG, min(kFitChunkG, q.nw - base));
                __syncthreads();
            }
        }
        if (tid < q.nx) out[tid] = s / scale1 / scale2;
    }
    __syncthreads();
}

__device__ __forceinline__ void eval_grad(const FitProb& q, double* out) {
    if (uniform_i32(q.tree))
        eval_grad_tree(q, out);
    else
        eval_grad_ref(q, out);
}

// ---- the minimiser wave and its workers
//
// Wave 0 runs frprmn / linmin / mnbrak / brent alone; the other waves only serve its block-wide
// operations (a FUNC evaluation, a gradient).  For each one wave 0 writes the command (and its
// arguments) to LDS and meets the workers at one barrier, after which every wave runs run_op with
// the same command, so all waves pass the same barriers inside it.  A worker reads the command
// before the first barrier inside the operation and wave 0 posts the next one only after that
// barrier.  kCmdStop ends the workers' loop; the kernel posts it on every path out of the
// minimiser.  The line-search logic, once repeated by all 16 waves (issue slots on every SIMD),
// now runs in one.
enum { kCmdStop = 0, kCmdFunc = 1, kCmdFuncListed = 2, kCmdFuncWide = 3, kCmdGrad = 4 };

__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// one block-wide operation; every wave of the block calls it with the same command
__device__ __forceinline__ double run_op(const FitProb& q, int cmd) {
    switch (cmd) {
        case kCmdFuncListed:  // FUNC at xt with the coupling lists wave 0 built (cached frequencies)
            return eval_func_tree_cached(q);
        case kCmdFuncWide:  // FUNC at pcom + cxv xicom, frequencies cached kGrp per thread
            return f1dim_tree_wide(q, S.cxv, S.cbuf);
        case kCmdFunc:  // FUNC at xt
            return eval_func(q);
        default:  // kCmdGrad: dfunc at xt into xi
            eval_grad(q, S.xi);
            return 0.0;
    }
}

// wave 0: post a command, meet the workers, run it
__device__ __forceinline__ double command(const FitProb& q, int cmd) {
    if (threadIdx.x == 0) S.cmd = cmd;
    __syncthreads();
    return cmd == kCmdStop ? 0.0 : run_op(q, cmd);
}

// waves 1..: serve commands until kCmdStop
__device__ void serve(const FitProb& q) {
    for (;;) {
        __syncthreads();
        const int cmd = uniform_i32(S.cmd);
        if (cmd == kCmdStop) return;
        (void)run_op(q, cmd);
    }
}

// f1dim (frprmn.f90:212-234), run by wave 0: the point, then one command
__device__ __forceinline__ double f1dim(const FitProb& q, double xv, int& nfev) {
#pragma clang fp contract(off)
    ++nfev;
    const int mode = uniform_i32(S.tc_mode);
    const int lane = threadIdx.x;
    // up to kGrp frequencies per thread: each wave forms the point itself (a non-finite xv takes
    // the path below, where every coupling is formed from xt)
    if (mode == 2 && isfinite(xv)) {
        if (lane == 0) {
            S.cxv = xv;
            S.cbuf = nfev & 1;
        }
        return command(q, kCmdFuncWide);
    }
    if (lane < q.nx) S.xt[lane] = S.pcom[lane] + xv * S.xicom[lane];
    wave_sync();
    if (mode == 1) {  // one frequency per thread: wave 0 builds the point's coupling lists
        nonzero_couplings_wave0(view(q));
        return command(q, kCmdFuncListed);
    }
    return command(q, kCmdFunc);
}

__device__ __forceinline__ double fsign(double a, double b) { return b >= 0.0 ? fabs(a) : -fabs(a); }

// mnbrak (frprmn.f90:236-310).  The reference's bracketing loop has no count limit; on a flat or
// unbounded direction its steps grow geometrically until f turns NaN and the loop ends.  The cap
// below is far beyond that and only turns a pathological input into an error instead of an endless
// kernel; false when it is reached.
constexpr int kBracketCap = 100000;
// f0 is FUNC at ax = 0, the line's start pcom, which the caller already holds: pcom is the last
// line search's minimum pcom' + xmin xicom', formed with the roundings f1dim used for its value
// there, and pcom + 0 xicom rounds back to pcom (only a -0.0 entry may turn +0.0, which FUNC only
// squares or subtracts from 0.0).  The reference evaluates it again; here it is counted, not
// recomputed, with the same result.
__device__ bool mnbrak(const FitProb q, double& ax, double& bx, double& cx, double& fa, double& fb,
                       double& fc, int& nfev, double f0) {
#pragma clang fp contract(off)
    const double GOLD = (double)1.618034f, GLIMIT = 100.0, TINY = (double)1.e-20f;
    double dum, fu, q_, r, u, ulim;
    ++nfev;  // fa = f1dim(ax = 0)
    fa = f0;
    fb = f1dim(q, bx, nfev);
    if (fb > fa) {
        dum = ax;
        ax = bx;
        bx = dum;
        dum = fb;
        fb = fa;
        fa = dum;
    }
    cx = bx + GOLD * (bx - ax);
    fc = f1dim(q, cx, nfev);
    for (int guard = 0; fb >= fc; ++guard) {
        if (guard == kBracketCap) return false;
        r = (bx - ax) * (fb - fc);
        q_ = (bx - cx) * (fb - fa);
        u = bx - ((bx - cx) * q_ - (bx - ax) * r) / (2. * fsign(fmax(fabs(q_ - r), TINY), q_ - r));
        ulim = bx + GLIMIT * (cx - bx);
        if ((bx - u) * (u - cx) > 0.) {
            fu = f1dim(q, u, nfev);
            if (fu < fc) {
                ax = bx;
                fa = fb;
                bx = u;
                fb = fu;
                return true;
            } else if (fu > fb) {
                cx = u;
                fc = fu;
                return true;
            }
            u = cx + GOLD * (cx - bx);
            fu = f1dim(q, u, nfev);
        } else if ((cx - u) * (u - ulim) > 0.) {
            fu = f1dim(q, u, nfev);
            if (fu < fc) {
                bx = cx;
                cx = u;
                u = cx + GOLD * (cx - bx);
                fb = fc;
                fc = fu;
                fu = f1dim(q, u, nfev);
            }
        } else if ((u - ulim) * (ulim - cx) >= 0.) {
            u = ulim;
            fu = f1dim(q, u, nfev);
        } else {
            u = cx + GOLD * (cx - bx);
            fu = f1dim(q, u, nfev);
        }
        ax = bx;
        bx = cx;
        cx = u;
        fa = fb;
        fb = fc;
        fc = fu;
    }
    return true;
}

// brent (frprmn.f90:119-210); false when its iteration limit is reached.  fbx is FUNC at bx, which
// mnbrak returned with it: mnbrak moves every abscissa together with its value, so fbx is the value
// f1dim gave at this exact bx.  The reference evaluates it again; here it is counted, not recomputed.
__device__ bool brent(const FitProb q, double ax, double bx, double cx, double tol, double& xmin,
                      double& fret, int& nfev, double fbx) {
#pragma clang fp contract(off)
    const double CGOLD = (double).3819660f, ZEPS = (double)1.0e-10f;
    double a = fmin(ax, cx), b = fmax(ax, cx);
    double d = 0.0, e = 0., etemp, fu, fv, fw, fx, p, q_, r, tol1, tol2, u, v, w, x, xm;
    v = bx;
    w = v;
    x = v;
    ++nfev;  // fx = f1dim(x = bx)
    fx = fbx;
    fv = fx;
    fw = fx;
    for (int iter = 1; iter <= kBrentItmax; ++iter) {
        xm = 0.5 * (a + b);
        tol1 = tol * fabs(x) + ZEPS;
        tol2 = 2. * tol1;
        if (fabs(x - xm) <= (tol2 - .5 * (b - a))) {
            xmin = x;
            fret = fx;
            return true;
        }
        bool golden = true;
        if (fabs(e) > tol1) {
            r = (x - w) * (fx - fv);
            q_ = (x - v) * (fx - fw);
            p = (x - v) * q_ - (x - w) * r;
            q_ = 2. * (q_ - r);
            if (q_ > 0.) p = -p;
            q_ = fabs(q_);
            etemp = e;
            e = d;
            if (!(fabs(p) >= fabs(.5 * q_ * etemp) || p <= q_ * (a - x) || p >= q_ * (b - x))) {
                d = p / q_;
                u = x + d;
                if (u - a < tol2 || b - u < tol2) d = fsign(tol1, xm - x);
                golden = false;
            }
        }
        if (golden) {
            e = (x >= xm) ? a - x : b - x;
            d = CGOLD * e;
        }
        u = (fabs(d) >= tol1) ? x + d : x + fsign(tol1, d);
        fu = f1dim(q, u, nfev);
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
    return false;
}

// frprmn (frprmn.f90:1-84) with linmin (:86-117) inlined, run by wave 0 of fit blockIdx.x; every
// path returns here, where the kernel stops the workers.  LDS written by wave 0 reaches the
// workers through the barrier of its next command.
__device__ __forceinline__ void minimise(const FitCgArgs& a, int its0, int nfev, double fp, double fret) {
#pragma clang fp contract(off)
    const FitProb q = a.q;
    const int f = blockIdx.x, lane = threadIdx.x, n = q.nx;  // n <= kFitNx = 64: one lane per entry
    int32_t* ist = a.ist + (size_t)f * kStateI;
    double* st = a.st + (size_t)f * kStateD;
    double* px = a.x + (size_t)f * n;
    const double EPS = 1.e-10, LTOL = (double)1.e-8f;  // frprmn.f90:12, :91
    if (uniform_i32(ist[0]) == kPhaseStart) {
        fp = command(q, kCmdFunc);  // FUNC at xt = the start point
        ++nfev;
        command(q, kCmdGrad);
        if (lane < n) {
            S.g[lane] = -S.xi[lane];
            S.h[lane] = S.g[lane];
            S.xi[lane] = S.h[lane];
        }
        wave_sync();
        its0 = 0;
    }
    int rc = 1, iter = its0;
    for (int its = its0 + 1; its <= kItmax; ++its) {
        iter = its;
        // ---- linmin
        if (lane < n) {
            S.pcom[lane] = S.p[lane];
            S.xicom[lane] = S.xi[lane];
        }
        wave_sync();
        if (uniform_i32(S.tc_mode) == 2) {
            line_couplings_wave0(view(q));
            wave_sync();
        }
        double ax = 0., xx = 1., bx, fa, fx, fb, xmin = 0.0, fr = 0.0;
        if (!mnbrak(q, ax, xx, bx, fa, fx, fb, nfev, fp) || !brent(q, ax, xx, bx, LTOL, xmin, fr, nfev, fx)) {
            rc = MOED_ECONV;  // 'brent exceed maximum iterations'
            break;
        }
        fret = fr;
        if (lane < n) {
            S.xi[lane] = xmin * S.xi[lane];
            S.p[lane] = S.p[lane] + S.xi[lane];
        }
        wave_sync();
        // ---- end linmin
        if (2.0 * fabs(fret - fp) <= a.ftol * (fabs(fret) + fabs(fp) + EPS)) {
            rc = MOED_OK;
            break;
        }
        fp = fret;
        if (lane < n) S.xt[lane] = S.p[lane];
        wave_sync();
        command(q, kCmdGrad);
        double gg = 0.0, dgg = 0.0;
        for (int j = 0; j < n; ++j) {
            gg = gg + S.g[j] * S.g[j];
            dgg = dgg + (S.xi[j] + S.g[j]) * S.xi[j];
        }
        if (gg == 0.0) {
            rc = MOED_OK;
            break;
        }
        const double gam = dgg / gg;
        wave_sync();
        if (lane < n) {
            S.g[lane] = -S.xi[lane];
            S.h[lane] = S.g[lane] + gam * S.h[lane];
            S.xi[lane] = S.h[lane];
        }
        wave_sync();
        if (its - its0 >= a.budget && its < kItmax) {  // suspend at the iteration boundary
            if (lane < n) {
                px[lane] = S.p[lane];
                st[lane] = S.xi[lane];
                st[kFitNx + lane] = S.g[lane];
                st[2 * kFitNx + lane] = S.h[lane];
            }
            if (lane == 0) {
                st[3 * kFitNx] = fp;
                st[3 * kFitNx + 1] = fret;
                ist[0] = kPhaseRun;
                ist[1] = its;
                ist[3] = nfev;
                atomicAdd(a.active, 1);
            }
            return;
        }
    }
    if (rc == 1) rc = MOED_ECONV;  // 'frprmn maximum iterations exceeded!'
    if (lane < n) px[lane] = S.p[lane];
    if (lane == 0) {
        st[3 * kFitNx + 1] = fret;
        ist[0] = kPhaseDone;
        ist[1] = iter;
        ist[2] = rc;
        ist[3] = nfev;
    }
}

__global__ void __launch_bounds__(kFitThreads) fit_cg_kernel(FitCgArgs a) {
    const FitProb q = a.q;
    const int f = blockIdx.x, tid = threadIdx.x, n = q.nx;
    const int32_t* ist = a.ist + (size_t)f * kStateI;
    const double* st = a.st + (size_t)f * kStateD;
    const double* px = a.x + (size_t)f * n;
    if (uniform_i32(ist[0]) == kPhaseDone) return;
    tree_cache_fill(q);
    const int its0 = uniform_i32(ist[1]), nfev = uniform_i32(ist[3]);
    const double fp = st[3 * kFitNx], fret = st[3 * kFitNx + 1];
    for (int j = tid; j < n; j += kFitThreads) {
        S.p[j] = px[j];
        S.xi[j] = st[j];
        S.g[j] = st[kFitNx + j];
        S.h[j] = st[2 * kFitNx + j];
        S.xt[j] = px[j];
    }
    __syncthreads();
    if (uniform_i32(tid >> 6) != 0) {  // waves 1..: the workers
        serve(q);
        return;
    }
    minimise(a, its0, nfev, fp, fret);
    (void)command(q, kCmdStop);
}

// FUNC and/or dfunc at x[blockIdx.x]
__global__ void __launch_bounds__(kFitThreads) fit_eval_kernel(FitEvalArgs a) {
    const int f = blockIdx.x, n = a.q.nx;
    tree_cache_fill(a.q);
    for (int j = threadIdx.x; j < n; j += kFitThreads) S.xt[j] = a.x[(size_t)f * n + j];
    __syncthreads();
    if (a.f) {
        const double v = eval_func(a.q);
        if (threadIdx.x == 0) a.f[f] = v;
    }
    if (a.g) {
        eval_grad(a.q, S.g);
        for (int j = threadIdx.x; j < n; j += kFitThreads) a.g[(size_t)f * n + j] = S.g[j];
    }
}

// ordered_sum over n terms from s0, kFitThreads no of them at a time (nseq one by one at the start):
// the block-wide replay checked against the sequential loop on chosen inputs
__global__ void __launch_bounds__(kFitThreads) ordered_sum_kernel(const double* T, int n, int no, int nseq,
                                                                   double s0, double* out) {
    double s = s0;
    const int chunk = kFitThreads * no;
    for (int base = 0; base < n; base += chunk) {
        const int m = min(chunk, n - base);
        for (int i = threadIdx.x; i < m; i += kFitThreads) S.u.os.T[i] = T[base + i];
        __syncthreads();
        s = ordered_sum(s, m, no, base == 0 ? nseq : 0);
    }
    if (threadIdx.x == 0) out[0] = s;
}

struct Buf {
    void* p = nullptr;
    ~Buf() {
        if (p) (void)hipFree(p);
    }
    hipError_t alloc(size_t b) { return hipMalloc(&p, std::max<size_t>(b, 256)); }
    template <class T>
    T* as() const {
        return reinterpret_cast<T*>(p);
    }
};

int hfail(const FitCtx& c, hipError_t e, const char* what) {
    if (c.err) *c.err = std::string(what) + ": " + hipGetErrorString(e);
    return e == hipErrorOutOfMemory ? MOED_ENOMEM : MOED_EHIP;
}
#define FCHK(expr)                                        \
    do {                                                  \
        hipError_t _e = (expr);                           \
        if (_e != hipSuccess) return hfail(c, _e, #expr); \
    } while (0)

int argfail(const FitCtx& c, const char* what) {
    if (c.err) *c.err = what;
    return MOED_EARG;
}

bool finite_all(const double* v, size_t n) {
    for (size_t i = 0; i < n; ++i)
        if (!std::isfinite(v[i])) return false;
    return true;
}

// shape and finiteness checks shared by the entry points (skipping zero couplings relies on
// finite inputs)
int check_problem(const FitCtx& c, int nfit, int nx, const double* x, int nw, const double* omega,
                  const double* Dev) {
    if (nx > kFitNx || c.nbath > kFitMaxBath || c.norb > kMaxOrb) return argfail(c, "fit: nx or Nbath too large");
    if (!finite_all(omega, nw) || !finite_all(Dev, (size_t)2 * c.norb * nw))
        return argfail(c, "fit: omega and Dev must be finite");
    for (int i = 0; i < nw; ++i)
        if (!(omega[i] > 0.0)) return argfail(c, "fit: omega must be positive");
    if (x && !finite_all(x, (size_t)nfit * nx)) return argfail(c, "fit: x must be finite");
    return MOED_OK;
}

FitProb make_prob(const FitCtx& c, int nx, int nw, const double* om, const double* dev) {
    return FitProb{nx, c.norb, c.nbath, nw, c.compat, c.tree, om, dev};
}

}  // namespace

int fit_eval(const FitCtx& c, int nfit, int nx, const double* x, int nw, const double* omega, const double* Dev,
             double* f, double* g) {
    int rc = check_problem(c, nfit, nx, x, nw, omega, Dev);
    if (rc) return rc;
    Buf dx, dom, ddev, df, dg;
    FCHK(dx.alloc((size_t)nfit * nx * sizeof(double)));
    FCHK(dom.alloc(nw * sizeof(double)));
    FCHK(ddev.alloc((size_t)2 * c.norb * nw * sizeof(double)));
    FCHK(df.alloc(nfit * sizeof(double)));
    FCHK(dg.alloc((size_t)nfit * nx * sizeof(double)));
    FCHK(hipMemcpyAsync(dx.p, x, (size_t)nfit * nx * sizeof(double), hipMemcpyHostToDevice, c.stream));
    FCHK(hipMemcpyAsync(dom.p, omega, nw * sizeof(double), hipMemcpyHostToDevice, c.stream));
    FCHK(hipMemcpyAsync(ddev.p, Dev, (size_t)2 * c.norb * nw * sizeof(double), hipMemcpyHostToDevice, c.stream));
    FitEvalArgs a{make_prob(c, nx, nw, dom.as<double>(), ddev.as<double>()), dx.as<double>(),
                  f ? df.as<double>() : nullptr, g ? dg.as<double>() : nullptr};
    hipLaunchKernelGGL(fit_eval_kernel, dim3(nfit), dim3(kFitThreads), 0, c.stream, a);
    FCHK(hipGetLastError());
    if (f) FCHK(hipMemcpyAsync(f, df.p, nfit * sizeof(double), hipMemcpyDeviceToHost, c.stream));
    if (g) FCHK(hipMemcpyAsync(g, dg.p, (size_t)nfit * nx * sizeof(double), hipMemcpyDeviceToHost, c.stream));
    FCHK(hipStreamSynchronize(c.stream));
    return MOED_OK;
}

int fit_delta(const FitCtx& c, int nx, const double* x, int nw, const double* omega, double* Dev) {
    // the residual pass with a zero target returns Delta itself
    Buf dx, dom, dzero, dout;
    FCHK(dx.alloc(nx * sizeof(double)));
    FCHK(dom.alloc(nw * sizeof(double)));
    FCHK(dzero.alloc((size_t)2 * c.norb * nw * sizeof(double)));
    FCHK(dout.alloc((size_t)2 * c.norb * nw * sizeof(double)));
    FCHK(hipMemcpyAsync(dx.p, x, nx * sizeof(double), hipMemcpyHostToDevice, c.stream));
    FCHK(hipMemcpyAsync(dom.p, omega, nw * sizeof(double), hipMemcpyHostToDevice, c.stream));
    FCHK(hipMemsetAsync(dzero.p, 0, (size_t)2 * c.norb * nw * sizeof(double), c.stream));
    FitArgs a{dx.as<double>(), nx, c.norb, c.nbath, dom.as<double>(), dzero.as<double>(), nw, c.compat,
              dout.as<double>()};
    FCHK(launch_fit_residual(a, c.stream));
    FCHK(hipMemcpyAsync(Dev, dout.p, (size_t)2 * c.norb * nw * sizeof(double), hipMemcpyDeviceToHost, c.stream));
    FCHK(hipStreamSynchronize(c.stream));
    return MOED_OK;
}

namespace {

__global__ void fit_zero_i32_kernel(int32_t* p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

hipError_t zero_i32(int32_t* p, int n, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(fit_zero_i32_kernel, dim3((n + 255) / 256), dim3(256), 0, s, p, n);
    return hipGetLastError();
}

// The CG launches of one minimisation over device buffers: st / ist hold each fit's state between
// launches (ist zero = not started), act counts the fits a launch leaves running.  Every launch
// ends each running fit after `budget` more iterations; kItmax bounds the loop.
int fit_run(const FitCtx& c, int nfit, int nx, double* dx, int nw, const double* dom, const double* ddev,
            double ftol, double* st, int32_t* ist, int32_t* act) {
    FCHK(zero_i32(reinterpret_cast<int32_t*>(st), nfit * kStateD * 2, c.stream));
    FCHK(zero_i32(ist, nfit * kStateI, c.stream));
    const int budget = c.budget > 0 ? c.budget : (c.tree ? 1024 : 128);
    FitCgArgs a{make_prob(c, nx, nw, dom, ddev), nfit, budget, ftol, dx, st, ist, act};
    for (int launch = 0; launch <= kItmax; ++launch) {
        FCHK(zero_i32(act, 1, c.stream));
        hipLaunchKernelGGL(fit_cg_kernel, dim3(nfit), dim3(kFitThreads), 0, c.stream, a);
        FCHK(hipGetLastError());
        int32_t active = 0;
        FCHK(hipMemcpyAsync(&active, act, sizeof(int32_t), hipMemcpyDeviceToHost, c.stream));
        FCHK(hipStreamSynchronize(c.stream));
        if (active == 0) break;
    }
    return MOED_OK;
}

}  // namespace

int fit_minimize(const FitCtx& c, int nfit, int nx, double* x, int nw, const double* omega, const double* Dev,
                 double ftol, double* fret, int* iter, int* status) {
    int rc = check_problem(c, nfit, nx, x, nw, omega, Dev);
    if (rc) return rc;
    Buf dx, dom, ddev, dst, dist, dact;
    FCHK(dx.alloc((size_t)nfit * nx * sizeof(double)));
    FCHK(dom.alloc(nw * sizeof(double)));
    FCHK(ddev.alloc((size_t)2 * c.norb * nw * sizeof(double)));
    FCHK(dst.alloc((size_t)nfit * kStateD * sizeof(double)));
    FCHK(dist.alloc((size_t)nfit * kStateI * sizeof(int32_t)));
    FCHK(dact.alloc(sizeof(int32_t)));
    FCHK(hipMemcpyAsync(dx.p, x, (size_t)nfit * nx * sizeof(double), hipMemcpyHostToDevice, c.stream));
    FCHK(hipMemcpyAsync(dom.p, omega, nw * sizeof(double), hipMemcpyHostToDevice, c.stream));
    FCHK(hipMemcpyAsync(ddev.p, Dev, (size_t)2 * c.norb * nw * sizeof(double), hipMemcpyHostToDevice, c.stream));
    rc = fit_run(c, nfit, nx, dx.as<double>(), nw, dom.as<double>(), ddev.as<double>(), ftol, dst.as<double>(),
                 dist.as<int32_t>(), dact.as<int32_t>());
    if (rc) return rc;
    std::vector<double> stv((size_t)nfit * kStateD);
    std::vector<int32_t> isv((size_t)nfit * kStateI);
    FCHK(hipMemcpyAsync(x, dx.p, (size_t)nfit * nx * sizeof(double), hipMemcpyDeviceToHost, c.stream));
    FCHK(hipMemcpyAsync(stv.data(), dst.p, stv.size() * sizeof(double), hipMemcpyDeviceToHost, c.stream));
    FCHK(hipMemcpyAsync(isv.data(), dist.p, isv.size() * sizeof(int32_t), hipMemcpyDeviceToHost, c.stream));
    FCHK(hipStreamSynchronize(c.stream));
    for (int i = 0; i < nfit; ++i) {
        const int32_t* s = isv.data() + (size_t)i * kStateI;
        if (s[0] != kPhaseDone) {
            if (c.err) *c.err = "fit: a fit did not finish";
            return MOED_ECONV;
        }
        if (fret) fret[i] = stv[(size_t)i * kStateD + 3 * kFitNx + 1];
        if (iter) iter[i] = s[1];
        if (status) status[i] = s[2];
        if (c.nfev) c.nfev[i] = s[3];
    }
    return MOED_OK;
}

// ------------------------------------------------------------------ allocation-free minimisation
//
// The CG state of up to kNoAllocFits fits lives in one file-scope device array, so that a caller
// who cannot allocate device memory (code under stream capture, serving loops) can still run the
// minimiser on its own device buffers.  Every GPU holds its own copy; its address is looked up once
// per device (fit_scratch_address, called when a handle is created) and only that address is
// handed to kernels and copies.  The array keeps the last call's contents, so each call clears the
// parts it reads before use.  Calls that overlap on one GPU share the array and race.

namespace {

struct NoAllocScratch {
    double st[(size_t)kNoAllocFits * kStateD];
    int32_t ist[(size_t)kNoAllocFits * kStateI];
    int32_t flag[64];  // [0] fits a launch leaves running, [1] input check bits, [2] unfinished fits
};

static __device__ __align__(256) NoAllocScratch scratch;  // ~1.6 MB per GPU

// input checks of check_problem on device data: bit 0 omega not positive or not finite, bit 1 Dev
// not finite, bit 2 x not finite
__global__ void fit_check_kernel(const double* omega, int nw, const double* dev, int ndev, const double* x, int nxall,
                                 int32_t* flag) {
    const int stride = gridDim.x * blockDim.x;
    int bad = 0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nw; i += stride)
        if (!(omega[i] > 0.0) || !isfinite(omega[i])) bad |= 1;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < ndev; i += stride)
        if (!isfinite(dev[i])) bad |= 2;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nxall; i += stride)
        if (!isfinite(x[i])) bad |= 4;
    if (bad) atomicOr(flag, bad);
}

// per fit: fret, iterations, status; unfinished fits are counted into *unfinished
__global__ void fit_results_kernel(const double* st, const int32_t* ist, int nfit, double* fret, int32_t* iter,
                                   int32_t* status, int32_t* unfinished) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nfit) return;
    const int32_t* s = ist + (size_t)i * kStateI;
    if (s[0] != kPhaseDone) atomicAdd(unfinished, 1);
    if (fret) fret[i] = st[(size_t)i * kStateD + 3 * kFitNx + 1];
    if (iter) iter[i] = s[1];
    if (status) status[i] = s[2];
}

}  // namespace

int fit_scratch_address(int device, void** out) {
    static std::mutex mu;
    static void* addr[kMaxDevices] = {};
    if (device < 0 || device >= kMaxDevices) return MOED_EARG;
    std::lock_guard<std::mutex> lk(mu);
    if (!addr[device]) {
        void* p = nullptr;
        // the device must be current: every GPU has its own copy of the array
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return MOED_EHIP;
        addr[device] = p;
    }
    *out = addr[device];
    return MOED_OK;
}

int fit_minimize_noalloc(const FitCtx& c, void* scratch_addr, int nfit, int nx, double* x, int nw,
                         const double* omega, const double* Dev, double ftol, double* fret, int32_t* iter,
                         int32_t* status) {
    if (nfit > kNoAllocFits) {
        if (c.err) *c.err = "minimize_device_noalloc: at most " + std::to_string(kNoAllocFits) + " fits per call";
        return MOED_EARG;
    }
    if (nx > kFitNx || c.nbath > kFitMaxBath || c.norb > kMaxOrb) return argfail(c, "fit: nx or Nbath too large");
    NoAllocScratch* sc = static_cast<NoAllocScratch*>(scratch_addr);
    int32_t* flag = sc->flag;
    FCHK(zero_i32(flag, 64, c.stream));
    const int ndev = 2 * c.norb * nw, nxall = nfit * nx;
    const int nb = std::min(1024, (std::max({nw, ndev, nxall}) + 255) / 256);
    hipLaunchKernelGGL(fit_check_kernel, dim3(nb), dim3(256), 0, c.stream, omega, nw, Dev, ndev, x, nxall, flag + 1);
    FCHK(hipGetLastError());
    int32_t bad = 0;
    FCHK(hipMemcpyAsync(&bad, flag + 1, sizeof(int32_t), hipMemcpyDeviceToHost, c.stream));
    FCHK(hipStreamSynchronize(c.stream));
    if (bad & 1) return argfail(c, "fit: omega must be positive");
    if (bad & 2) return argfail(c, "fit: omega and Dev must be finite");
    if (bad & 4) return argfail(c, "fit: x must be finite");
    int rc = fit_run(c, nfit, nx, x, nw, omega, Dev, ftol, sc->st, sc->ist, flag);
    if (rc) return rc;
    hipLaunchKernelGGL(fit_results_kernel, dim3((nfit + 255) / 256), dim3(256), 0, c.stream, sc->st, sc->ist, nfit,
                       fret, iter, status, flag + 2);
    FCHK(hipGetLastError());
    int32_t unfinished = 0;
    FCHK(hipMemcpyAsync(&unfinished, flag + 2, sizeof(int32_t), hipMemcpyDeviceToHost, c.stream));
    FCHK(hipStreamSynchronize(c.stream));
    if (unfinished) {
        if (c.err) *c.err = "fit: a fit did not finish";
        return MOED_ECONV;
    }
    return MOED_OK;
}

}  // namespace moed

// s0 + terms[0] + ... + terms[n-1], one rounding per add in this order, by the block-wide replay the
// reference-order chi^2 uses (no terms per thread, 1 <= no <= 8; the first nseq added one by one).
// Host arrays; synchronous on the null stream.  Not part of the solver's interface: the check of the
// replay against the plain loop on inputs chosen to exercise its fallbacks.
extern "C" int moed_check_ordered_sum(const double* terms, int n, int no, int nseq, double s0, double* out) {
    using namespace moed;
    if (n < 0 || no < 1 || no > 2 * kMaxOrb || nseq < 0 || !out || (n > 0 && !terms)) return MOED_EARG;
    FitCtx c{};
    Buf dt, dout;
    FCHK(dt.alloc((size_t)n * sizeof(double)));
    FCHK(dout.alloc(sizeof(double)));
    if (n > 0) FCHK(hipMemcpy(dt.p, terms, (size_t)n * sizeof(double), hipMemcpyHostToDevice));
    hipLaunchKernelGGL(ordered_sum_kernel, dim3(1), dim3(kFitThreads), 0, 0, dt.as<double>(), n, no, nseq, s0,
                       dout.as<double>());
    FCHK(hipGetLastError());
    FCHK(hipMemcpy(out, dout.p, sizeof(double), hipMemcpyDeviceToHost));
    return MOED_OK;
}
