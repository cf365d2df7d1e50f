// Host side of libmoed: the handle (the reference's ed_config.f90 module state), the per-popcount
// string tables, the Lanczos drivers, the sector eigensolver that replaces PARPACK's diag
// (ed_solver.F90:107-191), ed_solve / ed_calc_green_ftn drivers and the extern "C" ABI of moed.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <string>
#include <vector>
#include <atomic>
#include <thread>

#include "moed_fit.h"
#include "moed_internal.h"
#include "moed_ritzconv.h"
#include "tridiag.h"

using namespace moed;

namespace {

struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    hipError_t ensure(size_t b) {
        if (b <= bytes && p) return hipSuccess;
        release();
        hipError_t e = hipMalloc(&p, std::max<size_t>(b, 256));
        if (e != hipSuccess) {
            p = nullptr;
            return e;
        }
        bytes = std::max<size_t>(b, 256);
        return hipSuccess;
    }
    template <class T>
    T* as() const {
        return reinterpret_cast<T*>(p);
    }
};

int64_t binom(int n, int k) {
    if (k < 0 || k > n) return 0;
    int64_t r = 1;
    for (int i = 1; i <= k; ++i) r = r * (n - k + i) / i;
    return r;
}

// sign parity of moving one electron between bits p and q of string s: popcount of the bits strictly between
inline int move_parity(uint32_t s, int p, int q) {
    const int lo = std::min(p, q), hi = std::max(p, q);
    const uint32_t mask = ((1u << hi) - 1u) & ~((1u << (lo + 1)) - 1u);
    return __builtin_popcount(s & mask) & 1;
}

struct StrTable {
    int n = 0, count = 0, mh = 0, nslot = 0, rw = 0;
    std::vector<uint32_t> str;
    std::vector<int32_t> rank;  // [2^Ns]: rank within this popcount class, -1 otherwise
    DevBuf d_str, d_nhop, d_mv, d_mvc, d_imp, d_rec, d_rec8, d_cperm;
    StrDev dev() const {
        return {d_str.as<uint16_t>(), d_nhop.as<int32_t>(), d_mv.as<uint32_t>(), d_mvc.as<double>(),
                d_imp.as<int32_t>(),  d_rec.as<uint4>(),    d_rec8.as<uint4>(),  count, mh, nslot, rw, n,
                d_cperm.as<int32_t>()};
    }
};

struct SectorEig {
    int nu = 0, nd = 0;
    int64_t D = 0;
    int nfound = 0;
    std::vector<double> ev;  // nev values, DBL_MAX where none
    DevBuf vec;              // nfound vectors of D doubles
};

// One-body move slot of the H v row pass: bath slot (orbital bit p, bath bit q), applies when
// exactly one of the two is occupied; impurity slot (p = i, q = j), applies when j is occupied and
// i is empty (c+_i c_j).
struct Slot {
    int p, q;
    bool imp;
    double c;
};

struct GreenRecord {
    int iev, iorb, ispin;
    double E, p;
    int even, nsp, nsn;
    std::vector<double> ap, bp, an, bn;
};

// done-flag copies of the eigensolver's convergence tests in flight at once
constexpr int kFlagRing = 4;

}  // namespace

struct moed_handle {
    int device = 0;
    hipStream_t stream = nullptr;
    int norb = 0, nsite = 0, nbath = 0;
    double U = 0, J = 0, beta = 0, UmJ3 = 0, Um2J = 0;
    int nstep = 0, nev = 0, nw = 0;
    uint32_t flags = 0;
    double gs_tol = 1e-12;
    int gs_maxvec = 400;
    int eig_chunk = 0;  // Krylov basis vectors per allocation (0: what fits in kBasisBytes)
    std::vector<double> ek, vk, tij;
    bool has_tij = false;
    std::vector<int> nelec, nup;
    std::string err;
    std::vector<double> coef;
    DevBuf d_ek, d_coef;
    std::map<int, std::unique_ptr<StrTable>> tables;
    std::vector<std::unique_ptr<SectorEig>> eig;
    bool solved = false;
    std::vector<double> kept_val, kept_prob;
    std::vector<int> kept_sector, kept_level;
    std::vector<GreenRecord> records;
    std::vector<Slot> slots;
    DevBuf d_slotc;
    int hv_threads = 0, hv_strip = 0;
    int fit_budget = 0;            // CG iterations per fit and launch of moed_minimize (0: default)
    std::vector<int32_t> fit_nfev;  // function evaluations of each fit of the last moed_minimize
    void* fit_scratch = nullptr;    // this device's copy of the allocation-free minimiser's state array
    int32_t* pinned_i32 = nullptr;  // pinned host buffer for flags read back by the eigensolver
    hipStream_t stream2 = nullptr;  // second stream (default priority): the persistent eigensolver
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    // third stream (default priority): the eigensolver's convergence tests beside its steps
    hipStream_t stream3 = nullptr;
    hipEvent_t ev_step = nullptr, ev_convd = nullptr;
    hipEvent_t ev_flags[kFlagRing] = {};  // done-flag copies to pinned memory in flight
    int64_t persist_max = -1;       // largest sector dimension for the persistent launch (-1: default)
    size_t pinned_n = 0;
    int ncu = 0;  // compute units of the device
    DevBuf ws[9];
    DevBuf eng[8];  // multi-problem Lanczos engine (eig_batch)
    DevBuf gbuf[5];  // batched Green's-function chains (green_batch)
    // instantiated HIP graphs of the reference recurrence's steps 1..nstep (lanczos_ref_dev), per
    // sector, step count and buffer set; cleared whenever a captured pointer may change
    struct LzGraph {
        int nu, nd, nstep;
        int64_t ldx;
        const void* ptr[6];
        hipGraph_t graph;
        hipGraphExec_t exec;
    };
    std::vector<LzGraph> lz_graphs;
    void clear_graphs() {
        for (auto& g : lz_graphs) {
            (void)hipGraphExecDestroy(g.exec);
            (void)hipGraphDestroy(g.graph);
        }
        lz_graphs.clear();
    }
    std::vector<std::unique_ptr<DevBuf>> basis;  // its Krylov basis, in chunks of whole vectors
};

namespace {

enum { WS_X = 0, WS_Y, WS_PART, WS_SCAL, WS_V, WS_W, WS_MISC, WS_MISC2, WS_DOT };

// f(0..n-1) on up to 16 host threads (independent work items)
template <class F>
void parallel_for(int n, F&& f) {
    const int nt = std::max(1, std::min({n, 16, (int)std::thread::hardware_concurrency()}));
    if (nt <= 1) {
        for (int i = 0; i < n; ++i) f(i);
        return;
    }
    std::atomic<int> next{0};
    std::vector<std::thread> th;
    for (int t = 0; t < nt; ++t)
        th.emplace_back([&] {
            for (int i = next++; i < n; i = next++) f(i);
        });
    for (auto& t : th) t.join();
}

int fail(moed_handle* h, int code, const std::string& msg) {
    if (h) h->err = msg;
    return code;
}

int hipfail(moed_handle* h, hipError_t e, const char* what) {
    return fail(h, e == hipErrorOutOfMemory ? MOED_ENOMEM : MOED_EHIP,
                std::string(what) + ": " + hipGetErrorString(e));
}

#define HIPCHK(expr)                                   \
    do {                                               \
        hipError_t _e = (expr);                        \
        if (_e != hipSuccess) return hipfail(h, _e, #expr); \
    } while (0)

// U - 3*J and U - 2*J rounded exactly as the reference forms them (ed_operators.F90:159,162):
// the product first, then the difference, no fused multiply-add.
void interaction_constants(double U, double J, double* umj3, double* um2j) {
#pragma clang fp contract(off)
    *umj3 = U - 3.0 * J;
    *um2j = U - 2.0 * J;
}

int set_model(moed_handle* h, const double* ek, const double* vk) {
    h->ek.assign(ek, ek + h->nsite);
    h->vk.assign(vk, vk + (size_t)h->norb * h->nbath);
    for (double v : h->ek)
        if (!std::isfinite(v)) return fail(h, MOED_EARG, "ek contains a non-finite value");
    for (double v : h->vk)
        if (!std::isfinite(v)) return fail(h, MOED_EARG, "vk contains a non-finite value");
    if ((size_t)h->norb * h->nbath + (size_t)h->norb * h->norb >= (size_t)kMaxCoef) return fail(h, MOED_EARG, "too many hopping coefficients");
    h->coef.assign(kMaxCoef, 0.0);  // the last slot stays 0.0: padding entries of the move tables point at it
    for (size_t i = 0; i < h->vk.size(); ++i) h->coef[i] = h->vk[i];
    if (h->has_tij)
        for (int i = 0; i < h->norb * h->norb; ++i) h->coef[(size_t)h->norb * h->nbath + i] = h->tij[i];
    HIPCHK(h->d_ek.ensure(16 * sizeof(double)));
    std::vector<double> ek16(16, 0.0);
    std::copy(h->ek.begin(), h->ek.end(), ek16.begin());
    HIPCHK(hipMemcpy(h->d_ek.p, ek16.data(), 16 * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(h->d_coef.ensure(h->coef.size() * sizeof(double)));
    HIPCHK(hipMemcpy(h->d_coef.p, h->coef.data(), h->coef.size() * sizeof(double), hipMemcpyHostToDevice));
    // slots in the order of the move lists: (orbital, bath) pairs, then impurity pairs
    h->slots.clear();
    for (int o = 0; o < h->norb; ++o)
        for (int b = 0; b < h->nbath; ++b)
            if (h->coef[o + (size_t)h->norb * b] != 0.0)
                h->slots.push_back({o, h->norb + b, false, h->coef[o + (size_t)h->norb * b]});
    if (h->has_tij)
        for (int i = 0; i < h->norb; ++i)
            for (int j = 0; j < h->norb; ++j) {
                const double t = h->coef[(size_t)h->norb * h->nbath + i + h->norb * j];
                if (i != j && t != 0.0) h->slots.push_back({i, j, true, t});
            }
    if (h->slots.size() > 59) return fail(h, MOED_EARG, "too many one-body move slots");
    std::vector<double> slotc(64, 0.0);  // the row pass reads up to 59 slots; the rest are 0.0
    for (size_t k = 0; k < h->slots.size(); ++k) slotc[k] = h->slots[k].c;
    HIPCHK(h->d_slotc.ensure(slotc.size() * sizeof(double)));
    HIPCHK(hipMemcpy(h->d_slotc.p, slotc.data(), slotc.size() * sizeof(double), hipMemcpyHostToDevice));
    h->clear_graphs();  // captured launches hold table and buffer addresses
    h->tables.clear();  // onsite sums and move lists depend on ek, vk
    h->solved = false;
    return MOED_OK;
}

int build_table(moed_handle* h, int n, StrTable** out) {
    auto it = h->tables.find(n);
    if (it != h->tables.end()) {
        *out = it->second.get();
        return MOED_OK;
    }
    auto t = std::make_unique<StrTable>();
    const int ns = h->nsite, norb = h->norb, nbath = h->nbath;
    t->n = n;
    t->rank.assign((size_t)1 << ns, -1);
    for (uint32_t s = 0; s < (1u << ns); ++s)
        if (__builtin_popcount(s) == n) {
            t->rank[s] = (int32_t)t->str.size();
            t->str.push_back(s);
        }
    t->count = (int)t->str.size();
    const int C = t->count;
    std::vector<std::vector<uint32_t>> moves(C);
    std::vector<double> onsite(C);
    std::vector<int32_t> imp((size_t)norb * norb * C, -1);
    for (int r = 0; r < C; ++r) {
        const uint32_t s = t->str[r];
        double S = 0.0;
        for (int site = 0; site < ns; ++site)
            if ((s >> site) & 1u) S = S + h->ek[site];
        onsite[r] = S;
        auto& mv = moves[r];
        for (int o = 0; o < norb; ++o)
            for (int b = 0; b < nbath; ++b) {
                const int ci = o + norb * b;
                if (h->coef[ci] == 0.0) continue;
                const int sb = norb + b;
                const uint32_t ho = (s >> o) & 1u, hb = (s >> sb) & 1u;
                if (ho == hb) continue;
                const uint32_t tgt = s ^ (1u << o) ^ (1u << sb);
                mv.push_back((uint32_t)t->rank[tgt] | ((uint32_t)ci << 16) | ((uint32_t)move_parity(s, o, sb) << 31));
            }
        if (h->has_tij)
            for (int i = 0; i < norb; ++i)
                for (int j = 0; j < norb; ++j) {
                    const int ci = norb * nbath + i + norb * j;
                    if (i == j || h->coef[ci] == 0.0) continue;
                    if (((s >> j) & 1u) && !((s >> i) & 1u)) {
                        const uint32_t tgt = s ^ (1u << i) ^ (1u << j);
                        mv.push_back((uint32_t)t->rank[tgt] | ((uint32_t)ci << 16) |
                                     ((uint32_t)move_parity(s, i, j) << 31));
                    }
                }
        if ((int)mv.size() > kMaxHop) return fail(h, MOED_EARG, "too many one-body moves per string");
        t->mh = std::max(t->mh, (int)mv.size());
        for (int a = 0; a < norb; ++a)
            for (int b = 0; b < norb; ++b) {
                if (a == b) continue;
                if (((s >> b) & 1u) && !((s >> a) & 1u)) {
                    const uint32_t tgt = s ^ (1u << a) ^ (1u << b);
                    imp[(size_t)(a * norb + b) * C + r] = t->rank[tgt] | (move_parity(s, a, b) << 16);
                }
            }
    }
    // the strings in component order under the orbital-bath moves (coefficient index below
    // norb * nbath), components numbered by their lowest rank.  The strip pass sweeps its rows in
    // this order, so the rows live in one XCD's L2 at a time are one component's (at 14 sites the
    // largest holds 1225 of the 3432 rows; strip pass 54.2 -> 53.0 us).  Impurity hopping moves
    // cross components, and with them the order measured slower (15-site sector: 402 -> 421 us),
    // so those models keep the rank order (cperm null).
    if (!h->has_tij) {
        std::vector<int32_t> par(C);
        for (int r = 0; r < C; ++r) par[r] = r;
        auto find = [&](int32_t a) {
            while (par[a] != a) a = par[a] = par[par[a]];
            return a;
        };
        for (int r = 0; r < C; ++r)
            for (uint32_t e : moves[r])
                if ((int)((e >> 16) & 0xffu) < norb * nbath) {
                    const int32_t a = find(r), b = find((int32_t)(e & 0xffffu));
                    if (a != b) par[std::max(a, b)] = std::min(a, b);
                }
        std::vector<int32_t> cid(C), first(C, -1), cnt;
        for (int r = 0; r < C; ++r) {
            const int32_t root = find(r);
            if (first[root] < 0) {
                first[root] = (int32_t)cnt.size();
                cnt.push_back(0);
            }
            cid[r] = first[root];
            ++cnt[cid[r]];
        }
        std::vector<int32_t> off(cnt.size() + 1, 0), cperm(C);
        for (size_t c = 0; c < cnt.size(); ++c) off[c + 1] = off[c] + cnt[c];
        for (int r = 0; r < C; ++r) cperm[off[cid[r]]++] = r;
        HIPCHK(t->d_cperm.ensure(C * sizeof(int32_t)));
        HIPCHK(hipMemcpy(t->d_cperm.p, cperm.data(), C * sizeof(int32_t), hipMemcpyHostToDevice));
    }
    // row-major move lists, padded to a multiple of 8 (the strip pass reads them eight at a time)
    const int mh = (std::max(t->mh, 1) + 7) / 8 * 8;
    t->mh = mh;
    std::vector<uint32_t> mvt((size_t)mh * C);
    std::vector<double> mvc((size_t)mh * C, 0.0);
    std::vector<int32_t> nhop(C);
    std::vector<uint16_t> str16(C);
    for (int r = 0; r < C; ++r) {
        nhop[r] = (int32_t)moves[r].size();
        str16[r] = (uint16_t)t->str[r];
        for (int k = 0; k < mh; ++k) {
            if (k < nhop[r]) {
                const uint32_t e = moves[r][k];
                const double c = h->coef[(e >> 16) & 0xffu];
                mvt[(size_t)r * mh + k] = e & 0xffffu;
                mvc[(size_t)r * mh + k] = (e >> 31) ? -c : c;
            } else {
                mvt[(size_t)r * mh + k] = (uint32_t)r;
            }
        }
    }
    // row-pass records (moed_internal.h, StrDev::rec): onsite prefix, string, slot entries
    t->nslot = (int)h->slots.size();
    // record_dwords(nslot) dwords per record, back to back: the onsite prefix, the string and the
    // slots (the last slot positions are padding entries when nslot is even)
    t->rw = record_dwords(t->nslot);
    if (t->rw == 0) return fail(h, MOED_EARG, "too many move slots per string");
    const size_t ru16 = (size_t)t->rw * 2;  // uint16 per record
    std::vector<uint16_t> rec(ru16 * C, (uint16_t)C);
    for (int r = 0; r < C; ++r) {
        const uint32_t s = t->str[r];
        uint16_t* rr = rec.data() + ru16 * r;
        std::memcpy(rr, &onsite[r], sizeof(double));
        rr[4] = (uint16_t)s;
        for (int k = 0; k < t->nslot; ++k) {
            const Slot& sl = h->slots[k];
            const uint32_t bp = (s >> sl.p) & 1u, bq = (s >> sl.q) & 1u;
            const bool ok = sl.imp ? (bq && !bp) : (bp != bq);
            if (!ok) continue;  // entry C: the zero
            const uint32_t tgt = s ^ (1u << sl.p) ^ (1u << sl.q);
            rr[5 + k] = (uint16_t)(move_parity(s, sl.p, sl.q) ? row_neg(C) + t->rank[tgt] : t->rank[tgt]);
        }
    }
    HIPCHK(t->d_rec.ensure(rec.size() * sizeof(uint16_t)));
    HIPCHK(hipMemcpy(t->d_rec.p, rec.data(), rec.size() * sizeof(uint16_t), hipMemcpyHostToDevice));
    if (8 * (int64_t)row_len(C) <= 65536) {  // the same records with slot entries as LDS byte offsets
        for (int r = 0; r < C; ++r)  // every slot position, the padding slots' zero entries included
            for (size_t k = 5; k < ru16; ++k) rec[ru16 * r + k] = (uint16_t)(8 * rec[ru16 * r + k]);
        HIPCHK(t->d_rec8.ensure(rec.size() * sizeof(uint16_t)));
        HIPCHK(hipMemcpy(t->d_rec8.p, rec.data(), rec.size() * sizeof(uint16_t), hipMemcpyHostToDevice));
    }
    str16.resize(C + 2, 0);  // scalar reads fetch whole 4-byte words
    HIPCHK(t->d_str.ensure((C + 2) * sizeof(uint16_t)));
    HIPCHK(t->d_nhop.ensure(C * sizeof(int32_t)));
    HIPCHK(t->d_mv.ensure(mvt.size() * sizeof(uint32_t)));
    HIPCHK(t->d_mvc.ensure(mvc.size() * sizeof(double)));
    HIPCHK(t->d_imp.ensure(std::max<size_t>(imp.size(), 1) * sizeof(int32_t)));
    HIPCHK(hipMemcpy(t->d_str.p, str16.data(), (C + 2) * sizeof(uint16_t), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(t->d_nhop.p, nhop.data(), C * sizeof(int32_t), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(t->d_mv.p, mvt.data(), mvt.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(t->d_mvc.p, mvc.data(), mvc.size() * sizeof(double), hipMemcpyHostToDevice));
    if (!imp.empty()) HIPCHK(hipMemcpy(t->d_imp.p, imp.data(), imp.size() * sizeof(int32_t), hipMemcpyHostToDevice));
    *out = t.get();
    h->tables[n] = std::move(t);
    return MOED_OK;
}

struct Sector {
    int nu, nd;
    int64_t D;
    StrTable* up;
    StrTable* dn;
};

int get_sector(moed_handle* h, int nu, int nd, Sector* s) {
    if (nu < 0 || nd < 0 || nu > h->nsite || nd > h->nsite)
        return fail(h, MOED_ESECTOR, "sector (" + std::to_string(nu) + "," + std::to_string(nd) + ") out of range");
    int rc = build_table(h, nu, &s->up);
    if (rc) return rc;
    rc = build_table(h, nd, &s->dn);
    if (rc) return rc;
    s->nu = nu;
    s->nd = nd;
    s->D = (int64_t)s->up->count * s->dn->count;
    return MOED_OK;
}

// ldx: row stride of the vectors in doubles (0: nU, the reference layout)
HvPlan hv_plan_of(moed_handle* h, const Sector& s, int64_t ldx = 0) {
    return hv_plan(s.up->count, s.dn->count, h->hv_threads, h->hv_strip, h->ncu, ldx > 0 ? ldx : s.up->count);
}

// number of x.y partials hv_dev writes to dot_part
int hv_groups(moed_handle* h, const Sector& s, int64_t ldx = 0) { return hv_plan_of(h, s, ldx).grid_b; }

// y = H x (x != y) on stream st; optional fused y *= 1/yscale[0] and x.y partials.  ldx: row stride
// of x and y (0: nU); padding columns of x must hold zeros and are never written in y.
HvArgs hv_args(moed_handle* h, const Sector& s, const double* x, double* y, double* dot_part, const double* yscale,
               const int32_t* stop, int64_t ldx) {
    HvArgs a{};
    a.x = x;
    a.y = y;
    a.ldx = ldx > 0 ? ldx : s.up->count;
    a.up = s.up->dev();
    a.dn = s.dn->dev();
    a.ek = h->d_ek.as<double>();
    a.coef = h->d_coef.as<double>();
    a.slotc = h->d_slotc.as<double>();
    a.U = h->U;
    a.UmJ3 = h->UmJ3;
    a.Um2J = h->Um2J;
    a.J = h->J;
    a.dot_part = dot_part;
    a.yscale = yscale;
    a.stop = stop;
    return a;
}

int hv_dev(moed_handle* h, const Sector& s, const double* x, double* y, double* dot_part, const double* yscale,
           const int32_t* stop, hipStream_t st, int64_t ldx = 0) {
    HvArgs a{};
    a.x = x;
    a.y = y;
    a.ldx = ldx > 0 ? ldx : s.up->count;
    a.up = s.up->dev();
    a.dn = s.dn->dev();
    a.ek = h->d_ek.as<double>();
    a.coef = h->d_coef.as<double>();
    a.slotc = h->d_slotc.as<double>();
    a.U = h->U;
    a.UmJ3 = h->UmJ3;
    a.Um2J = h->Um2J;
    a.J = h->J;
    a.dot_part = dot_part;
    a.yscale = yscale;
    a.stop = stop;
    HIPCHK(launch_hv(a, hv_plan_of(h, s, a.ldx), h->norb, st));
    return MOED_OK;
}

// target sector of c+ (pm=1) / c (pm=2) on spin ispin
void target_sector(int pm, int ispin, int nu, int nd, int* nuo, int* ndo) {
    const int d = pm == 1 ? 1 : -1;
    *nuo = nu + (ispin == 1 ? d : 0);
    *ndo = nd + (ispin == 2 ? d : 0);
}

// y = c(+)_{iorb,ispin} x: sign of the operator evaluated on the target state, as apply_c does
// (ed_operators.F90:44-58 calls destruction_op/creation_op on basis_out states).
int apply_c_dev(moed_handle* h, int pm, int iorb, int ispin, const Sector& in, const Sector& out, const double* x,
                double* y, hipStream_t st) {
    const int o = iorb - 1;
    const StrTable* tin = ispin == 1 ? in.up : in.dn;
    const StrTable* tout = ispin == 1 ? out.up : out.dn;
    std::vector<int32_t> map(tout->count, -1);
    for (int r = 0; r < tout->count; ++r) {
        const uint32_t s = tout->str[r];
        const bool occ = (s >> o) & 1u;
        if ((pm == 1) != occ) continue;  // c+ |src> = |tgt> needs bit o set in tgt; c needs it clear
        const uint32_t src = s ^ (1u << o);
        int par = __builtin_popcount(s & ((1u << o) - 1u)) & 1;
        if (ispin == 2) par ^= (in.nu & 1);  // all up bits lie below the down bits (ed_operators.F90:65)
        map[r] = tin->rank[src] | (par << 30);
    }
    DevBuf& mb = h->ws[WS_MISC2];
    HIPCHK(mb.ensure(map.size() * sizeof(int32_t)));
    HIPCHK(hipMemcpyAsync(mb.p, map.data(), map.size() * sizeof(int32_t), hipMemcpyHostToDevice, st));
    ApplyCArgs a{x, y, mb.as<int32_t>(), out.up->count, out.dn->count, in.up->count, ispin == 1 ? 0 : 1};
    HIPCHK(launch_apply_c(a, st));
    HIPCHK(hipStreamSynchronize(st));  // map buffer is reused by the next call
    return MOED_OK;
}

int dot_dev(moed_handle* h, const double* a, const double* b, int64_t n, double* out_host, hipStream_t st) {
    DevBuf& part = h->ws[WS_DOT];
    HIPCHK(part.ensure((kRedBlocks + 8) * sizeof(double)));
    double* p = part.as<double>();
    HIPCHK(launch_dot_partial(a, b, n, p, st));
    HIPCHK(launch_reduce_final(p, kRedBlocks, p + kRedBlocks, 0, 0, nullptr, st));
    HIPCHK(hipMemcpyAsync(out_host, p + kRedBlocks, sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    return MOED_OK;
}

const double kBreakdown = (double)1.e-12f;  // ed_lanczos.F90:62, a REAL*4 literal

// lanczos_iteration (ed_lanczos.F90:14-68) with the vectors kept on the device.
int lanczos_ref_dev(moed_handle* h, const Sector& s, const double* u0, int nstep, int* nsc, double* a, double* b) {
    if (nstep < 1) return fail(h, MOED_EARG, "nstep must be >= 1");
    hipStream_t st = h->stream;
    // rows of odd length run padded to an even stride, so that the H v strip pass reads 16-byte
    // pairs; the padding columns hold zeros throughout (H v never writes them and the three-term
    // updates of zeros stay zero), so every sum is over the sector's own elements.  With three or
    // more orbitals the two-column strip kernel (85 VGPRs, 5 waves per SIMD) is slower than the
    // one-column kernel on the unpadded rows (15-site G stage 1.55 against 1.51 s), so those stay
    // as they are.
    const int64_t nU = s.up->count, nD = s.dn->count;
    const int64_t ldx = (nU % 2 == 1 && h->hv_strip != 64 && h->norb <= 2) ? nU + 1 : nU;
    const int64_t D = nD * ldx;
    const int ngrp = hv_groups(h, s, ldx);
    DevBuf& V = h->ws[WS_V];
    DevBuf& W = h->ws[WS_W];
    DevBuf& part = h->ws[WS_PART];
    DevBuf& scal = h->ws[WS_SCAL];
    HIPCHK(V.ensure((size_t)3 * D * sizeof(double)));
    HIPCHK(W.ensure((size_t)D * sizeof(double)));
    HIPCHK(part.ensure((size_t)(std::max(kRedBlocks, ngrp) + 8) * sizeof(double)));
    HIPCHK(scal.ensure((size_t)(2 * (nstep + 1) + 2) * sizeof(double)));
    double* u[3] = {V.as<double>(), V.as<double>() + D, V.as<double>() + 2 * D};
    double* w = W.as<double>();
    double* p = part.as<double>();
    LzScal sc{scal.as<double>(), scal.as<double>() + (nstep + 1),
              reinterpret_cast<int32_t*>(scal.as<double>() + 2 * (nstep + 1)),
              reinterpret_cast<int32_t*>(scal.as<double>() + 2 * (nstep + 1)) + 1};
    HIPCHK(launch_fill(0.0, 2 * (nstep + 1), scal.as<double>(), st));
    const int32_t init[2] = {0, nstep};
    HIPCHK(hipMemcpyAsync(sc.stop, init, sizeof(init), hipMemcpyHostToDevice, st));
    if (ldx != nU) {
        HIPCHK(launch_fill(0.0, 3 * D, V.as<double>(), st));
        HIPCHK(launch_fill(0.0, D, w, st));
        HIPCHK(launch_pad_rows(u0, nU, nD, u[0], ldx, st));
    } else {
        HIPCHK(hipMemcpyAsync(u[0], u0, D * sizeof(double), hipMemcpyDeviceToDevice, st));
    }
    // B[0] = |u_0|^2 ; w = H u_0 ; A[0] = <u_0, w> / B[0]
    HIPCHK(launch_dot_partial(u[0], u[0], D, p, st));
    HIPCHK(launch_lz_finish_b(0, p, kRedBlocks, sc, st));
    int rc = hv_dev(h, s, u[0], w, p, nullptr, sc.stop, st, ldx);
    if (rc) return rc;
    HIPCHK(launch_lz_finish_a(0, p, ngrp, sc, kBreakdown, st));
    // steps 1..nstep: a fixed sequence of kernel launches (nothing else), captured once per sector,
    // step count and buffer set into a HIP graph and replayed as one launch on later calls
    const void* key_ptr[6] = {V.p, W.p, part.p, scal.p, s.up->d_rec.p, s.dn->d_mv.p};
    moed_handle::LzGraph* gr = nullptr;
    for (auto& g : h->lz_graphs)
        if (g.nu == s.nu && g.nd == s.nd && g.nstep == nstep && g.ldx == ldx &&
            std::equal(key_ptr, key_ptr + 6, g.ptr))
            gr = &g;
    if (!gr) {
        moed_handle::LzGraph g{s.nu, s.nd, nstep, ldx, {}, nullptr, nullptr};
        std::copy(key_ptr, key_ptr + 6, g.ptr);
        HIPCHK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        hipError_t ce = hipSuccess;
        for (int j = 1; j <= nstep && ce == hipSuccess; ++j) {
            ce = launch_lz_update(j, w, u[(j - 1) % 3], u[(j + 1) % 3], u[j % 3], D, sc, p, st);
            if (ce == hipSuccess) ce = launch_lz_finish_b(j, p, kRedBlocks, sc, st);
            if (ce == hipSuccess) {
                HvArgs a = hv_args(h, s, u[j % 3], w, p, nullptr, sc.stop, ldx);
                ce = launch_hv(a, hv_plan_of(h, s, ldx), h->norb, st);
            }
            if (ce == hipSuccess) ce = launch_lz_finish_a(j, p, ngrp, sc, kBreakdown, st);
        }
        const hipError_t ee = hipStreamEndCapture(st, &g.graph);  // always ends the capture
        if (ce == hipSuccess) ce = ee;
        if (ce == hipSuccess) ce = hipGraphInstantiate(&g.exec, g.graph, nullptr, nullptr, 0);
        if (ce != hipSuccess) {
            if (g.graph) (void)hipGraphDestroy(g.graph);
            return hipfail(h, ce, "capturing the Lanczos steps in a HIP graph");
        }
        h->lz_graphs.push_back(g);
        gr = &h->lz_graphs.back();
    }
    HIPCHK(hipGraphLaunch(gr->exec, st));
    std::vector<double> hs(2 * (nstep + 1) + 2);
    HIPCHK(hipMemcpyAsync(hs.data(), scal.p, hs.size() * sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    for (int j = 0; j <= nstep; ++j) {
        a[j] = hs[j];
        b[j] = hs[nstep + 1 + j];
    }
    b[0] = 0.0;  // ed_lanczos.F90:40
    int32_t flags[2];
    std::memcpy(flags, &hs[2 * (nstep + 1)], sizeof(flags));
    *nsc = flags[1];
    // entries past a breakdown are zero, as the reference leaves them (ed_lanczos.F90:27-28)
    for (int j = *nsc + 1; j <= nstep; ++j) a[j] = b[j] = 0.0;
    return MOED_OK;
}

// ----------------------------------------------------------------- eigensolver (replaces PARPACK diag)

// Lanczos with partial reorthogonalisation for several sectors at once; one sector is the
// per-sector form of the same code.  Every problem p runs the same iteration as alone: start
// vector u_0 (counter hash, seed of its sector), then per step w = H u_k / |u_k|, u_{k+1} = w minus
// its components along u_k, u_{k-1}; where the estimated loss of orthogonality exceeds sqrt(eps)
// (moed_eig.hip, eig_finish_kernel), u_{k+1} -= sum_i <u_i, u_{k+1}>/|u_i|^2 u_i over the whole basis
// (second pass when |u_{k+1}|^2 drops below half); alpha_k, beta_k = |u_{k+1}|.  A problem stops when the Krylov space is exhausted
// (beta_k <= 1e-14 |T|), at its basis cap, or when the lowest nev Ritz pairs have relative
// residual |beta_m z_{m,i}| <= tol |theta_i| (tested on the host after steps m < 20 and every
// fifth step after, with one synchronisation for all problems).  Ritz vectors V S on the device.
// The vectors are packed end to end; every launch serves all problems (moed_eig.hip).
enum { EG_VT = 0, EG_W, EG_I32, EG_I64, EG_DBL, EG_DESC, EG_OUT, EG_N };

// The Krylov basis is allocated as the iteration reaches it, in chunks of whole vectors
// (eig_chunk vectors; default: what fits in kBasisBytes), and the handle keeps the chunks for
// later calls.  Its cap (gs_maxvec + 1 vectors: 401 vectors of 331 MB at the 15-site sector) is
// far more than a solve uses, and device allocations cost ~30 ms/GiB once the memory has been used
// before (the driver clears it), so allocating the cap up front made the first solve 2-4 s instead
// of 0.3 s.  Kernels address the vectors through a device table of their addresses.
constexpr size_t kBasisBytes = (size_t)4 << 30;

// persistent: problems small enough for one workgroup (D <= kPersistMaxDim, at most kPersistMaxU up
// strings) run their whole iteration in one persistent work-queue launch (moed_persist.hip) instead
// of step by step; the rest go step by step as before.
// Deflation (degeneracy check): the iteration runs on H' = H + sigma_p sum_i L_i L_i^T, with the
// locked vectors L_i packed like the basis (zero where a problem has fewer than nl).
// A problem only asks whether H' has a level below above[p]: its Ritz pairs whose residual interval
// lies wholly above it need not converge.
struct Deflation {
    const double* const* L;  // device table of nl packed vectors
    int nl;
    const double* sigma;     // device [P]
    const double* above;     // device [P]
    const double* above_h;   // the same, host [P]
};

int eig_batch(moed_handle* h, const std::vector<std::pair<int, int>>& secs, int nev, std::vector<SectorEig*>& out,
              bool persistent = false, const Deflation* defl = nullptr, uint64_t seed_salt = 0) {
    hipStream_t st = h->stream;
    const int P = (int)secs.size();
    if (P == 0) return MOED_OK;
    std::vector<Sector> S(P);
    std::vector<int64_t> off(P), dim(P);
    std::vector<int32_t> mmax(P);
    int64_t dtot = 0;
    int M = 1;
    for (int p = 0; p < P; ++p) {
        int rc = get_sector(h, secs[p].first, secs[p].second, &S[p]);
        if (rc) return rc;
        dim[p] = S[p].D;
        off[p] = dtot;
        dtot += dim[p] + (dim[p] & 1);  // every problem starts on an even element: the H v strip pass
                                        // reads 16-byte pairs (the gap element is never touched)
        mmax[p] = (int32_t)std::min<int64_t>(h->gs_maxvec, dim[p]);
        M = std::max(M, (int)mmax[p]);
    }
    const int ldh = M + 2;
    // vector items and H v row items (sectors up to 2^20 elements: one batched launch; larger
    // sectors: the two-pass product, one launch pair per sector)
    constexpr int64_t kSmall = (int64_t)1 << 20;
    std::vector<int32_t> item_prob, item_len, pitem0(P), pnitem(P), row_prob, row_id;
    std::vector<int64_t> item_start;
    std::vector<BatchProbDev> desc(P);
    int max_nU = 1;
    // problems for the persistent work-queue launch, largest first
    std::vector<char> pers(P, 0);
    std::vector<int32_t> pqueue;
    int pers_nU = 1, pers_M = 1;
    const int64_t pmax = std::min<int64_t>(h->persist_max >= 0 ? h->persist_max : kPersistDefaultDim, kPersistMaxDim);
    if (persistent && M + 2 <= kPersistMaxBasis) {
        for (int p = 0; p < P; ++p)
            if (dim[p] <= pmax && S[p].up->count <= kPersistMaxU) {
                pers[p] = 1;
                pqueue.push_back(p);
                pers_nU = std::max(pers_nU, S[p].up->count);
                pers_M = std::max(pers_M, (int)mmax[p]);
            }
        std::stable_sort(pqueue.begin(), pqueue.end(), [&](int a, int b) { return dim[a] > dim[b]; });
    }
    for (int p = 0; p < P; ++p) {
        pitem0[p] = (int32_t)item_prob.size();
        for (int64_t s0 = 0; s0 < dim[p]; s0 += kEigChunk) {
            item_prob.push_back(p);
            item_start.push_back(off[p] + s0);
            item_len.push_back((int32_t)std::min<int64_t>(kEigChunk, dim[p] - s0));
        }
        pnitem[p] = (int32_t)item_prob.size() - pitem0[p];
        desc[p] = {S[p].up->dev(), S[p].dn->dev(), off[p]};
        if (pers[p]) continue;  // its rows are the persistent kernel's
        if (dim[p] <= kSmall) {
            for (int id = 0; id < S[p].dn->count; ++id) {
                row_prob.push_back(p);
                row_id.push_back(id);
            }
            max_nU = std::max(max_nU, S[p].up->count);
        }
    }
    const int ni = (int)item_prob.size(), nr = (int)row_prob.size();
    // device memory (handle-owned, grown as needed)
    DevBuf* g = h->eng;
    const size_t npart = std::max<size_t>((size_t)ldh * ni, (size_t)nev * ni);
    HIPCHK(g[EG_VT].ensure((size_t)(M + 1) * sizeof(double*)));
    HIPCHK(g[EG_W].ensure((size_t)dtot * sizeof(double)));
    const size_t n32 = 2 * (size_t)ni + 10 * (size_t)P + 2 * (size_t)nr + pqueue.size() + 2;
    HIPCHK(g[EG_I32].ensure(n32 * sizeof(int32_t)));
    HIPCHK(g[EG_I64].ensure(((size_t)ni + 3 * P) * sizeof(int64_t)));
    const size_t nzs = (size_t)P * nev * 3 * ldh;  // convergence test workspace
    const size_t ndbl = 8 * (size_t)P * ldh + 4 * (size_t)P + npart + (size_t)P * nev * ldh + (size_t)P * nev + nzs;
    HIPCHK(g[EG_DBL].ensure(ndbl * sizeof(double)));
    HIPCHK(g[EG_DESC].ensure((size_t)P * sizeof(BatchProbDev)));
    HIPCHK(g[EG_OUT].ensure((size_t)P * sizeof(double*)));
    int32_t* i32 = g[EG_I32].as<int32_t>();
    int32_t *d_item_prob = i32, *d_item_len = d_item_prob + ni, *d_pitem0 = d_item_len + ni, *d_pnitem = d_pitem0 + P,
            *d_mmax = d_pnitem + P, *d_done = d_mmax + P, *d_pass2 = d_done + P, *d_mfinal = d_pass2 + P,
            *d_nf = d_mfinal + P, *d_row_prob = d_nf + P, *d_row_id = d_row_prob + nr, *d_full = d_row_id + nr,
            *d_force = d_full + P, *d_cflag = d_force + P, *d_pqueue = d_cflag + P, *d_head = d_pqueue + pqueue.size();
    int64_t *d_item_start = g[EG_I64].as<int64_t>(), *d_poff = d_item_start + ni, *d_pdim = d_poff + P,
            *d_seed = d_pdim + P;
    double* dbl = g[EG_DBL].as<double>();
    double *d_alpha = dbl, *d_beta = d_alpha + (size_t)P * ldh, *d_nrm = d_beta + (size_t)P * ldh,
           *d_h = d_nrm + (size_t)P * ldh, *d_c = d_h + (size_t)P * ldh, *d_om = d_c + (size_t)P * ldh,
           *d_aux = d_om + 3 * (size_t)P * ldh,
           *d_part = d_aux + 4 * (size_t)P, *d_S = d_part + npart, *d_n2 = d_S + (size_t)P * nev * ldh,
           *d_zs = d_n2 + (size_t)P * nev;
    std::vector<int32_t> zeroP(P, 0);
    HIPCHK(hipMemcpyAsync(d_item_prob, item_prob.data(), ni * sizeof(int32_t), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(d_item_len, item_len.data(), ni * sizeof(int32_t), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(d_pitem0, pitem0.data(), P * sizeof(int32_t), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(d_pnitem, pnitem.data(), P * sizeof(int32_t), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(d_mmax, mmax.data(), P * sizeof(int32_t), hipMemcpyHostToDevice, st));
    // the problems of the persistent launch count as done for the step-by-step kernels from the
    // start (the persistent kernel keeps its own flags and sets done and mfinal when it finishes)
    std::vector<int32_t> done0(P);
    for (int p = 0; p < P; ++p) done0[p] = pers[p];
    HIPCHK(hipMemcpyAsync(d_done, done0.data(), P * sizeof(int32_t), hipMemcpyHostToDevice, st));
    for (int p = 0; p < P; ++p) done0[p] = -done0[p];  // cflag: -1 for the persistent launch's problems
    HIPCHK(hipMemcpyAsync(d_cflag, done0.data(), P * sizeof(int32_t), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(d_pass2, zeroP.data(), P * sizeof(int32_t), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(d_mfinal, zeroP.data(), P * sizeof(int32_t), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(d_full, zeroP.data(), P * sizeof(int32_t), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(d_force, zeroP.data(), P * sizeof(int32_t), hipMemcpyHostToDevice, st));
    if (nr > 0) {
        HIPCHK(hipMemcpyAsync(d_row_prob, row_prob.data(), nr * sizeof(int32_t), hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync(d_row_id, row_id.data(), nr * sizeof(int32_t), hipMemcpyHostToDevice, st));
    }
    HIPCHK(hipMemcpyAsync(d_item_start, item_start.data(), ni * sizeof(int64_t), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(d_poff, off.data(), P * sizeof(int64_t), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(d_pdim, dim.data(), P * sizeof(int64_t), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(g[EG_DESC].p, desc.data(), P * sizeof(BatchProbDev), hipMemcpyHostToDevice, st));
    HIPCHK(launch_fill(0.0, (int64_t)(8 * (size_t)P * ldh + 4 * (size_t)P), dbl, st));

    EigArgs e{};
    e.P = P;
    e.ni = ni;
    e.ldh = ldh;
    e.item_prob = d_item_prob;
    e.item_start = d_item_start;
    e.item_len = d_item_len;
    e.pitem0 = d_pitem0;
    e.pnitem = d_pnitem;
    e.poff = d_poff;
    e.pdim = d_pdim;
    e.mmax = d_mmax;
    e.done = d_done;
    e.pass2 = d_pass2;
    e.mfinal = d_mfinal;
    e.alpha = d_alpha;
    e.beta = d_beta;
    e.nrm = d_nrm;
    e.h = d_h;
    e.c = d_c;
    e.aux = d_aux;
    e.part = d_part;
    e.om = d_om;
    e.full = d_full;
    e.force = d_force;
    e.full_mode = (h->flags & MOED_F_FULL_REORTH) ? 1 : 0;
    e.above = defl ? defl->above : nullptr;
    e.cflag = d_cflag;

    // basis vectors: those the handle's chunks already hold, then new chunks as steps need them
    const size_t vbytes = (size_t)dtot * sizeof(double);
    const int64_t per = h->eig_chunk > 0 ? h->eig_chunk : std::max<int64_t>(1, (int64_t)(kBasisBytes / vbytes));
    double** vt = g[EG_VT].as<double*>();
    std::vector<double*> vp;
    vp.reserve((size_t)M + 1);  // the table's host copy never moves (queued copies read it)
    size_t vp_dev = 0;          // entries already queued to the device table
    for (auto& c : h->basis)
        for (size_t o = 0; o + vbytes <= c->bytes && (int64_t)vp.size() < M + 1; o += vbytes)
            vp.push_back(reinterpret_cast<double*>(static_cast<char*>(c->p) + o));
    auto reserve = [&](int64_t need) -> int {
        while ((int64_t)vp.size() < need) {
            const int64_t nv = std::min<int64_t>(per, M + 1 - (int64_t)vp.size());
            auto c = std::make_unique<DevBuf>();
            if (c->ensure((size_t)nv * vbytes) != hipSuccess)
                return fail(h, MOED_ENOMEM, "eigensolver: allocating " + std::to_string(nv) + " Krylov vectors of " +
                                                std::to_string(vbytes) + " bytes failed");
            for (int64_t j = 0; j < nv; ++j)
                vp.push_back(reinterpret_cast<double*>(static_cast<char*>(c->p) + (size_t)j * vbytes));
            h->basis.push_back(std::move(c));
        }
        if (vp.size() > vp_dev) {
            HIPCHK(hipMemcpyAsync(vt + vp_dev, vp.data() + vp_dev, (vp.size() - vp_dev) * sizeof(double*),
                                  hipMemcpyHostToDevice, st));
            vp_dev = vp.size();
        }
        return MOED_OK;
    };
    if (int rc = reserve(2)) return rc;
    double* W = g[EG_W].as<double>();
    for (int p = 0; p < P; ++p)
        if (!pers[p])
            HIPCHK(launch_fill_start((0x5eed0000ull + (uint64_t)secs[p].first * 131 + secs[p].second) ^ seed_salt,
                                     dim[p], vp[0] + off[p], st));
    HIPCHK(launch_eig_segdot(e, vp[0], vp[0], d_nrm, ldh, 1, st));  // nrm[p][0] = |u_0|

    HvBatchArgs hb{};
    hb.prob = g[EG_DESC].as<BatchProbDev>();
    hb.row_prob = d_row_prob;
    hb.row_id = d_row_id;
    hb.y = W;
    hb.ek = h->d_ek.as<double>();
    hb.slotc = h->d_slotc.as<double>();
    hb.U = h->U;
    hb.UmJ3 = h->UmJ3;
    hb.Um2J = h->Um2J;
    hb.J = h->J;
    hb.scale_stride = ldh;
    hb.skip = d_done;

    std::vector<int32_t> done_h(P, 0), mfin_h(P, 0);
    std::vector<double> al_h((size_t)P * ldh), be_h((size_t)P * ldh);
    // pinned: the host's view of the done flags, then kFlagRing copies in flight
    const size_t npin = (size_t)(kFlagRing + 1) * P;
    if (h->pinned_n < npin) {
        if (h->pinned_i32) (void)hipHostFree(h->pinned_i32);
        h->pinned_i32 = nullptr;
        h->pinned_n = 0;
        HIPCHK(hipHostMalloc((void**)&h->pinned_i32, npin * sizeof(int32_t), hipHostMallocDefault));
        h->pinned_n = npin;
    }
    int32_t* flags = h->pinned_i32;
    int32_t* ring = h->pinned_i32 + P;
    for (int p = 0; p < P; ++p) flags[p] = pers[p];
    int ntest = 0;  // device convergence tests launched
    int active = P - (int)pqueue.size();
    if (!pqueue.empty()) {
        // the whole iteration of every queued problem in one launch; its basis is allocated up front
        if (int rc = reserve(pers_M + 1)) return rc;
        std::vector<uint64_t> seeds(P);
        for (int p = 0; p < P; ++p)
            seeds[p] = (0x5eed0000ull + (uint64_t)secs[p].first * 131 + secs[p].second) ^ seed_salt;
        const int32_t zero2[2] = {0, 0};
        HIPCHK(hipMemcpyAsync(d_pqueue, pqueue.data(), pqueue.size() * sizeof(int32_t), hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync(d_head, zero2, sizeof(zero2), hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync(d_seed, seeds.data(), P * sizeof(uint64_t), hipMemcpyHostToDevice, st));
        PersistArgs pa{};
        pa.e = e;
        pa.prob = g[EG_DESC].as<BatchProbDev>();
        pa.V = vt;
        pa.W = W;
        pa.ek = h->d_ek.as<double>();
        pa.slotc = h->d_slotc.as<double>();
        pa.U = h->U;
        pa.UmJ3 = h->UmJ3;
        pa.Um2J = h->Um2J;
        pa.J = h->J;
        pa.seed = reinterpret_cast<const uint64_t*>(d_seed);
        pa.order = d_pqueue;
        pa.nq = (int32_t)pqueue.size();
        pa.head = d_head;
        pa.nev = nev;
        pa.max_nU = pers_nU;
        pa.tol = h->gs_tol;
        pa.zs = d_zs;
        const int grid = std::max(1, std::min((int)pqueue.size(), h->ncu));
        // on the second stream, concurrent with the step-by-step iteration of the larger problems
        // (the two touch disjoint rows of every table and disjoint ranges of every vector)
        HIPCHK(hipEventRecord(h->ev_fork, st));
        HIPCHK(hipStreamWaitEvent(h->stream2, h->ev_fork, 0));
        HIPCHK(launch_eig_persistent(pa, grid, h->norb, (int)h->slots.size(), h->stream2));
        HIPCHK(hipEventRecord(h->ev_join, h->stream2));
    }
    for (int k = 0; k < M && active > 0; ++k) {
        if (int rc = reserve(k + 2)) return rc;  // step k writes vector k + 1
        const double* Vk = vp[k];
        hb.x = Vk;
        hb.scale = d_nrm + k;
        HIPCHK(launch_hv_batch(hb, nr, max_nU, h->norb, (int)h->slots.size(), st));
        for (int p = 0; p < P; ++p) {
            if (dim[p] <= kSmall || flags[p]) continue;  // flags as the host last read them
            int rc = hv_dev(h, S[p], Vk + off[p], W + off[p], nullptr, d_nrm + (size_t)p * ldh + k, d_done + p, st);
            if (rc) return rc;
        }
        if (defl) HIPCHK(launch_eig_deflate(e, defl->L, defl->nl, defl->sigma, Vk, W, k, st));
        HIPCHK(launch_eig_pass(e, vt, vp[k + 1], W, k, 0, 0.5, 1e-14, st));  // three-term part
        HIPCHK(launch_eig_pass(e, vt, vp[k + 1], W, k, 1, 0.5, 1e-14, st));  // whole basis
        HIPCHK(launch_eig_pass(e, vt, vp[k + 1], W, k, 2, 0.5, 1e-14, st));  // again where DGKS asks
        const int m = k + 1;
        if (m < 20 || m % 5 == 0 || m == M) {
            if (P > 1) {
                // several problems: the convergence test of every running problem on the device, on
                // the third stream beside the steps that follow; its verdicts enter done / mfinal in
                // stream order at the next test (a problem that passed at m stops with mfinal = m
                // whatever steps ran meanwhile: they write nothing below m).  The host reads the done
                // flags one test behind, so the steps stay queued on the GPU; steps queued after
                // every problem has stopped do nothing.
                if (ntest > 0) {
                    HIPCHK(hipStreamWaitEvent(st, h->ev_convd, 0));
                    HIPCHK(launch_eig_conv_apply(e, st));
                }
                const int slot = ntest % kFlagRing;
                HIPCHK(hipMemcpyAsync(ring + (size_t)slot * P, d_done, P * sizeof(int32_t), hipMemcpyDeviceToHost, st));
                HIPCHK(hipEventRecord(h->ev_flags[slot], st));
                HIPCHK(hipEventRecord(h->ev_step, st));
                HIPCHK(hipStreamWaitEvent(h->stream3, h->ev_step, 0));
                HIPCHK(launch_eig_conv(e, m, nev, h->gs_tol, d_zs, h->stream3));
                HIPCHK(hipEventRecord(h->ev_convd, h->stream3));
                ++ntest;
                if (ntest >= 2) {
                    const int ps = (ntest - 2) % kFlagRing;
                    HIPCHK(hipEventSynchronize(h->ev_flags[ps]));
                    std::memcpy(flags, ring + (size_t)ps * P, P * sizeof(int32_t));
                    active = 0;
                    for (int p = 0; p < P; ++p) active += flags[p] == 0;
                }
            } else {
                // one problem: its tridiagonal QL on the host (microseconds) beats one wave on the device
                HIPCHK(hipMemcpyAsync(done_h.data(), d_done, P * sizeof(int32_t), hipMemcpyDeviceToHost, st));
                HIPCHK(hipMemcpyAsync(mfin_h.data(), d_mfinal, P * sizeof(int32_t), hipMemcpyDeviceToHost, st));
                HIPCHK(hipMemcpyAsync(al_h.data(), d_alpha, al_h.size() * sizeof(double), hipMemcpyDeviceToHost, st));
                HIPCHK(hipMemcpyAsync(be_h.data(), d_beta, be_h.size() * sizeof(double), hipMemcpyDeviceToHost, st));
                HIPCHK(hipStreamSynchronize(st));
                // convergence of the lowest nev Ritz pairs, every active problem in parallel
                std::vector<int> conv(P, 0), qlbad(P, 0);
                parallel_for(P, [&](int p) {
                    if (done_h[p] || m < nev) return;
                    std::vector<double> d(al_h.begin() + (size_t)p * ldh, al_h.begin() + (size_t)p * ldh + m);
                    std::vector<double> ee(be_h.begin() + (size_t)p * ldh, be_h.begin() + (size_t)p * ldh + (m - 1));
                    std::vector<double> Z((size_t)m, 0.0);
                    Z[m - 1] = 1.0;
                    if (!tridiag_ql(d, ee, 1, &Z)) {
                        qlbad[p] = 1;
                        return;
                    }
                    const double bk = be_h[(size_t)p * ldh + m - 1];
                    bool ok = true;
                    const double above = defl ? defl->above_h[p] : INFINITY;
                    for (int i = 0; i < std::min(nev, m) && ok; ++i) {
                        const double res = std::fabs(bk * Z[i]);
                        const bool pass = res <= h->gs_tol * std::max(std::fabs(d[i]), 1e-300) ||
                                          (res <= kAboveTol * std::max(std::fabs(d[i]), 1.0) && d[i] - res > above);
                        if (!pass) ok = false;
                    }
                    conv[p] = ok;
                });
                bool changed = false;
                active = 0;
                for (int p = 0; p < P; ++p) {
                    if (qlbad[p]) return fail(h, MOED_ECONV, "tridiagonal QL did not converge");
                    if (done_h[p]) continue;
                    if (conv[p]) {
                        done_h[p] = 1;
                        mfin_h[p] = m;
                        changed = true;
                        continue;
                    }
                    ++active;
                }
                if (changed) {
                    HIPCHK(hipMemcpyAsync(d_done, done_h.data(), P * sizeof(int32_t), hipMemcpyHostToDevice, st));
                    HIPCHK(hipMemcpyAsync(d_mfinal, mfin_h.data(), P * sizeof(int32_t), hipMemcpyHostToDevice, st));
                }
                flags[0] = done_h[0];
            }
        }
    }
    if (ntest > 0) {  // the last test's verdicts
        HIPCHK(hipStreamWaitEvent(st, h->ev_convd, 0));
        HIPCHK(launch_eig_conv_apply(e, st));
    }
    if (!pqueue.empty()) HIPCHK(hipStreamWaitEvent(st, h->ev_join, 0));
    // final Ritz pairs: T_m of every problem on the host (in parallel), S = Z / |u_k|
    HIPCHK(hipMemcpyAsync(done_h.data(), d_done, P * sizeof(int32_t), hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(mfin_h.data(), d_mfinal, P * sizeof(int32_t), hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(al_h.data(), d_alpha, al_h.size() * sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(be_h.data(), d_beta, be_h.size() * sizeof(double), hipMemcpyDeviceToHost, st));
    std::vector<double> nrm_h((size_t)P * ldh);
    HIPCHK(hipMemcpyAsync(nrm_h.data(), d_nrm, nrm_h.size() * sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    std::vector<double> S_h((size_t)P * nev * ldh, 0.0);
    std::vector<int32_t> nf(P, 0);
    std::vector<int> ok(P, 1);
    // eigenvalues of T_m by QL, the wanted eigenvectors by inverse iteration
    auto ritz = [&](int p) {
        const int m = mfin_h[p];
        const double* al = al_h.data() + (size_t)p * ldh;
        const double* be = be_h.data() + (size_t)p * ldh;
        std::vector<double> d(al, al + m);
        if (!tridiag_ql(d, std::vector<double>(be, be + (m - 1)), 0, nullptr)) {
            ok[p] = 0;
            return;
        }
        nf[p] = std::min(nev, m);
        std::vector<double> Z;
        tridiag_vectors(al, be, m, d.data(), nf[p], Z);
        out[p]->ev.assign(nev, DBL_MAX);
        for (int i = 0; i < nf[p]; ++i) {
            out[p]->ev[i] = d[i];
            for (int kk = 0; kk < m; ++kk)
                S_h[((size_t)p * nev + i) * ldh + kk] = Z[(size_t)i * m + kk] / nrm_h[(size_t)p * ldh + kk];
        }
    };
    parallel_for(P, ritz);
    for (int p = 0; p < P; ++p)
        if (!ok[p]) return fail(h, MOED_ECONV, "tridiagonal QL did not converge");
    std::vector<double*> outp(P);
    for (int p = 0; p < P; ++p) {
        out[p]->nu = secs[p].first;
        out[p]->nd = secs[p].second;
        out[p]->D = dim[p];
        out[p]->nfound = nf[p];
        HIPCHK(out[p]->vec.ensure((size_t)std::max(nf[p], 1) * dim[p] * sizeof(double)));
        outp[p] = out[p]->vec.as<double>();
    }
    HIPCHK(hipMemcpyAsync(d_S, S_h.data(), S_h.size() * sizeof(double), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(d_nf, nf.data(), P * sizeof(int32_t), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(g[EG_OUT].p, outp.data(), P * sizeof(double*), hipMemcpyHostToDevice, st));
    HIPCHK(launch_eig_ritz(e, vt, d_S, ldh, nev, d_nf, g[EG_OUT].as<double*>(), d_n2, st));
    HIPCHK(hipStreamSynchronize(st));
    return MOED_OK;
}

// Levels a single-start Lanczos run can miss: its Krylov space holds one direction per eigenspace,
// so a degenerate level (identical orbitals, for one) is found once.  For every sector solved by
// Lanczos with nev >= 2: run the iteration again, from another start vector, on H' = H + sigma
// sum_i y_i y_i^T with the sector's Ritz vectors y_i shifted out of the way (sigma > the spread of
// the found values), for its lowest Ritz pair only.  A value below the sector's nev-th (by more than
// 1e-9 relative) is a missed level: it joins the sector's list with its vector, and the check runs
// again; a sector is done when its check finds nothing below.  Every sector's check runs in one
// multi-problem run, which stops a sector's iteration as soon as its lowest Ritz value lies above
// the sector's nev-th level with its whole residual interval.
// a level found by the check joins the sector's list when it lies below this
double insert_below(const SectorEig& E, int nev) {
    if (E.nfound < nev) return DBL_MAX;
    const double top = E.ev[nev - 1];
    return top - 1e-9 * std::max(1.0, std::fabs(top));
}

int degeneracy_check(moed_handle* h, int nev, std::vector<int> cand) {
    hipStream_t st = h->stream;
    for (int round = 0; round < 2 * nev + 2 && !cand.empty(); ++round) {
        const int P = (int)cand.size();
        std::vector<std::pair<int, int>> secs(P);
        std::vector<int64_t> off(P);
        int64_t dtot = 0;
        int nl = 1;
        for (int p = 0; p < P; ++p) {
            const SectorEig& E = *h->eig[cand[p]];
            secs[p] = {E.nu, E.nd};
            off[p] = dtot;  // the engine's packed layout: every problem on an even element
            dtot += E.D + (E.D & 1);
            nl = std::max(nl, E.nfound);
        }
        DevBuf& Lb = h->ws[WS_MISC];
        HIPCHK(Lb.ensure((size_t)nl * dtot * sizeof(double) + (size_t)nl * sizeof(double*) + 2 * P * sizeof(double)));
        double* L = Lb.as<double>();
        double** Lt = reinterpret_cast<double**>(L + (size_t)nl * dtot);
        double* sig = reinterpret_cast<double*>(Lt + nl);
        double* abv = sig + P;
        HIPCHK(launch_fill(0.0, (int64_t)nl * dtot, L, st));
        std::vector<double*> lt(nl);
        std::vector<double> sg(P), ab(P);
        for (int i = 0; i < nl; ++i) lt[i] = L + (size_t)i * dtot;
        for (int p = 0; p < P; ++p) {
            const SectorEig& E = *h->eig[cand[p]];
            for (int i = 0; i < E.nfound; ++i)
                HIPCHK(hipMemcpyAsync(lt[i] + off[p], E.vec.as<double>() + (size_t)i * E.D, E.D * sizeof(double),
                                      hipMemcpyDeviceToDevice, st));
            const double lo = E.ev[0], top = E.ev[E.nfound - 1];
            sg[p] = (top - lo) + std::max(1.0, std::fabs(top)) + std::max(1.0, std::fabs(lo));
            ab[p] = insert_below(E, nev);
        }
        HIPCHK(hipMemcpyAsync(Lt, lt.data(), nl * sizeof(double*), hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync(sig, sg.data(), P * sizeof(double), hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync(abv, ab.data(), P * sizeof(double), hipMemcpyHostToDevice, st));
        std::vector<std::unique_ptr<SectorEig>> chk(P);
        std::vector<SectorEig*> outs(P);
        for (int p = 0; p < P; ++p) {
            chk[p] = std::make_unique<SectorEig>();
            outs[p] = chk[p].get();
        }
        const Deflation d{Lt, nl, sig, abv, ab.data()};
        if (int rc = eig_batch(h, secs, 1, outs, false, &d, 0x9e3779b97f4a7c15ull * (uint64_t)(round + 1))) return rc;
        std::vector<int> next;
        for (int p = 0; p < P; ++p) {
            SectorEig& E = *h->eig[cand[p]];
            const SectorEig& C = *chk[p];
            if (C.nfound < 1) continue;
            const double th = C.ev[0];
            if (!(th < ab[p])) continue;
            // insert (th, its vector) in order; the list keeps the lowest nev
            const int nf = std::min(nev, E.nfound + 1);
            int j = 0;
            while (j < E.nfound && E.ev[j] <= th) ++j;
            DevBuf nv;
            HIPCHK(nv.ensure((size_t)nf * E.D * sizeof(double)));
            double* dst = nv.as<double>();
            const double* src = E.vec.as<double>();
            const size_t vb = E.D * sizeof(double);
            if (j > 0) HIPCHK(hipMemcpyAsync(dst, src, (size_t)j * vb, hipMemcpyDeviceToDevice, st));
            HIPCHK(hipMemcpyAsync(dst + (size_t)j * E.D, C.vec.as<double>(), vb, hipMemcpyDeviceToDevice, st));
            if (nf - j - 1 > 0)
                HIPCHK(hipMemcpyAsync(dst + (size_t)(j + 1) * E.D, src + (size_t)j * E.D, (size_t)(nf - j - 1) * vb,
                                      hipMemcpyDeviceToDevice, st));
            HIPCHK(hipStreamSynchronize(st));
            std::swap(E.vec.p, nv.p);
            std::swap(E.vec.bytes, nv.bytes);
            std::vector<double> ev(nev, DBL_MAX);
            for (int i = 0, q = 0; i < nf; ++i) ev[i] = i == j ? th : E.ev[q++];
            E.ev = ev;
            E.nfound = nf;
            next.push_back(cand[p]);
        }
        cand = next;
    }
    return MOED_OK;
}

// Every sector of at most 64 states of a solve at once: their matrices from one batched product
// (each column of each such sector a problem of hv_batch_kernel applied to a unit vector; its terms
// per element are those of the two-pass product), then cyclic Jacobi on the host, the sectors in
// parallel.
int eig_dense_batch(moed_handle* h, const std::vector<int>& secs, int nev) {
    if (secs.empty()) return MOED_OK;
    hipStream_t st = h->stream;
    const int ns = (int)secs.size();
    std::vector<Sector> S(ns);
    std::vector<BatchProbDev> desc;
    std::vector<int32_t> row_prob, row_id;
    std::vector<int64_t> off0(ns);
    int64_t tot = 0;
    int max_nU = 1;
    for (int q = 0; q < ns; ++q) {
        const int is = secs[q];
        if (int rc = get_sector(h, h->nup[is], h->nelec[is] - h->nup[is], &S[q])) return rc;
        const int64_t D = S[q].D, ld = D + (D & 1);  // every column on an even element
        off0[q] = tot;
        for (int64_t j = 0; j < D; ++j) {
            const int pid = (int)desc.size();
            desc.push_back({S[q].up->dev(), S[q].dn->dev(), tot + j * ld});
            for (int id = 0; id < S[q].dn->count; ++id) {
                row_prob.push_back(pid);
                row_id.push_back(id);
            }
        }
        tot += D * ld;
        max_nU = std::max(max_nU, S[q].up->count);
    }
    std::vector<double> X((size_t)tot, 0.0);
    for (int q = 0; q < ns; ++q) {
        const int64_t D = S[q].D, ld = D + (D & 1);
        for (int64_t j = 0; j < D; ++j) X[(size_t)(off0[q] + j * ld + j)] = 1.0;
    }
    const int nr = (int)row_prob.size();
    DevBuf& Xd = h->ws[WS_X];
    DevBuf& Yd = h->ws[WS_Y];
    DevBuf& Md = h->ws[WS_MISC];
    HIPCHK(Xd.ensure((size_t)tot * sizeof(double)));
    HIPCHK(Yd.ensure((size_t)tot * sizeof(double)));
    const size_t dbytes = desc.size() * sizeof(BatchProbDev);
    HIPCHK(Md.ensure(dbytes + 2 * (size_t)nr * sizeof(int32_t)));
    int32_t* d_rp = reinterpret_cast<int32_t*>(static_cast<char*>(Md.p) + dbytes);
    HIPCHK(hipMemcpyAsync(Xd.p, X.data(), X.size() * sizeof(double), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(Md.p, desc.data(), dbytes, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(d_rp, row_prob.data(), nr * sizeof(int32_t), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(d_rp + nr, row_id.data(), nr * sizeof(int32_t), hipMemcpyHostToDevice, st));
    HvBatchArgs hb{};
    hb.prob = Md.as<BatchProbDev>();
    hb.row_prob = d_rp;
    hb.row_id = d_rp + nr;
    hb.x = Xd.as<double>();
    hb.y = Yd.as<double>();
    hb.ek = h->d_ek.as<double>();
    hb.slotc = h->d_slotc.as<double>();
    hb.U = h->U;
    hb.UmJ3 = h->UmJ3;
    hb.Um2J = h->Um2J;
    hb.J = h->J;
    HIPCHK(launch_hv_batch(hb, nr, max_nU, h->norb, (int)h->slots.size(), st));
    std::vector<double> Y((size_t)tot);
    HIPCHK(hipMemcpyAsync(Y.data(), Yd.p, Y.size() * sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    std::vector<std::vector<double>> vecs(ns);
    parallel_for(ns, [&](int q) {
        SectorEig& E = *h->eig[secs[q]];
        const int D = (int)S[q].D, ld = D + (D & 1);
        std::vector<double> Hm((size_t)D * D), w, Vm;  // column j of H; H is symmetric: row-major
        for (int j = 0; j < D; ++j)
            std::copy(Y.begin() + off0[q] + (size_t)j * ld, Y.begin() + off0[q] + (size_t)j * ld + D,
                      Hm.begin() + (size_t)j * D);
        jacobi_eig(Hm, D, w, Vm);
        const int nf = std::min(nev, D);
        E.nu = S[q].nu;
        E.nd = S[q].nd;
        E.D = D;
        E.nfound = nf;
        E.ev.assign(nev, DBL_MAX);
        vecs[q].resize((size_t)nf * D);
        for (int i = 0; i < nf; ++i) {
            E.ev[i] = w[i];
            for (int r = 0; r < D; ++r) vecs[q][(size_t)i * D + r] = Vm[(size_t)r * D + i];
        }
    });
    for (int q = 0; q < ns; ++q) {
        SectorEig& E = *h->eig[secs[q]];
        HIPCHK(E.vec.ensure(vecs[q].size() * sizeof(double)));
        HIPCHK(hipMemcpyAsync(E.vec.p, vecs[q].data(), vecs[q].size() * sizeof(double), hipMemcpyHostToDevice, st));
    }
    HIPCHK(hipStreamSynchronize(st));
    return MOED_OK;
}

int eig_sector(moed_handle* h, int nu, int nd, int nev, SectorEig& E) {
    Sector s;
    int rc = get_sector(h, nu, nd, &s);
    if (rc) return rc;
    E.nu = nu;
    E.nd = nd;
    E.D = s.D;
    std::vector<SectorEig*> out{&E};
    return eig_batch(h, {{nu, nd}}, nev, out);
}

// SORT of func.f90:301-353: heapsort returning a 1-based index, with the same sift order so that
// ties come out in the same order.
void heapsort_index(int n, const double* a, int* indx) {
    const double* A = a - 1;
    int* X = indx - 1;
    if (n == 1) {
        X[1] = 1;
        return;
    }
    for (int j = 1; j <= n; ++j) X[j] = j;
    int l = n / 2 + 1, ir = n;
    while (true) {
        int indxt;
        double q;
        if (l > 1) {
            --l;
            indxt = X[l];
            q = A[indxt];
        } else {
            indxt = X[ir];
            q = A[indxt];
            X[ir] = X[1];
            if (--ir == 1) {
                X[1] = indxt;
                return;
            }
        }
        int i = l, j = l + l;
        while (j <= ir) {
            if (j < ir && A[X[j]] < A[X[j + 1]]) ++j;
            if (q < A[X[j]]) {
                X[i] = X[j];
                i = j;
                j += j;
            } else {
                j = ir + 1;
            }
        }
        X[i] = indxt;
    }
}

// boltzmann_factor (ed_solver.F90:193-219)
void boltzmann(const std::vector<double>& e, double beta, std::vector<double>& p) {
    const int n = (int)e.size();
    p.assign(n, 0.0);
    for (int i = 0; i < n; ++i) {
        double Z = 0.0;
        bool cut = false;
        for (int k = 0; k < n; ++k) {
            if (-beta * (e[k] - e[i]) > 10.0) {
                cut = true;
                break;
            }
            Z = Z + std::exp(-beta * (e[k] - e[i]));
        }
        p[i] = cut ? 0.0 : 1.0 / Z;
    }
}

int green_diag(moed_handle* h, const Sector& s, const double* psi, double E, double weight, int iorb, int ispin,
               const double* d_omega, int nw, double* d_G, GreenRecord& rec) {
    hipStream_t st = h->stream;
    const int compat = (h->flags & MOED_F_EXACT_CMPLX) ? 0 : 1;
    const int nstep = h->nstep;
    double normsq = 0.0;
    for (int pm = 1; pm <= 2; ++pm) {
        int nuo, ndo;
        target_sector(pm, ispin, s.nu, s.nd, &nuo, &ndo);
        std::vector<double> a(nstep + 1, 0.0), b(nstep + 1, 0.0);
        int nsc = 0;
        bool have = nuo >= 0 && ndo >= 0 && nuo <= h->nsite && ndo <= h->nsite;
        double n2 = 0.0;
        if (have) {
            Sector so;
            int rc = get_sector(h, nuo, ndo, &so);
            if (rc) return rc;
            DevBuf& X = h->ws[WS_X];
            HIPCHK(X.ensure((size_t)so.D * sizeof(double)));
            rc = apply_c_dev(h, pm, iorb, ispin, s, so, psi, X.as<double>(), st);
            if (rc) return rc;
            rc = dot_dev(h, X.as<double>(), X.as<double>(), so.D, &n2, st);
            if (rc) return rc;
            if (n2 > 0.0) {
                rc = lanczos_ref_dev(h, so, X.as<double>(), nstep, &nsc, a.data(), b.data());
                if (rc) return rc;
            } else {
                have = false;  // c(+)|psi> = 0: no contribution
            }
        }
        if (pm == 1) normsq = n2;
        b[0] = pm == 1 ? normsq : 1.0 - normsq;  // ed_green.F90:158, :174
        if (have) {
            DevBuf& AB = h->ws[WS_MISC];
            HIPCHK(AB.ensure((size_t)2 * (nstep + 1) * sizeof(double)));
            HIPCHK(hipMemcpyAsync(AB.p, a.data(), (nstep + 1) * sizeof(double), hipMemcpyHostToDevice, st));
            HIPCHK(hipMemcpyAsync(AB.as<double>() + nstep + 1, b.data(), (nstep + 1) * sizeof(double),
                                  hipMemcpyHostToDevice, st));
            CfArgs cf{AB.as<double>(), AB.as<double>() + nstep + 1, nsc, pm == 1 ? 1 : -1, E, weight, d_omega, nw,
                      compat, d_G + 2 * (iorb - 1), h->norb};
            HIPCHK(launch_cf(cf, st));
            HIPCHK(hipStreamSynchronize(st));
        }
        if (pm == 1) {
            rec.nsp = nsc;
            rec.ap = a;
            rec.bp = b;
        } else {
            rec.nsn = nsc;
            rec.an = a;
            rec.bn = b;
        }
    }
    return MOED_OK;
}

// apply_c's map for one chain (as apply_c_dev builds it): per target string of the changed spin,
// source rank | parity << 30, or -1
void apply_c_map(int pm, int iorb, int ispin, const Sector& in, const Sector& out, int32_t* map) {
    const int o = iorb - 1;
    const StrTable* tin = ispin == 1 ? in.up : in.dn;
    const StrTable* tout = ispin == 1 ? out.up : out.dn;
    for (int r = 0; r < tout->count; ++r) {
        map[r] = -1;
        const uint32_t s = tout->str[r];
        const bool occ = (s >> o) & 1u;
        if ((pm == 1) != occ) continue;
        const uint32_t src = s ^ (1u << o);
        int par = __builtin_popcount(s & ((1u << o) - 1u)) & 1;
        if (ispin == 2) par ^= (in.nu & 1);
        map[r] = tin->rank[src] | (par << 30);
    }
}

// ed_calc_green_ftn (ed_green.F90:60-134) with every chain -- kept state x spin x orbital x (c+, c)
// -- in one multi-problem run of the reference recurrence: the start vectors c(+)|psi> of all
// chains packed end to end, one launch per operation for every chain, no host round trip per step
// (one to read every chain's |c+ psi|^2 before the recurrence, one for the coefficients after it).
// The continued fractions are then summed into G per orbital in the sequential path's order.
int green_batch(moed_handle* h, int nw, const double* d_omega, double* d_G) {
    hipStream_t st = h->stream;
    const int nstep = h->nstep, norb = h->norb;
    const int compat = (h->flags & MOED_F_EXACT_CMPLX) ? 0 : 1;
    struct Chain {
        int iev, iorb, ispin, pm;
        double E, weight, p;
        bool even, valid;
        Sector src, dst;
        const double* psi;
    };
    std::vector<Chain> ch;
    for (size_t iev = 0; iev < h->kept_val.size(); ++iev) {
        const int is = h->kept_sector[iev] - 1, il = h->kept_level[iev] - 1;
        const int nu = h->nup[is], nd = h->nelec[is] - h->nup[is];
        const bool even = nd == nu;
        Sector s;
        if (int rc = get_sector(h, nu, nd, &s)) return rc;
        const SectorEig& E = *h->eig[is];
        if (il >= E.nfound) return fail(h, MOED_ESTATE, "kept state has no eigenvector");
        for (int ispin = 1; ispin <= (even ? 1 : 2); ++ispin)
            for (int iorb = 1; iorb <= norb; ++iorb)
                for (int pm = 1; pm <= 2; ++pm) {
                    Chain c{};
                    c.iev = (int)iev;
                    c.iorb = iorb;
                    c.ispin = ispin;
                    c.pm = pm;
                    c.E = h->kept_val[iev];
                    c.p = h->kept_prob[iev];
                    c.weight = h->kept_prob[iev] * (even ? 1.0 : 0.5);
                    c.even = even;
                    c.src = s;
                    c.psi = E.vec.as<double>() + (size_t)il * E.D;
                    int nuo, ndo;
                    target_sector(pm, ispin, nu, nd, &nuo, &ndo);
                    c.valid = nuo >= 0 && ndo >= 0 && nuo <= h->nsite && ndo <= h->nsite;
                    if (c.valid)
                        if (int rc = get_sector(h, nuo, ndo, &c.dst)) return rc;
                    ch.push_back(c);
                }
    }
    const int P = (int)ch.size();
    if (P == 0) return MOED_OK;
    // packed layout (every chain starts on an even element), items, H v row items, apply_c maps
    constexpr int64_t kSmall = (int64_t)1 << 20;
    std::vector<int64_t> off(P, 0), dim(P, 0), item_start;
    std::vector<int32_t> item_prob, item_len, pitem0(P), pnitem(P), row_prob, row_id, map_off(P, 0);
    std::vector<BatchProbDev> desc(P);
    int64_t dtot = 0, nmap = 0;
    int max_nU = 1;
    for (int p = 0; p < P; ++p) {
        Chain& c = ch[p];
        pitem0[p] = (int32_t)item_prob.size();
        if (c.valid) {
            dim[p] = c.dst.D;
            off[p] = dtot;
            dtot += dim[p] + (dim[p] & 1);
            for (int64_t s0 = 0; s0 < dim[p]; s0 += kEigChunk) {
                item_prob.push_back(p);
                item_start.push_back(off[p] + s0);
                item_len.push_back((int32_t)std::min<int64_t>(kEigChunk, dim[p] - s0));
            }
            desc[p] = {c.dst.up->dev(), c.dst.dn->dev(), off[p]};
            if (dim[p] <= kSmall) {
                for (int id = 0; id < c.dst.dn->count; ++id) {
                    row_prob.push_back(p);
                    row_id.push_back(id);
                }
                max_nU = std::max(max_nU, c.dst.up->count);
            }
            map_off[p] = (int32_t)nmap;
            nmap += c.ispin == 1 ? c.dst.up->count : c.dst.dn->count;
        } else {
            desc[p] = BatchProbDev{};
        }
        pnitem[p] = (int32_t)item_prob.size() - pitem0[p];
    }
    const int ni = (int)item_prob.size(), nr = (int)row_prob.size();
    std::vector<int32_t> maps((size_t)std::max<int64_t>(nmap, 1));
    for (int p = 0; p < P; ++p)
        if (ch[p].valid) apply_c_map(ch[p].pm, ch[p].iorb, ch[p].ispin, ch[p].src, ch[p].dst, maps.data() + map_off[p]);
    const int ldh = nstep + 1;
    // device buffers (handle-owned, grown as needed)
    DevBuf* g = h->gbuf;
    HIPCHK(g[0].ensure((size_t)std::max<int64_t>(4 * dtot, 1) * sizeof(double)));  // u0, u1, u2, w
    const size_t n32 = (size_t)2 * ni + 2 * (size_t)P + 2 * (size_t)nr + 8 * (size_t)P + maps.size();
    HIPCHK(g[1].ensure(n32 * sizeof(int32_t)));
    HIPCHK(g[2].ensure(((size_t)ni + 2 * P + 1) * sizeof(int64_t)));
    const size_t ndbl = 2 * (size_t)P * ldh + (size_t)std::max(ni, 1) + 4 * (size_t)P;
    HIPCHK(g[3].ensure(ndbl * sizeof(double)));
    HIPCHK(g[4].ensure((size_t)P * sizeof(BatchProbDev)));
    double* U[3] = {g[0].as<double>(), g[0].as<double>() + dtot, g[0].as<double>() + 2 * dtot};
    double* W = g[0].as<double>() + 3 * dtot;
    int32_t* i32 = g[1].as<int32_t>();
    int32_t *d_item_prob = i32, *d_item_len = d_item_prob + ni, *d_pitem0 = d_item_len + ni, *d_pnitem = d_pitem0 + P,
            *d_row_prob = d_pnitem + P, *d_row_id = d_row_prob + nr, *d_stop = d_row_id + nr, *d_nsc = d_stop + P,
            *d_sign = d_nsc + P, *d_iorb = d_sign + P, *d_use = d_iorb + P, *d_maps = d_use + P;
    int64_t *d_item_start = g[2].as<int64_t>();
    double* dbl = g[3].as<double>();
    double *d_A = dbl, *d_B = d_A + (size_t)P * ldh, *d_part = d_B + (size_t)P * ldh, *d_b0 = d_part + std::max(ni, 1),
           *d_E = d_b0 + P, *d_w = d_E + P, *d_n2 = d_w + P;
    if (ni > 0) {
        HIPCHK(hipMemcpyAsync(d_item_prob, item_prob.data(), ni * sizeof(int32_t), hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync(d_item_len, item_len.data(), ni * sizeof(int32_t), hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync(d_item_start, item_start.data(), ni * sizeof(int64_t), hipMemcpyHostToDevice, st));
    }
    HIPCHK(hipMemcpyAsync(d_pitem0, pitem0.data(), P * sizeof(int32_t), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(d_pnitem, pnitem.data(), P * sizeof(int32_t), hipMemcpyHostToDevice, st));
    if (nr > 0) {
        HIPCHK(hipMemcpyAsync(d_row_prob, row_prob.data(), nr * sizeof(int32_t), hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync(d_row_id, row_id.data(), nr * sizeof(int32_t), hipMemcpyHostToDevice, st));
    }
    HIPCHK(hipMemcpyAsync(d_maps, maps.data(), maps.size() * sizeof(int32_t), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(g[4].p, desc.data(), P * sizeof(BatchProbDev), hipMemcpyHostToDevice, st));
    HIPCHK(launch_fill(0.0, (int64_t)2 * P * ldh, d_A, st));
    EigArgs e{};
    e.P = P;
    e.ni = ni;
    e.item_prob = d_item_prob;
    e.item_start = d_item_start;
    e.item_len = d_item_len;
    e.pitem0 = d_pitem0;
    e.pnitem = d_pnitem;
    e.part = d_part;
    // start vectors c(+)|psi> and their squared norms (one host round trip)
    for (int p = 0; p < P; ++p) {
        const Chain& c = ch[p];
        if (!c.valid) continue;
        ApplyCArgs a{c.psi, U[0] + off[p], d_maps + map_off[p], c.dst.up->count, c.dst.dn->count, c.src.up->count,
                     c.ispin == 1 ? 0 : 1};
        HIPCHK(launch_apply_c(a, st));
    }
    std::vector<int32_t> zero(P, 0);
    HIPCHK(hipMemcpyAsync(d_stop, zero.data(), P * sizeof(int32_t), hipMemcpyHostToDevice, st));
    if (ni > 0) HIPCHK(launch_eig_segdot(e, U[0], U[0], d_n2, 1, 0, st));
    std::vector<double> n2(P, 0.0);
    HIPCHK(hipMemcpyAsync(n2.data(), d_n2, P * sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    std::vector<int32_t> stop(P), nsc(P), sign(P), iorb(P), use(P);
    std::vector<double> b0(P), Ev(P), wv(P), normsq(P);
    for (int p = 0; p < P; ++p) {
        const Chain& c = ch[p];
        const double v = c.valid ? n2[p] : 0.0;
        const bool have = c.valid && v > 0.0;  // c(+)|psi> = 0: no contribution (green_diag)
        // b(0) of green_diag: |c+ psi|^2 for the particle chain, 1 - |c+ psi|^2 for the hole chain
        // (ed_green.F90:158, :174); the hole chain follows its particle chain
        normsq[p] = c.pm == 1 ? v : normsq[p - 1];
        b0[p] = c.pm == 1 ? normsq[p] : 1.0 - normsq[p];
        stop[p] = have ? 0 : 1;
        nsc[p] = have ? nstep : 0;
        use[p] = have;
        sign[p] = c.pm == 1 ? 1 : -1;
        iorb[p] = c.iorb - 1;
        Ev[p] = c.E;
        wv[p] = c.weight;
    }
    HIPCHK(hipMemcpyAsync(d_stop, stop.data(), P * sizeof(int32_t), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(d_nsc, nsc.data(), P * sizeof(int32_t), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(d_sign, sign.data(), P * sizeof(int32_t), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(d_iorb, iorb.data(), P * sizeof(int32_t), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(d_use, use.data(), P * sizeof(int32_t), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(d_b0, b0.data(), P * sizeof(double), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(d_E, Ev.data(), P * sizeof(double), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(d_w, wv.data(), P * sizeof(double), hipMemcpyHostToDevice, st));
    // H v of every chain: one launch for the sectors up to 2^20 elements, the two-pass product per
    // larger sector (each skipped once its chain has stopped)
    HvBatchArgs hb{};
    hb.prob = g[4].as<BatchProbDev>();
    hb.row_prob = d_row_prob;
    hb.row_id = d_row_id;
    hb.ek = h->d_ek.as<double>();
    hb.slotc = h->d_slotc.as<double>();
    hb.U = h->U;
    hb.UmJ3 = h->UmJ3;
    hb.Um2J = h->Um2J;
    hb.J = h->J;
    hb.skip = d_stop;
    auto hv_all = [&](const double* x) -> int {
        hb.x = x;
        hb.y = W;
        HIPCHK(launch_hv_batch(hb, nr, max_nU, h->norb, (int)h->slots.size(), st));
        for (int p = 0; p < P; ++p)
            if (use[p] && dim[p] > kSmall)
                if (int rc = hv_dev(h, ch[p].dst, x + off[p], W + off[p], nullptr, nullptr, d_stop + p, st)) return rc;
        return MOED_OK;
    };
    if (ni > 0) {
        // j = 0: B[0] = |u_0|^2, w = H u_0, A[0] = <u_0, w> / B[0]
        HIPCHK(launch_lzb_dot(e, U[0], U[0], d_stop, st));
        HIPCHK(launch_lzb_finish(e, 0, 0, d_A, d_B, ldh, d_stop, d_nsc, kBreakdown, st));
        if (int rc = hv_all(U[0])) return rc;
        HIPCHK(launch_lzb_dot(e, U[0], W, d_stop, st));
        HIPCHK(launch_lzb_finish(e, 0, 1, d_A, d_B, ldh, d_stop, d_nsc, kBreakdown, st));
        for (int j = 1; j <= nstep; ++j) {
            double* uj = U[j % 3];
            HIPCHK(launch_lzb_update(e, j, W, U[(j - 1) % 3], U[(j + 1) % 3], uj, d_A, d_B, ldh, d_stop, st));
            HIPCHK(launch_lzb_finish(e, j, 0, d_A, d_B, ldh, d_stop, d_nsc, kBreakdown, st));
            if (int rc = hv_all(uj)) return rc;
            HIPCHK(launch_lzb_dot(e, uj, W, d_stop, st));
            HIPCHK(launch_lzb_finish(e, j, 1, d_A, d_B, ldh, d_stop, d_nsc, kBreakdown, st));
        }
    }
    CfBatchArgs cf{P, ldh, norb, d_A, d_B, d_b0, d_nsc, d_sign, d_iorb, d_use, d_E, d_w, d_omega, nw, compat, d_G};
    HIPCHK(launch_cf_batch(cf, st));
    // the records (apbpanbn.dat order: per state, spin, orbital; particle then hole)
    std::vector<double> Ah((size_t)P * ldh), Bh((size_t)P * ldh);
    HIPCHK(hipMemcpyAsync(Ah.data(), d_A, Ah.size() * sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(Bh.data(), d_B, Bh.size() * sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(nsc.data(), d_nsc, P * sizeof(int32_t), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    h->records.clear();
    for (int p = 0; p < P; p += 2) {
        GreenRecord rec{};
        rec.iev = ch[p].iev + 1;
        rec.iorb = ch[p].iorb;
        rec.ispin = ch[p].ispin;
        rec.E = ch[p].E;
        rec.p = ch[p].p;
        rec.even = ch[p].even;
        for (int q = p; q < p + 2; ++q) {
            const int n = use[q] ? nsc[q] : 0;
            std::vector<double> a(nstep + 1, 0.0), b(nstep + 1, 0.0);
            if (use[q])
                for (int j = 0; j <= n; ++j) {
                    a[j] = Ah[(size_t)q * ldh + j];
                    b[j] = Bh[(size_t)q * ldh + j];
                }
            b[0] = b0[q];
            if (ch[q].pm == 1) {
                rec.nsp = n;
                rec.ap = std::move(a);
                rec.bp = std::move(b);
            } else {
                rec.nsn = n;
                rec.an = std::move(a);
                rec.bn = std::move(b);
            }
        }
        h->records.push_back(std::move(rec));
    }
    return MOED_OK;
}

}  // namespace

static FitCtx fit_ctx(moed_handle* h) {
    return FitCtx{h->stream, h->norb, h->nbath, (h->flags & MOED_F_EXACT_CMPLX) ? 0 : 1,
                  (h->flags & MOED_F_FIT_TREE) ? 1 : 0, h->fit_budget, nullptr, &h->err};
}

int fit_delta_host(moed_handle* h, int nx, const double* x, int nw, const double* omega, double* Dev) {
    (void)hipSetDevice(h->device);
    return fit_delta(fit_ctx(h), nx, x, nw, omega, Dev);
}

int fit_eval_host(moed_handle* h, int nfit, int nx, const double* x, int nw, const double* omega, const double* Dev,
                  double* f, double* g) {
    (void)hipSetDevice(h->device);
    return fit_eval(fit_ctx(h), nfit, nx, x, nw, omega, Dev, f, g);
}

int fit_minimize_host(moed_handle* h, int nfit, int nx, double* x, int nw, const double* omega, const double* Dev,
                      double ftol, double* fret, int* iter, int* status) {
    (void)hipSetDevice(h->device);
    FitCtx c = fit_ctx(h);
    h->fit_nfev.assign(nfit, 0);
    c.nfev = h->fit_nfev.data();
    return fit_minimize(c, nfit, nx, x, nw, omega, Dev, ftol, fret, iter, status);
}

// =================================================================== C ABI

extern "C" {

static void destroy_streams(moed_handle* h) {
    for (hipEvent_t& ev : h->ev_flags)
        if (ev) (void)hipEventDestroy(ev);
    for (hipEvent_t ev : {h->ev_convd, h->ev_step, h->ev_join, h->ev_fork})
        if (ev) (void)hipEventDestroy(ev);
    for (hipStream_t s : {h->stream3, h->stream2, h->stream})
        if (s) (void)hipStreamDestroy(s);
}

int moed_create(const moed_params* p, moed_handle** out) {
    if (!p || !out) return MOED_EARG;
    *out = nullptr;
    if (p->norb < 1 || p->norb > kMaxOrb || p->nsite < p->norb || p->nsite > kMaxSite || !p->ek || !p->vk)
        return MOED_EARG;
    if (p->nstep < 0 || p->nev < 0 || p->nw < 0 || p->nsector < 0) return MOED_EARG;
    auto* h = new moed_handle();
    h->device = p->device;
    h->norb = p->norb;
    h->nsite = p->nsite;
    h->nbath = p->nsite - p->norb;
    h->U = p->U;
    h->J = p->J;
    h->beta = p->beta;
    h->nstep = p->nstep;
    h->nev = p->nev;
    h->nw = p->nw;
    h->flags = p->flags;
    h->gs_tol = p->gs_tol > 0 ? p->gs_tol : 1e-12;
    h->gs_maxvec = p->gs_maxvec > 0 ? p->gs_maxvec : 400;
    if (p->tij) {
        h->has_tij = true;
        h->tij.assign(p->tij, p->tij + p->norb * p->norb);
    }
    for (int i = 0; i < p->nsector; ++i) {
        h->nelec.push_back(p->nelec[i]);
        h->nup.push_back(p->nup[i]);
    }
    interaction_constants(h->U, h->J, &h->UmJ3, &h->Um2J);
    hipError_t e = hipSetDevice(h->device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&h->stream2, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&h->ev_fork, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&h->ev_join, hipEventDisableTiming);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&h->stream3, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&h->ev_step, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&h->ev_convd, hipEventDisableTiming);
    for (int i = 0; i < kFlagRing && e == hipSuccess; ++i)
        e = hipEventCreateWithFlags(&h->ev_flags[i], hipEventDisableTiming);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&h->ncu, hipDeviceAttributeMultiprocessorCount, h->device);
    if (e == hipSuccess && fit_scratch_address(h->device, &h->fit_scratch) != MOED_OK) e = hipErrorInvalidValue;
    if (e != hipSuccess) {  // (device is current: fit_scratch_address looks up this GPU's array)
        destroy_streams(h);
        delete h;
        return e == hipErrorOutOfMemory ? MOED_ENOMEM : MOED_EHIP;
    }
    int rc = set_model(h, p->ek, p->vk);
    if (rc) {
        moed_destroy(h);
        return rc;
    }
    *out = h;
    return MOED_OK;
}

int moed_destroy(moed_handle* h) {
    if (!h) return MOED_OK;
    (void)hipSetDevice(h->device);
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    h->clear_graphs();
    h->tables.clear();
    h->eig.clear();
    h->basis.clear();
    for (auto& w : h->ws) w.release();
    if (h->pinned_i32) (void)hipHostFree(h->pinned_i32);
    h->d_ek.release();
    h->d_coef.release();
    if (h->stream2) (void)hipStreamSynchronize(h->stream2);
    if (h->stream3) (void)hipStreamSynchronize(h->stream3);
    destroy_streams(h);
    delete h;
    return MOED_OK;
}

int moed_last_error(const moed_handle* h, char* buf, int32_t len) {
    if (!h || !buf || len <= 0) return MOED_EARG;
    std::snprintf(buf, (size_t)len, "%s", h->err.c_str());
    return MOED_OK;
}

int moed_set_bath(moed_handle* h, const double* ek, const double* vk) {
    if (!h || !ek || !vk) return MOED_EARG;
    (void)hipSetDevice(h->device);
    return set_model(h, ek, vk);
}

int moed_sector_dim(moed_handle* h, int32_t nu, int32_t nd, int64_t* dim) {
    if (!h || !dim) return MOED_EARG;
    if (nu < 0 || nd < 0 || nu > h->nsite || nd > h->nsite) return fail(h, MOED_ESECTOR, "sector out of range");
    *dim = binom(h->nsite, nu) * binom(h->nsite, nd);
    return MOED_OK;
}

int moed_basis(moed_handle* h, int32_t nu, int32_t nd, int32_t* up_states, int32_t* dn_states) {
    if (!h) return MOED_EARG;
    (void)hipSetDevice(h->device);
    Sector s;
    int rc = get_sector(h, nu, nd, &s);
    if (rc) return rc;
    if (up_states)
        for (int i = 0; i < s.up->count; ++i) up_states[i] = (int32_t)s.up->str[i];
    if (dn_states)
        for (int i = 0; i < s.dn->count; ++i) dn_states[i] = (int32_t)s.dn->str[i];
    return MOED_OK;
}

int moed_hv_device(moed_handle* h, int32_t nu, int32_t nd, const double* x, double* y, void* stream) {
    if (!h || !x || !y) return MOED_EARG;
    (void)hipSetDevice(h->device);
    Sector s;
    int rc = get_sector(h, nu, nd, &s);
    if (rc) return rc;
    return hv_dev(h, s, x, y, nullptr, nullptr, nullptr, stream ? (hipStream_t)stream : h->stream);
}

int moed_hv(moed_handle* h, int32_t nu, int32_t nd, const double* x, double* y) {
    if (!h || !x || !y) return MOED_EARG;
    (void)hipSetDevice(h->device);
    Sector s;
    int rc = get_sector(h, nu, nd, &s);
    if (rc) return rc;
    hipStream_t st = h->stream;
    DevBuf& X = h->ws[WS_X];
    DevBuf& Y = h->ws[WS_Y];
    HIPCHK(X.ensure(s.D * sizeof(double)));
    HIPCHK(Y.ensure(s.D * sizeof(double)));
    HIPCHK(hipMemcpyAsync(X.p, x, s.D * sizeof(double), hipMemcpyHostToDevice, st));
    rc = hv_dev(h, s, X.as<double>(), Y.as<double>(), nullptr, nullptr, nullptr, st);
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(y, Y.p, s.D * sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    return MOED_OK;
}

int moed_apply_c(moed_handle* h, int32_t pm, int32_t iorb, int32_t ispin, int32_t nu, int32_t nd, const double* x,
                 double* y) {
    if (!h || !x || !y || (pm != 1 && pm != 2) || iorb < 1 || iorb > h->norb || (ispin != 1 && ispin != 2))
        return h ? fail(h, MOED_EARG, "apply_c: bad arguments") : MOED_EARG;
    (void)hipSetDevice(h->device);
    Sector si, so;
    int rc = get_sector(h, nu, nd, &si);
    if (rc) return rc;
    int nuo, ndo;
    target_sector(pm, ispin, nu, nd, &nuo, &ndo);
    rc = get_sector(h, nuo, ndo, &so);
    if (rc) return rc;
    hipStream_t st = h->stream;
    DevBuf& X = h->ws[WS_X];
    DevBuf& Y = h->ws[WS_Y];
    HIPCHK(X.ensure(si.D * sizeof(double)));
    HIPCHK(Y.ensure(so.D * sizeof(double)));
    HIPCHK(hipMemcpyAsync(X.p, x, si.D * sizeof(double), hipMemcpyHostToDevice, st));
    rc = apply_c_dev(h, pm, iorb, ispin, si, so, X.as<double>(), Y.as<double>(), st);
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(y, Y.p, so.D * sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    return MOED_OK;
}

int moed_lanczos(moed_handle* h, int32_t nu, int32_t nd, const double* vec, int32_t nstep, int32_t* nsc, double* a,
                 double* b) {
    if (!h || !vec || !nsc || !a || !b || nstep < 1) return h ? fail(h, MOED_EARG, "lanczos: bad arguments") : MOED_EARG;
    (void)hipSetDevice(h->device);
    Sector s;
    int rc = get_sector(h, nu, nd, &s);
    if (rc) return rc;
    DevBuf& X = h->ws[WS_X];
    HIPCHK(X.ensure(s.D * sizeof(double)));
    HIPCHK(hipMemcpyAsync(X.p, vec, s.D * sizeof(double), hipMemcpyHostToDevice, h->stream));
    double n2;
    rc = dot_dev(h, X.as<double>(), X.as<double>(), s.D, &n2, h->stream);
    if (rc) return rc;
    if (!(n2 > 0.0)) return fail(h, MOED_EZERO, "lanczos: zero start vector");
    int c = 0;
    rc = lanczos_ref_dev(h, s, X.as<double>(), nstep, &c, a, b);
    *nsc = c;
    return rc;
}

int moed_solve(moed_handle* h, int32_t* nev_calc, double* val, double* prob, int32_t* sector, int32_t* level) {
    if (!h || !nev_calc) return MOED_EARG;
    (void)hipSetDevice(h->device);
    const int nsec = (int)h->nelec.size(), nev = h->nev;
    if (nsec < 1 || nev < 1) return fail(h, MOED_EARG, "solve: need nsector >= 1 and nev >= 1");
    h->solved = false;
    std::vector<double> all((size_t)nsec * nev);
    // the sectors' results of the last solve are overwritten in place: their vector buffers are
    // kept (a device free synchronises the device, and most solves of a run have the same shapes)
    if ((int)h->eig.size() != nsec) {
        h->eig.clear();
        for (int is = 0; is < nsec; ++is) h->eig.push_back(std::make_unique<SectorEig>());
    }
    for (auto& E : h->eig) {
        E->nfound = 0;
        E->ev.assign(nev, DBL_MAX);
    }
    std::vector<int> dense;
    std::vector<std::pair<int, int>> big;
    std::vector<int> big_is;
    for (int is = 0; is < nsec; ++is) {
        const int nu = h->nup[is], nd = h->nelec[is] - h->nup[is];
        Sector s;
        if (int rc = get_sector(h, nu, nd, &s)) return rc;
        if (s.D <= 64) {
            dense.push_back(is);
        } else {
            big.push_back({nu, nd});
            big_is.push_back(is);
        }
    }
    if (int rc = eig_dense_batch(h, dense, nev)) return rc;
    if (h->flags & MOED_F_BATCHED) {
        // every sector above the dense size in one multi-problem Lanczos run
        std::vector<SectorEig*> bout;
        for (int is : big_is) bout.push_back(h->eig[is].get());
        if (int rc = eig_batch(h, big, nev, bout, true)) return rc;
    } else {
        for (size_t q = 0; q < big.size(); ++q)
            if (int rc = eig_sector(h, big[q].first, big[q].second, nev, *h->eig[big_is[q]])) return rc;
    }
    if (nev >= 2 && !(h->flags & MOED_F_NO_DEGENERACY_CHECK)) {
        // the sectors solved by Lanczos; every check in one run with MOED_F_BATCHED, else sector by sector
        std::vector<int> cand;
        for (int is = 0; is < nsec; ++is)
            if (h->eig[is]->D > 64 && h->eig[is]->nfound >= 1) cand.push_back(is);
        if (h->flags & MOED_F_BATCHED) {
            if (int rc = degeneracy_check(h, nev, cand)) return rc;
        } else {
            for (int is : cand)
                if (int rc = degeneracy_check(h, nev, {is})) return rc;
        }
    }
    for (int is = 0; is < nsec; ++is)
        for (int i = 0; i < nev; ++i) all[(size_t)is * nev + i] = h->eig[is]->ev[i];
    // ed_solver.F90:65-90
    std::vector<int> ind(all.size());
    heapsort_index((int)all.size(), all.data(), ind.data());
    std::vector<double> pev;
    boltzmann(all, h->beta, pev);
    int nc = 0;
    for (int i = 0; i < nev && i < (int)all.size(); ++i)
        if (pev[ind[i] - 1] > 0.001) ++nc;
    h->kept_val.assign(nc, 0.0);
    h->kept_prob.assign(nc, 0.0);
    h->kept_sector.assign(nc, 0);
    h->kept_level.assign(nc, 0);
    double Z = 0.0;
    for (int i = 0; i < nc; ++i) {
        const int k = ind[i] - 1;
        h->kept_val[i] = all[k];
        h->kept_prob[i] = pev[k];
        h->kept_sector[i] = k / nev + 1;
        h->kept_level[i] = k % nev + 1;
        Z = Z + pev[k];
    }
    for (int i = 0; i < nc; ++i) h->kept_prob[i] = h->kept_prob[i] / Z;
    *nev_calc = nc;
    for (int i = 0; i < nc; ++i) {
        if (val) val[i] = h->kept_val[i];
        if (prob) prob[i] = h->kept_prob[i];
        if (sector) sector[i] = h->kept_sector[i];
        if (level) level[i] = h->kept_level[i];
    }
    h->solved = true;
    return MOED_OK;
}

int moed_sector_eigenvalues(moed_handle* h, int32_t isector, double* ev) {
    if (!h || !ev) return MOED_EARG;
    if (!h->solved) return fail(h, MOED_ESTATE, "call moed_solve first");
    if (isector < 1 || isector > (int)h->eig.size()) return fail(h, MOED_EARG, "sector index out of range");
    const auto& E = *h->eig[isector - 1];
    for (int i = 0; i < h->nev; ++i) ev[i] = E.ev[i];
    return MOED_OK;
}

int moed_eigvec(moed_handle* h, int32_t isector, int32_t ilevel, double* vec) {
    if (!h || !vec) return MOED_EARG;
    if (!h->solved) return fail(h, MOED_ESTATE, "call moed_solve first");
    if (isector < 1 || isector > (int)h->eig.size()) return fail(h, MOED_EARG, "sector index out of range");
    const auto& E = *h->eig[isector - 1];
    if (ilevel < 1 || ilevel > E.nfound) return fail(h, MOED_EARG, "level not available");
    (void)hipSetDevice(h->device);
    HIPCHK(hipMemcpy(vec, E.vec.as<double>() + (size_t)(ilevel - 1) * E.D, E.D * sizeof(double),
                     hipMemcpyDeviceToHost));
    return MOED_OK;
}

int moed_green(moed_handle* h, int32_t nw, const double* omega, double* Gr) {
    if (!h || !omega || !Gr || nw < 1) return h ? fail(h, MOED_EARG, "green: bad arguments") : MOED_EARG;
    if (!h->solved) return fail(h, MOED_ESTATE, "call moed_solve first");
    (void)hipSetDevice(h->device);
    hipStream_t st = h->stream;
    const int norb = h->norb;
    DevBuf om, G;
    HIPCHK(om.ensure(nw * sizeof(double)));
    HIPCHK(G.ensure((size_t)2 * norb * nw * sizeof(double)));
    HIPCHK(hipMemcpyAsync(om.p, omega, nw * sizeof(double), hipMemcpyHostToDevice, st));
    HIPCHK(launch_fill(0.0, (int64_t)2 * norb * nw, G.as<double>(), st));
    h->records.clear();
    // every chain in one multi-problem run where all target sectors take the one-launch H v (up to
    // 2^20 elements); larger chains are bandwidth-bound on their own and run one at a time, on
    // rows padded to an even stride where that lets the H v read 16-byte pairs
    bool small = true;
    for (size_t iev = 0; iev < h->kept_val.size() && small; ++iev) {
        const int is = h->kept_sector[iev] - 1;
        const int nu = h->nup[is], nd = h->nelec[is] - h->nup[is];
        for (int d = -1; d <= 1; d += 2) {
            int64_t D1 = 0, D2 = 0;
            (void)moed_sector_dim(h, std::clamp(nu + d, 0, h->nsite), nd, &D1);
            (void)moed_sector_dim(h, nu, std::clamp(nd + d, 0, h->nsite), &D2);
            if (std::max(D1, D2) > ((int64_t)1 << 20)) small = false;
        }
    }
    if ((h->flags & MOED_F_BATCHED) && small) {
        if (int rc = green_batch(h, nw, om.as<double>(), G.as<double>())) return rc;
        HIPCHK(hipMemcpyAsync(Gr, G.p, (size_t)2 * norb * nw * sizeof(double), hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        return MOED_OK;
    }
    for (size_t iev = 0; iev < h->kept_val.size(); ++iev) {
        const int is = h->kept_sector[iev] - 1, il = h->kept_level[iev] - 1;
        const int nu = h->nup[is], nd = h->nelec[is] - h->nup[is];
        const bool even = nd == nu;  // ed_green.F90:87-94
        const double factor = even ? 1.0 : 0.5;
        const double weight = h->kept_prob[iev] * factor;
        Sector s;
        int rc = get_sector(h, nu, nd, &s);
        if (rc) return rc;
        const SectorEig& E = *h->eig[is];
        if (il >= E.nfound) return fail(h, MOED_ESTATE, "kept state has no eigenvector");
        const double* psi = E.vec.as<double>() + (size_t)il * E.D;
        for (int ispin = 1; ispin <= (even ? 1 : 2); ++ispin)
            for (int iorb = 1; iorb <= norb; ++iorb) {
                GreenRecord rec{};
                rec.iev = (int)iev + 1;
                rec.iorb = iorb;
                rec.ispin = ispin;
                rec.E = h->kept_val[iev];
                rec.p = h->kept_prob[iev];
                rec.even = even;
                rc = green_diag(h, s, psi, h->kept_val[iev], weight, iorb, ispin, om.as<double>(), nw, G.as<double>(),
                                rec);
                if (rc) return rc;
                h->records.push_back(std::move(rec));
            }
    }
    HIPCHK(hipMemcpyAsync(Gr, G.p, (size_t)2 * norb * nw * sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    return MOED_OK;
}

int moed_green_nrecords(moed_handle* h, int32_t* nrec) {
    if (!h || !nrec) return MOED_EARG;
    *nrec = (int32_t)h->records.size();
    return MOED_OK;
}

int moed_green_record(moed_handle* h, int32_t irec, int32_t* iev, int32_t* iorb, int32_t* ispin, double* E, double* p,
                      int32_t* even, int32_t* nsp, int32_t* nsn, double* ap, double* bp, double* an, double* bn) {
    if (!h) return MOED_EARG;
    if (irec < 0 || irec >= (int)h->records.size()) return fail(h, MOED_EARG, "record index out of range");
    const GreenRecord& r = h->records[irec];
    if (iev) *iev = r.iev;
    if (iorb) *iorb = r.iorb;
    if (ispin) *ispin = r.ispin;
    if (E) *E = r.E;
    if (p) *p = r.p;
    if (even) *even = r.even;
    if (nsp) *nsp = r.nsp;
    if (nsn) *nsn = r.nsn;
    const size_t n = h->nstep + 1;
    if (ap) std::copy(r.ap.begin(), r.ap.begin() + n, ap);
    if (bp) std::copy(r.bp.begin(), r.bp.begin() + n, bp);
    if (an) std::copy(r.an.begin(), r.an.begin() + n, an);
    if (bn) std::copy(r.bn.begin(), r.bn.begin() + n, bn);
    return MOED_OK;
}

int moed_green_cf(moed_handle* h, int32_t sign, double E, int32_t nsc, const double* a, const double* b, int32_t nw,
                  const double* omega, double* g) {
    if (!h || !a || !b || !omega || !g || nsc < 0 || nw < 1) return MOED_EARG;
    (void)hipSetDevice(h->device);
    hipStream_t st = h->stream;
    DevBuf ab, om, G;
    HIPCHK(ab.ensure((size_t)2 * (nsc + 1) * sizeof(double)));
    HIPCHK(om.ensure(nw * sizeof(double)));
    HIPCHK(G.ensure((size_t)2 * nw * sizeof(double)));
    HIPCHK(hipMemcpyAsync(ab.p, a, (nsc + 1) * sizeof(double), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(ab.as<double>() + nsc + 1, b, (nsc + 1) * sizeof(double), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(om.p, omega, nw * sizeof(double), hipMemcpyHostToDevice, st));
    HIPCHK(launch_fill(0.0, 2 * nw, G.as<double>(), st));
    const int compat = (h->flags & MOED_F_EXACT_CMPLX) ? 0 : 1;
    CfArgs cf{ab.as<double>(), ab.as<double>() + nsc + 1, nsc, sign >= 0 ? 1 : -1, E, 1.0, om.as<double>(), nw, compat,
              G.as<double>(), 1};
    HIPCHK(launch_cf(cf, st));
    HIPCHK(hipMemcpyAsync(g, G.p, (size_t)2 * nw * sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    return MOED_OK;
}

static int dos_record(moed_handle* h, int nw, const double* d_om, double small, double ev, int nsp, int nsn,
                      const double* ap, const double* bp, const double* an, const double* bn, double weight,
                      int accumulate, double* d_out, int stride) {
    hipStream_t st = h->stream;
    DevBuf ab;
    const size_t np = nsp + 1, nn = nsn + 1;
    HIPCHK(ab.ensure(2 * (np + nn) * sizeof(double)));
    double* d = ab.as<double>();
    HIPCHK(hipMemcpyAsync(d, ap, np * sizeof(double), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(d + np, bp, np * sizeof(double), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(d + 2 * np, an, nn * sizeof(double), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(d + 2 * np + nn, bn, nn * sizeof(double), hipMemcpyHostToDevice, st));
    const int compat = (h->flags & MOED_F_EXACT_CMPLX) ? 0 : 1;
    DosArgs a{d, d + np, d + 2 * np, d + 2 * np + nn, nsp, nsn, ev, small, weight, d_om, nw, compat, accumulate,
              d_out, stride};
    HIPCHK(launch_dos(a, st));
    HIPCHK(hipStreamSynchronize(st));  // ab is released on return
    return MOED_OK;
}

int moed_dos_diag(moed_handle* h, int32_t nw, const double* omega_real, double small, double ev, int32_t nsp,
                  int32_t nsn, const double* ap, const double* bp, const double* an, const double* bn, double* aff) {
    if (!h) return MOED_EARG;
    if (nw < 1 || nsp < 0 || nsn < 0 || !omega_real || !ap || !bp || !an || !bn || !aff)
        return fail(h, MOED_EARG, "dos_diag: bad arguments");
    (void)hipSetDevice(h->device);
    hipStream_t st = h->stream;
    DevBuf om, out;
    HIPCHK(om.ensure(nw * sizeof(double)));
    HIPCHK(out.ensure(nw * sizeof(double)));
    HIPCHK(hipMemcpyAsync(om.p, omega_real, nw * sizeof(double), hipMemcpyHostToDevice, st));
    int rc = dos_record(h, nw, om.as<double>(), small, ev, nsp, nsn, ap, bp, an, bn, 1.0, 0, out.as<double>(), 1);
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(aff, out.p, nw * sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    return MOED_OK;
}

int moed_dos(moed_handle* h, int32_t nw, const double* omega_real, double small, int32_t weighted, double* Aff) {
    if (!h) return MOED_EARG;
    if (nw < 1 || !omega_real || !Aff) return fail(h, MOED_EARG, "dos: bad arguments");
    if (h->records.empty()) return fail(h, MOED_ESTATE, "dos: no Green's-function records (call moed_green first)");
    (void)hipSetDevice(h->device);
    hipStream_t st = h->stream;
    const int norb = h->norb;
    DevBuf om, out;
    HIPCHK(om.ensure(nw * sizeof(double)));
    HIPCHK(out.ensure((size_t)norb * nw * sizeof(double)));
    HIPCHK(hipMemcpyAsync(om.p, omega_real, nw * sizeof(double), hipMemcpyHostToDevice, st));
    HIPCHK(launch_fill(0.0, (size_t)norb * nw, out.as<double>(), st));
    // Aff(iorb, :) column-major (norb, nw): element (o, w) at o + norb w
    for (size_t i = 0; i < h->records.size(); ++i) {
        const GreenRecord& r = h->records[i];
        if (!weighted) {  // every record overwrites its orbital: only the last one per orbital remains
            bool later = false;
            for (size_t j = i + 1; j < h->records.size() && !later; ++j) later = h->records[j].iorb == r.iorb;
            if (later) continue;
        }
        const double weight = r.p * (r.even ? 1.0 : 0.5);
        int rc = dos_record(h, nw, om.as<double>(), small, r.E, r.nsp, r.nsn, r.ap.data(), r.bp.data(), r.an.data(),
                            r.bn.data(), weight, weighted ? 1 : 0, out.as<double>() + (r.iorb - 1), norb);
        if (rc) return rc;
    }
    HIPCHK(hipMemcpyAsync(Aff, out.p, (size_t)norb * nw * sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    return MOED_OK;
}

// ---------------------------------------------------------------- on-disk records

// Fortran unformatted sequential files: every record is a 4-byte little-endian length, the
// payload, and the length again (records up to 2^31 - 1 bytes)
extern "C++" {
namespace {
struct RecordWriter {
    std::string buf;
    template <class T>
    void put(const T& v) {
        buf.append(reinterpret_cast<const char*>(&v), sizeof(T));
    }
    void put(const double* v, size_t n) { buf.append(reinterpret_cast<const char*>(v), n * sizeof(double)); }
    bool flush(FILE* f) {
        if (buf.size() > 0x7fffffffull) return false;
        const int32_t n = (int32_t)buf.size();
        const bool ok = fwrite(&n, 4, 1, f) == 1 && (buf.empty() || fwrite(buf.data(), buf.size(), 1, f) == 1) &&
                        fwrite(&n, 4, 1, f) == 1;
        buf.clear();
        return ok;
    }
};

// the next record's payload, or false at the end of the file or on a malformed record
bool read_record(FILE* f, std::string* out) {
    int32_t n = 0, m = 0;
    if (fread(&n, 4, 1, f) != 1 || n < 0) return false;
    out->resize(n);
    if (n > 0 && fread(&(*out)[0], n, 1, f) != 1) return false;
    return fread(&m, 4, 1, f) == 1 && m == n;
}

// Fortran edit descriptor Iw.m for a non-negative field: at least m digits in w columns, w
// asterisks when the value does not fit (as the reference's file names do for sector >= 100)
std::string fortran_i(int64_t v, int w, int m) {
    char b[32];
    std::snprintf(b, sizeof b, "%0*lld", m, (long long)v);
    std::string s(b);
    if (v < 0 || (int)s.size() > w) return std::string(w, '*');
    return std::string(w - s.size(), ' ') + s;
}

// eigvec_filename (ed_io.F90:141-147): "eigvec-" I2.2 "-" I3.3 "-" I4.4
std::string eigvec_path(const char* dir, int isector, int iev, int node) {
    std::string name = "eigvec-" + fortran_i(isector, 2, 2) + "-" + fortran_i(iev, 3, 3) + "-" + fortran_i(node, 4, 4);
    if (dir && *dir) return std::string(dir) + "/" + name;
    return name;
}
}  // namespace
}  // extern "C++"

int moed_eigvec_filename(int32_t isector, int32_t iev, int32_t node, char* buf, int32_t len) {
    if (!buf || len <= 0) return MOED_EARG;
    std::snprintf(buf, (size_t)len, "%s", eigvec_path(nullptr, isector, iev, node).c_str());
    return MOED_OK;
}

int moed_export_eigvec(const char* dir, int32_t isector, int32_t iev, int32_t node, int64_t nloc, const double* vec) {
    if (nloc < 0 || (nloc > 0 && !vec) || nloc * 8 > 0x7fffffffll) return MOED_EARG;
    FILE* f = std::fopen(eigvec_path(dir, isector, iev, node).c_str(), "wb");
    if (!f) return MOED_EARG;
    RecordWriter w;
    w.put((int32_t)nloc);
    bool ok = w.flush(f);
    w.put(vec, (size_t)nloc);
    ok = ok && w.flush(f);
    ok = (std::fclose(f) == 0) && ok;
    return ok ? MOED_OK : MOED_EARG;
}

int moed_import_eigvec(const char* dir, int32_t isector, int32_t iev, int32_t node, int64_t* nloc, double* vec) {
    if (!nloc) return MOED_EARG;
    FILE* f = std::fopen(eigvec_path(dir, isector, iev, node).c_str(), "rb");
    if (!f) return MOED_EARG;  // 'import_eigvec: File not found' (ed_io.F90:120-124)
    std::string r1, r2;
    const bool ok = read_record(f, &r1) && r1.size() == 4 && read_record(f, &r2);
    std::fclose(f);
    if (!ok) return MOED_EARG;
    int32_t n;
    std::memcpy(&n, r1.data(), 4);
    if (n < 0 || r2.size() != (size_t)n * 8) return MOED_EARG;
    *nloc = n;
    if (vec) std::memcpy(vec, r2.data(), r2.size());
    return MOED_OK;
}

int moed_export_green(moed_handle* h, const char* path) {
    if (!h || !path) return MOED_EARG;
    if (h->records.empty()) return fail(h, MOED_ESTATE, "export_green: no records (call moed_green first)");
    FILE* f = std::fopen(path, "wb");
    if (!f) return fail(h, MOED_EARG, std::string("export_green: cannot open ") + path);
    bool ok = true;
    RecordWriter w;
    for (const GreenRecord& r : h->records) {  // ed_green.F90:112-114
        w.put((int32_t)r.nsp);
        w.put((int32_t)r.nsn);
        ok = ok && w.flush(f);
        w.put((int32_t)r.iev);
        w.put((int32_t)r.iorb);
        w.put(r.E);
        w.put(r.p);
        w.put((int32_t)(r.even ? 1 : 0));  // default LOGICAL, 4 bytes
        w.put(r.ap.data(), r.nsp + 1);
        w.put(r.bp.data(), r.nsp + 1);
        w.put(r.an.data(), r.nsn + 1);
        w.put(r.bn.data(), r.nsn + 1);
        ok = ok && w.flush(f);
    }
    ok = (std::fclose(f) == 0) && ok;
    return ok ? MOED_OK : fail(h, MOED_EARG, "export_green: write failed");
}

int moed_import_green(moed_handle* h, const char* path) {
    if (!h || !path) return MOED_EARG;
    FILE* f = std::fopen(path, "rb");
    if (!f) return fail(h, MOED_EARG, std::string("import_green: cannot open ") + path);
    std::vector<GreenRecord> recs;
    std::string a, b;
    bool ok = true;
    int last_iev = 0, count_in_iev = 0;
    while (read_record(f, &a)) {
        int32_t hd[2];
        if (a.size() != 8 || !read_record(f, &b)) {
            ok = false;
            break;
        }
        std::memcpy(hd, a.data(), 8);
        GreenRecord r{};
        r.nsp = hd[0];
        r.nsn = hd[1];
        const size_t need = 4 + 4 + 8 + 8 + 4 + 8 * (size_t)(2 * (r.nsp + 1) + 2 * (r.nsn + 1));
        if (r.nsp < 0 || r.nsn < 0 || b.size() != need) {
            ok = false;
            break;
        }
        const char* q = b.data();
        int32_t i4;
        std::memcpy(&i4, q, 4), q += 4, r.iev = i4;
        std::memcpy(&i4, q, 4), q += 4, r.iorb = i4;
        std::memcpy(&r.E, q, 8), q += 8;
        std::memcpy(&r.p, q, 8), q += 8;
        std::memcpy(&i4, q, 4), q += 4, r.even = i4 != 0;
        // moed_green_record hands out nstep + 1 coefficients per array: longer records do not fit
        if (r.iorb < 1 || r.iorb > h->norb || r.nsp > h->nstep || r.nsn > h->nstep) {
            ok = false;
            break;
        }
        // every array holds nstep + 1 entries, zero past the record's own steps (as green_diag
        // leaves the entries past a breakdown, ed_lanczos.F90:27-28)
        auto take = [&](std::vector<double>& v, int n) {
            v.assign((size_t)h->nstep + 1, 0.0);
            std::memcpy(v.data(), q, 8 * (size_t)n);
            q += 8 * (size_t)n;
        };
        take(r.ap, r.nsp + 1);
        take(r.bp, r.nsp + 1);
        take(r.an, r.nsn + 1);
        take(r.bn, r.nsn + 1);
        // the file does not store the spin: per state, Norb spin-up records, then (N_up != N_dn)
        // Norb spin-down records (ed_green.F90:107-131)
        count_in_iev = r.iev == last_iev ? count_in_iev + 1 : 0;
        last_iev = r.iev;
        r.ispin = count_in_iev < h->norb ? 1 : 2;
        recs.push_back(std::move(r));
    }
    std::fclose(f);
    if (!ok) return fail(h, MOED_EARG, "import_green: malformed apbpanbn record, or orbital / steps beyond the handle's");
    h->records = std::move(recs);
    return MOED_OK;
}

int moed_green_cf_adjoint(moed_handle* h, int32_t sign, double E, int32_t nsc, const double* a, const double* b,
                          int32_t nw, const double* omega, const double* gbar, double* ga, double* gb) {
    if (!h || !a || !b || !omega || !gbar || !ga || !gb || nsc < 0 || nw < 1) return MOED_EARG;
    (void)hipSetDevice(h->device);
    hipStream_t st = h->stream;
    const size_t n1 = nsc + 1;
    DevBuf ab, om, gbr, scr, out;
    HIPCHK(ab.ensure(2 * n1 * sizeof(double)));
    HIPCHK(om.ensure(nw * sizeof(double)));
    HIPCHK(gbr.ensure((size_t)2 * nw * sizeof(double)));
    HIPCHK(scr.ensure(n1 * nw * 6 * sizeof(double)));
    HIPCHK(out.ensure(2 * n1 * sizeof(double)));
    HIPCHK(hipMemcpyAsync(ab.p, a, n1 * sizeof(double), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(ab.as<double>() + n1, b, n1 * sizeof(double), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(om.p, omega, nw * sizeof(double), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(gbr.p, gbar, (size_t)2 * nw * sizeof(double), hipMemcpyHostToDevice, st));
    const int compat = (h->flags & MOED_F_EXACT_CMPLX) ? 0 : 1;
    CfArgs cf{ab.as<double>(), ab.as<double>() + n1, nsc, sign >= 0 ? 1 : -1, E, 1.0, om.as<double>(), nw, compat,
              nullptr, 1};
    HIPCHK(launch_cf_adjoint(cf, gbr.as<double>(), scr.as<double>(), out.as<double>(), out.as<double>() + n1, st));
    HIPCHK(hipMemcpyAsync(ga, out.p, n1 * sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(gb, out.as<double>() + n1, n1 * sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    return MOED_OK;
}

int moed_calc_dev(moed_handle* h, const double* ek, const double* vk, const double* ef, int32_t nw, const double* omega,
                  double* Dev) {
    if (!h || !ek || !vk || !ef || !omega || !Dev || nw < 1) return MOED_EARG;
    // Delta(iw) of calc_dev is the fit's Delta at x = ev_to_x(ek, vk, ef) with a zero target
    const int norb = h->norb, nbath = h->nbath, nx = nbath + nbath * norb + norb;
    std::vector<double> x(nx);
    for (int i = 0; i < nbath; ++i) x[i] = ek[norb + i];
    for (int o = 0; o < norb; ++o)
        for (int k = 0; k < nbath; ++k) x[nbath + o * nbath + k] = vk[o + (size_t)norb * k];
    for (int o = 0; o < norb; ++o) x[nx - norb + o] = ef[o];
    return fit_delta_host(h, nx, x.data(), nw, omega, Dev);
}

int moed_chi2_grad(moed_handle* h, int32_t nfit, int32_t nx, const double* x, int32_t nw, const double* omega,
                   const double* Dev, double* f, double* g) {
    if (!h || !x || !omega || !Dev || nfit < 1 || nw < 1) return MOED_EARG;
    if (nx != h->nbath * (1 + h->norb) + h->norb) return fail(h, MOED_EARG, "nx != Nbath*(1+Norb)+Norb");
    return fit_eval_host(h, nfit, nx, x, nw, omega, Dev, f, g);
}

int moed_minimize(moed_handle* h, int32_t nfit, int32_t nx, double* x, int32_t nw, const double* omega,
                  const double* Dev, double ftol, double* fret, int32_t* iter, int32_t* status) {
    if (!h || !x || !omega || !Dev || nfit < 1 || nw < 1) return MOED_EARG;
    if (nx != h->nbath * (1 + h->norb) + h->norb) return fail(h, MOED_EARG, "nx != Nbath*(1+Norb)+Norb");
    return fit_minimize_host(h, nfit, nx, x, nw, omega, Dev, ftol, fret, iter, status);
}

int moed_minimize_device_noalloc(moed_handle* h, int32_t nfit, int32_t nx, double* x_device, int32_t nw,
                                 const double* omega_device, const double* Dev_device, double ftol,
                                 double* fret_device, int32_t* iter_device, int32_t* status_device, void* stream) {
    if (!h) return MOED_EARG;
    if (!x_device || !omega_device || !Dev_device || nfit < 1 || nw < 1)
        return fail(h, MOED_EARG, "minimize_device_noalloc: bad arguments");
    if (nx != h->nbath * (1 + h->norb) + h->norb) return fail(h, MOED_EARG, "nx != Nbath*(1+Norb)+Norb");
    (void)hipSetDevice(h->device);
    FitCtx c = fit_ctx(h);
    if (stream) c.stream = (hipStream_t)stream;
    return fit_minimize_noalloc(c, h->fit_scratch, nfit, nx, x_device, nw, omega_device, Dev_device, ftol,
                                fret_device, iter_device, status_device);
}

int moed_band_tb(int32_t norb, int32_t nq, double* hk) {
    if (norb < 1 || nq < 1 || !hk) return MOED_EARG;
    const double pi = 4.0 * std::atan(1.0);
    const int nk = (int)std::sqrt((double)nq);
    if (nk * nk != nq) return MOED_EARG;  // the reference stops unless nk*nk = Nq (ed_hamiltonian.F90:51)
    const double dq = 2.0 * pi / (double)(float)nk;
    int ik = 0;
    for (int i = 1; i <= nk; ++i)
        for (int j = 1; j <= nk; ++j, ++ik) {
            const double kx = dq * (double)(float)(i - 1), ky = dq * (double)(float)(j - 1);
            const double v = -0.5 * (std::cos(kx) + std::cos(ky));
            for (int o = 0; o < norb; ++o) hk[(size_t)ik * norb + o] = v;
        }
    return MOED_OK;
}

int moed_delta_new(moed_handle* h, int32_t nw, int32_t nq, double rmu, const double* hk, const double* Gr,
                   double* Dev, double* Gksum, double* nocc) {
    if (!h) return MOED_EARG;
    if (nw < 1 || nq < 1 || !hk || !Gr || !Dev) return fail(h, MOED_EARG, "delta_new: bad arguments");
    (void)hipSetDevice(h->device);
    hipStream_t st = h->stream;
    const int norb = h->norb;
    const size_t ncx = (size_t)2 * norb * nw;
    DevBuf b;
    HIPCHK(b.ensure(((size_t)nq * norb + 3 * ncx) * sizeof(double)));
    double* d_hk = b.as<double>();
    double* d_gr = d_hk + (size_t)nq * norb;
    double* d_dev = d_gr + ncx;
    double* d_gks = d_dev + ncx;
    HIPCHK(hipMemcpyAsync(d_hk, hk, (size_t)nq * norb * sizeof(double), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(d_gr, Gr, ncx * sizeof(double), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(d_dev, Dev, ncx * sizeof(double), hipMemcpyHostToDevice, st));
    DeltaArgs a{norb, nw, nq, rmu, d_hk, d_gr, d_dev, d_gks};
    HIPCHK(launch_delta_new(a, st));
    std::vector<double> gks(ncx);
    HIPCHK(hipMemcpyAsync(Dev, d_dev, ncx * sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(gks.data(), d_gks, ncx * sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    if (Gksum) std::copy(gks.begin(), gks.end(), Gksum);
    if (nocc)
        for (int o = 0; o < norb; ++o) {
            double sum = 0.0;
            for (int w = 0; w < nw; ++w) sum = sum + gks[2 * ((size_t)w * norb + o)];
            nocc[o] = 2.0 * sum / h->beta + 0.50;
        }
    return MOED_OK;
}

int moed_omega_grid(int32_t nw, double beta, double* omega) {
    if (nw < 0 || !omega || !(beta > 0)) return MOED_EARG;
    const double pi = 4.0 * std::atan(1.0);
    for (int i = 0; i < nw; ++i) omega[i] = (2.0 * (double)(float)i + 1) * pi / beta;
    return MOED_OK;
}

int moed_fill_start_device(moed_handle* h, uint64_t seed, int64_t n, double* x, void* stream) {
    if (!h || !x || n < 0) return MOED_EARG;
    (void)hipSetDevice(h->device);
    HIPCHK(launch_fill_start(seed, n, x, stream ? (hipStream_t)stream : h->stream));
    return MOED_OK;
}

int moed_set_fit_launch(moed_handle* h, int32_t iterations) {
    if (!h) return MOED_EARG;
    if (iterations < 0) return fail(h, MOED_EARG, "iterations must be >= 0");
    h->fit_budget = iterations;
    return MOED_OK;
}

int moed_fit_nfev(moed_handle* h, int32_t nfit, int32_t* nfev) {
    if (!h || !nfev || nfit < 0) return MOED_EARG;
    if ((size_t)nfit > h->fit_nfev.size()) return fail(h, MOED_ESTATE, "fit_nfev: fewer fits in the last moed_minimize");
    std::copy(h->fit_nfev.begin(), h->fit_nfev.begin() + nfit, nfev);
    return MOED_OK;
}

int moed_set_persistent(moed_handle* h, int64_t max_dim) {
    if (!h) return MOED_EARG;
    if (max_dim < -1 || max_dim > kPersistMaxDim)
        return fail(h, MOED_EARG, "max_dim must be -1 (default), 0 (off) or at most 2^20");
    h->persist_max = max_dim;
    return MOED_OK;
}

int moed_set_hv_shape(moed_handle* h, int32_t threads, int32_t strip) {
    if (!h) return MOED_EARG;
    if (threads != 0 && threads != 256 && threads != 512 && threads != 1024)
        return fail(h, MOED_EARG, "threads must be 0, 256, 512 or 1024");
    if (strip != 0 && strip != 64 && strip != 128) return fail(h, MOED_EARG, "strip must be 0, 64 or 128");
    h->hv_threads = threads;
    h->hv_strip = strip;
    h->clear_graphs();
    return MOED_OK;
}

int moed_set_eig_chunk(moed_handle* h, int32_t vectors) {
    if (!h) return MOED_EARG;
    if (vectors < 0) return fail(h, MOED_EARG, "vectors must be 0 (default) or positive");
    (void)hipSetDevice(h->device);
    (void)hipStreamSynchronize(h->stream);
    h->eig_chunk = vectors;
    h->basis.clear();
    return MOED_OK;
}

void* moed_stream(moed_handle* h) { return h ? (void*)h->stream : nullptr; }

}  // extern "C"
