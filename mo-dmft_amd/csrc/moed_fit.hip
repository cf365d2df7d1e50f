// Bath fit driver: batched device evaluation of FUNC/dfunc and the reference's conjugate-gradient
// minimiser restated as one C++20 coroutine per fit.  Every fit suspends at each function
// evaluation; the driver gathers the pending points of all fits into one launch, so one kernel
// pair evaluates up to nfit chi^2 (and gradients) at a time.
#include <hip/hip_runtime.h>

#include <cmath>
#include <coroutine>
#include <exception>
#include <memory>
#include <vector>

#include "moed_fit.h"
#include "moed_internal.h"

namespace moed {
namespace {

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
#define FCHK(expr)                                   \
    do {                                             \
        hipError_t _e = (expr);                      \
        if (_e != hipSuccess) return hfail(c, _e, #expr); \
    } while (0)

// device-resident problem of one batch of evaluations
struct FitDev {
    Buf x, om, dev, f, g, work, act;
    int nfit = 0, nx = 0, nw = 0;
};

int fitdev_init(const FitCtx& c, FitDev& d, int nfit, int nx, int nw, const double* omega, const double* Dev) {
    d.nfit = nfit;
    d.nx = nx;
    d.nw = nw;
    FCHK(d.x.alloc((size_t)nfit * nx * sizeof(double)));
    FCHK(d.om.alloc(nw * sizeof(double)));
    FCHK(d.dev.alloc((size_t)2 * c.norb * nw * sizeof(double)));
    FCHK(d.f.alloc(nfit * sizeof(double)));
    FCHK(d.g.alloc((size_t)nfit * nx * sizeof(double)));
    FCHK(d.work.alloc((size_t)nfit * c.norb * nw * 2 * sizeof(double)));
    FCHK(d.act.alloc(nfit * sizeof(int32_t)));
    FCHK(hipMemcpyAsync(d.om.p, omega, nw * sizeof(double), hipMemcpyHostToDevice, c.stream));
    FCHK(hipMemcpyAsync(d.dev.p, Dev, (size_t)2 * c.norb * nw * sizeof(double), hipMemcpyHostToDevice, c.stream));
    return MOED_OK;
}

// evaluate: x_host[nfit*nx]; act[i] bit0 = FUNC, bit1 = dfunc
int fitdev_eval(const FitCtx& c, FitDev& d, const double* xh, const int32_t* acth, double* fh, double* gh) {
    FCHK(hipMemcpyAsync(d.x.p, xh, (size_t)d.nfit * d.nx * sizeof(double), hipMemcpyHostToDevice, c.stream));
    FCHK(hipMemcpyAsync(d.act.p, acth, d.nfit * sizeof(int32_t), hipMemcpyHostToDevice, c.stream));
    FitArgs a{d.x.as<double>(), d.nfit, d.nx, c.norb, c.nbath, d.om.as<double>(), d.dev.as<double>(), d.nw,
              c.compat, d.f.as<double>(), d.g.as<double>(), d.work.as<double>(), d.act.as<int32_t>()};
    FCHK(launch_fit(a, c.stream));
    if (fh) FCHK(hipMemcpyAsync(fh, d.f.p, d.nfit * sizeof(double), hipMemcpyDeviceToHost, c.stream));
    if (gh) FCHK(hipMemcpyAsync(gh, d.g.p, (size_t)d.nfit * d.nx * sizeof(double), hipMemcpyDeviceToHost, c.stream));
    FCHK(hipStreamSynchronize(c.stream));
    return MOED_OK;
}

// ------------------------------------------------------------------ coroutine plumbing

struct FitTask {
    struct promise_type {
        int rc = 0;
        FitTask get_return_object() { return FitTask{std::coroutine_handle<promise_type>::from_promise(*this)}; }
        std::suspend_always initial_suspend() noexcept { return {}; }
        std::suspend_always final_suspend() noexcept { return {}; }
        void return_value(int r) { rc = r; }
        void unhandled_exception() { std::terminate(); }
    };
    std::coroutine_handle<promise_type> h;
    explicit FitTask(std::coroutine_handle<promise_type> hh) : h(hh) {}
    FitTask(FitTask&& o) noexcept : h(o.h) { o.h = nullptr; }
    FitTask(const FitTask&) = delete;
    ~FitTask() {
        if (h) h.destroy();
    }
};

// one fit's pending evaluation request and its answer
struct FitSlot {
    const double* x = nullptr;  // point to evaluate (lives in the coroutine frame)
    double* gout = nullptr;     // gradient destination, or null for FUNC only
    double f = 0.0;
    bool pending = false;
    int iter = 0;
    double fret = 0.0;
};

struct Eval {
    FitSlot* s;
    const double* x;
    double* gout;
    bool await_ready() const noexcept { return false; }
    void await_suspend(std::coroutine_handle<>) noexcept {
        s->x = x;
        s->gout = gout;
        s->pending = true;
    }
    double await_resume() const noexcept { return s->f; }
};

// frprmn + linmin + mnbrak + brent + f1dim (frprmn.f90:1-310) for one fit.  The constants
// that the reference writes as REAL*4 literals keep their single-precision values.
FitTask frprmn_task(FitSlot* S, double* p, int n, double ftol) {
    const int ITMAX = 20000, BITMAX = 100;
    const double EPS = 1.e-10;
    const double LTOL = (double)1.e-8f, CGOLD = (double).3819660f, ZEPS = (double)1.0e-10f;
    const double GOLD = (double)1.618034f, GLIMIT = 100.0, TINY = (double)1.e-20f;
    std::vector<double> g(n), h(n), xi(n), pcom(n), xicom(n), xt(n);
    auto fsign = [](double a, double b) { return b >= 0.0 ? std::fabs(a) : -std::fabs(a); };
#define F1DIM(res, xv)                                                     \
    do {                                                                   \
        const double _xv = (xv);                                           \
        for (int _j = 0; _j < n; ++_j) xt[_j] = pcom[_j] + _xv * xicom[_j]; \
        res = co_await Eval{S, xt.data(), nullptr};                        \
    } while (0)

    double fp = co_await Eval{S, p, nullptr};
    (void)co_await Eval{S, p, xi.data()};
    for (int j = 0; j < n; ++j) {
        g[j] = -xi[j];
        h[j] = g[j];
        xi[j] = h[j];
    }
    for (int its = 1; its <= ITMAX; ++its) {
        S->iter = its;
        // ---- linmin (frprmn.f90:86-117)
        for (int j = 0; j < n; ++j) {
            pcom[j] = p[j];
            xicom[j] = xi[j];
        }
        double ax = 0., bx = 1., cx, fa, fb, fc, fu, u, ulim, q, r, dum;
        // ---- mnbrak (frprmn.f90:236-310)
        F1DIM(fa, ax);
        F1DIM(fb, bx);
        if (fb > fa) {
            dum = ax;
            ax = bx;
            bx = dum;
            dum = fb;
            fb = fa;
            fa = dum;
        }
        cx = bx + GOLD * (bx - ax);
        F1DIM(fc, cx);
        while (fb >= fc) {
            r = (bx - ax) * (fb - fc);
            q = (bx - cx) * (fb - fa);
            u = bx - ((bx - cx) * q - (bx - ax) * r) / (2. * fsign(std::fmax(std::fabs(q - r), TINY), q - r));
            ulim = bx + GLIMIT * (cx - bx);
            bool ret = false;
            if ((bx - u) * (u - cx) > 0.) {
                F1DIM(fu, u);
                if (fu < fc) {
                    ax = bx;
                    fa = fb;
                    bx = u;
                    fb = fu;
                    ret = true;
                } else if (fu > fb) {
                    cx = u;
                    fc = fu;
                    ret = true;
                } else {
                    u = cx + GOLD * (cx - bx);
                    F1DIM(fu, u);
                }
            } else if ((cx - u) * (u - ulim) > 0.) {
                F1DIM(fu, u);
                if (fu < fc) {
                    bx = cx;
                    cx = u;
                    u = cx + GOLD * (cx - bx);
                    fb = fc;
                    fc = fu;
                    F1DIM(fu, u);
                }
            } else if ((u - ulim) * (ulim - cx) >= 0.) {
                u = ulim;
                F1DIM(fu, u);
            } else {
                u = cx + GOLD * (cx - bx);
                F1DIM(fu, u);
            }
            if (ret) break;
            ax = bx;
            bx = cx;
            cx = u;
            fa = fb;
            fb = fc;
            fc = fu;
        }
        // ---- brent (frprmn.f90:119-210) on (ax, bx, cx)
        double a = std::fmin(ax, cx), b = std::fmax(ax, cx);
        double v = bx, w = v, x = v, e = 0., d = 0.0, etemp, fv, fw, fx, pp, qq, rr, tol1, tol2, xm;
        F1DIM(fx, x);
        fv = fx;
        fw = fx;
        bool found = false;
        double xmin = 0.0, fret = 0.0;
        for (int iter = 1; iter <= BITMAX; ++iter) {
            xm = 0.5 * (a + b);
            tol1 = LTOL * std::fabs(x) + ZEPS;
            tol2 = 2. * tol1;
            if (std::fabs(x - xm) <= (tol2 - .5 * (b - a))) {
                xmin = x;
                fret = fx;
                found = true;
                break;
            }
            bool golden = true;
            if (std::fabs(e) > tol1) {
                rr = (x - w) * (fx - fv);
                qq = (x - v) * (fx - fw);
                pp = (x - v) * qq - (x - w) * rr;
                qq = 2. * (qq - rr);
                if (qq > 0.) pp = -pp;
                qq = std::fabs(qq);
                etemp = e;
                e = d;
                if (!(std::fabs(pp) >= std::fabs(.5 * qq * etemp) || pp <= qq * (a - x) || pp >= qq * (b - x))) {
                    d = pp / qq;
                    u = x + d;
                    if (u - a < tol2 || b - u < tol2) d = fsign(tol1, xm - x);
                    golden = false;
                }
            }
            if (golden) {
                e = (x >= xm) ? a - x : b - x;
                d = CGOLD * e;
            }
            u = (std::fabs(d) >= tol1) ? x + d : x + fsign(tol1, d);
            F1DIM(fu, u);
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
        if (!found) co_return MOED_ECONV;  // 'brent exceed maximum iterations'
        for (int j = 0; j < n; ++j) {
            xi[j] = xmin * xi[j];
            p[j] = p[j] + xi[j];
        }
        S->fret = fret;
        // ---- end linmin
        if (2.0 * std::fabs(fret - fp) <= ftol * (std::fabs(fret) + std::fabs(fp) + EPS)) co_return MOED_OK;
        fp = fret;
        (void)co_await Eval{S, p, xi.data()};
        double gg = 0.0, dgg = 0.0;
        for (int j = 0; j < n; ++j) {
            gg = gg + g[j] * g[j];
            dgg = dgg + (xi[j] + g[j]) * xi[j];
        }
        if (gg == 0.0) co_return MOED_OK;
        const double gam = dgg / gg;
        for (int j = 0; j < n; ++j) {
            g[j] = -xi[j];
            h[j] = g[j] + gam * h[j];
            xi[j] = h[j];
        }
    }
#undef F1DIM
    co_return MOED_ECONV;  // 'frprmn maximum iterations exceeded!'
}

}  // namespace

int fit_eval(const FitCtx& c, int nfit, int nx, const double* x, int nw, const double* omega, const double* Dev,
             double* f, double* g) {
    FitDev d;
    int rc = fitdev_init(c, d, nfit, nx, nw, omega, Dev);
    if (rc) return rc;
    std::vector<int32_t> act(nfit, (f ? 1 : 0) | (g ? 2 : 0));
    return fitdev_eval(c, d, x, act.data(), f, g);
}

int fit_delta(const FitCtx& c, int nx, const double* x, int nw, const double* omega, double* Dev) {
    // the residual pass with a zero target returns Delta itself
    FitDev d;
    std::vector<double> zero((size_t)2 * c.norb * nw, 0.0);
    int rc = fitdev_init(c, d, 1, nx, nw, omega, zero.data());
    if (rc) return rc;
    FCHK(hipMemcpyAsync(d.x.p, x, nx * sizeof(double), hipMemcpyHostToDevice, c.stream));
    FitArgs a{d.x.as<double>(), 1, nx, c.norb, c.nbath, d.om.as<double>(), d.dev.as<double>(), nw, c.compat,
              nullptr, nullptr, d.work.as<double>(), nullptr};
    FCHK(launch_fit(a, c.stream));
    FCHK(hipMemcpyAsync(Dev, d.work.p, (size_t)2 * c.norb * nw * sizeof(double), hipMemcpyDeviceToHost, c.stream));
    FCHK(hipStreamSynchronize(c.stream));
    return MOED_OK;
}

int fit_minimize(const FitCtx& c, int nfit, int nx, double* x, int nw, const double* omega, const double* Dev,
                 double ftol, double* fret, int* iter, int* status) {
    FitDev d;
    int rc = fitdev_init(c, d, nfit, nx, nw, omega, Dev);
    if (rc) return rc;
    std::vector<FitSlot> slots(nfit);
    std::vector<FitTask> tasks;
    tasks.reserve(nfit);
    for (int i = 0; i < nfit; ++i) tasks.push_back(frprmn_task(&slots[i], x + (size_t)i * nx, nx, ftol));
    std::vector<double> xb((size_t)nfit * nx, 0.0), fb(nfit), gb((size_t)nfit * nx);
    std::vector<int32_t> act(nfit);
    for (int i = 0; i < nfit; ++i) tasks[i].h.resume();  // run to the first evaluation
    while (true) {
        bool any = false, anyg = false;
        for (int i = 0; i < nfit; ++i) {
            act[i] = 0;
            if (tasks[i].h.done() || !slots[i].pending) continue;
            any = true;
            act[i] = slots[i].gout ? 2 : 1;
            anyg = anyg || slots[i].gout;
            std::copy(slots[i].x, slots[i].x + nx, xb.begin() + (size_t)i * nx);
        }
        if (!any) break;
        rc = fitdev_eval(c, d, xb.data(), act.data(), fb.data(), anyg ? gb.data() : nullptr);
        if (rc) return rc;
        for (int i = 0; i < nfit; ++i) {
            if (!act[i]) continue;
            FitSlot& s = slots[i];
            s.pending = false;
            if (s.gout)
                std::copy(gb.begin() + (size_t)i * nx, gb.begin() + (size_t)(i + 1) * nx, s.gout);
            else
                s.f = fb[i];
            tasks[i].h.resume();
        }
    }
    for (int i = 0; i < nfit; ++i) {
        if (fret) fret[i] = slots[i].fret;
        if (iter) iter[i] = slots[i].iter;
        if (status) status[i] = tasks[i].h.promise().rc;
    }
    return MOED_OK;
}

}  // namespace moed
