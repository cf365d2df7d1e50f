// Bath fit on the device: batched chi^2 and gradient (func.f90:1-142) and the conjugate-gradient
// minimiser (frprmn.f90), one workgroup per independent starting point, all of them in one launch.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

struct moed_handle;

namespace moed {

struct FitCtx {
    hipStream_t stream;
    int norb, nbath, compat;
    int tree;      // sums in a fixed tree order instead of the reference's sequential order
    int budget;    // CG iterations per fit and launch (0: default)
    int* nfev;     // optional [nfit]: function evaluations of each fit
    std::string* err;
};

// f[nfit], g[nfit*nx] (either may be null); x[nfit*nx]; Dev interleaved complex (norb, nw) column-major
int fit_eval(const FitCtx& c, int nfit, int nx, const double* x, int nw, const double* omega, const double* Dev,
             double* f, double* g);
// Delta(iw) (norb, nw) interleaved complex for one parameter vector x
int fit_delta(const FitCtx& c, int nx, const double* x, int nw, const double* omega, double* Dev);
int fit_minimize(const FitCtx& c, int nfit, int nx, double* x, int nw, const double* omega, const double* Dev,
                 double ftol, double* fret, int* iter, int* status);

// Allocation-free minimisation: x (in/out), omega, Dev and the outputs are device pointers, the CG
// state of at most kNoAllocFits fits lives in a file-scope device array whose address on the current
// device fit_scratch_address returns (looked up once per device).  No device allocation or free.
constexpr int kNoAllocFits = 1024;
constexpr int kMaxDevices = 64;
int fit_scratch_address(int device, void** out);
int fit_minimize_noalloc(const FitCtx& c, void* scratch_addr, int nfit, int nx, double* x, int nw,
                         const double* omega, const double* Dev, double ftol, double* fret, int32_t* iter,
                         int32_t* status);

}  // namespace moed

int fit_delta_host(moed_handle* h, int nx, const double* x, int nw, const double* omega, double* Dev);
int fit_eval_host(moed_handle* h, int nfit, int nx, const double* x, int nw, const double* omega, const double* Dev,
                  double* f, double* g);
int fit_minimize_host(moed_handle* h, int nfit, int nx, double* x, int nw, const double* omega, const double* Dev,
                      double ftol, double* fret, int* iter, int* status);
