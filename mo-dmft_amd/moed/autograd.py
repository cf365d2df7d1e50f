"""torch.autograd.Function wrappers of the device adjoint kernels, for testing them with torch's own
gradient checks.  The forward and backward passes both run in the HIP library; torch tensors are
only the carrier (CPU, float64 / complex128).

* ``Chi2.apply(x, solver, omega, Dev)``: the hybridisation misfit of each row of x (nfit, nx),
  FUNC (func.f90:1-45); its gradient is the dfunc kernel's (func.f90:47-142).
* ``GreenCF.apply(a, b, solver, sign, E, omega)``: the continued fraction G(iw) of green_diag
  (ed_green.F90:136-185) for Lanczos coefficients a, b (0..nsc); its gradient with respect to a
  and b is the adjoint kernel's (moed_green_cf_adjoint).  Build the solver with
  MOED_F_EXACT_CMPLX so that z = E + i w is not rounded to single precision, which the reference
  does and which makes G piecewise constant in E.
"""
import numpy as np
import torch


class Chi2(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, solver, omega, Dev):
        f, g = solver.func_dfunc(x.detach().cpu().numpy(), omega, Dev)
        ctx.save_for_backward(torch.from_numpy(g))
        return torch.from_numpy(f)

    @staticmethod
    def backward(ctx, grad_f):
        (g,) = ctx.saved_tensors
        return grad_f[:, None] * g, None, None, None


class GreenCF(torch.autograd.Function):
    @staticmethod
    def forward(ctx, a, b, solver, sign, E, omega):
        nsc = a.shape[0] - 1
        an, bn = a.detach().cpu().numpy(), b.detach().cpu().numpy()
        ctx.args = (solver, sign, E, nsc, an, bn, np.asarray(omega, np.float64))
        return torch.from_numpy(solver.green_cf(sign, E, nsc, an, bn, omega))

    @staticmethod
    def backward(ctx, grad_g):
        # for a real loss torch hands over dL/dRe g + i dL/dIm g, the adjoint kernel's input
        solver, sign, E, nsc, an, bn, omega = ctx.args
        ga, gb = solver.green_cf_adjoint(sign, E, nsc, an, bn, omega, grad_g.detach().cpu().numpy())
        return torch.from_numpy(ga), torch.from_numpy(gb), None, None, None, None
