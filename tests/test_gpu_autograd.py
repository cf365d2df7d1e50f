"""torch.autograd.gradcheck of the device adjoint kernels through moed.autograd: the chi^2
hybridisation misfit (FUNC with the dfunc kernel's gradient, both summation orders) and the
continued-fraction G(iw) with respect to the Lanczos coefficients (the adjoint kernel).  gradcheck
compares each backward pass with central finite differences of the forward pass."""
import numpy as np
import pytest
import torch

from moed import EDSolver, MOED_F_EXACT_CMPLX, MOED_F_FIT_TREE, dmft
from moed.autograd import Chi2, GreenCF
from moed.model import omega_grid

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("flags", [0, MOED_F_FIT_TREE])
def test_chi2_gradcheck(flags):
    norb, nbath = 2, 4
    rng = np.random.default_rng(31)
    om = omega_grid(256, 40.0)
    ek = rng.uniform(-2, 2, norb + nbath)
    vk = rng.uniform(0.2, 0.8, (norb, nbath))
    x0 = dmft.ev_to_x(ek, vk, ek[:norb], norb, nbath)
    X = torch.tensor(x0[None, :] * (1.0 + 0.05 * rng.standard_normal((3, len(x0)))), dtype=torch.float64,
                     requires_grad=True)
    with EDSolver(norb, nbath, 0.0, 0.0, np.zeros(norb + nbath), np.zeros((norb, nbath)), flags=flags) as s:
        Dev = s.calc_dev(ek, vk, ek[:norb], om)  # the target: x0's own hybridisation function
        assert torch.autograd.gradcheck(lambda x: Chi2.apply(x, s, om, Dev), (X,), eps=1e-6, atol=1e-9, rtol=1e-5)
        # and the minimum: zero misfit and zero gradient at x0 itself
        x = torch.tensor(x0[None, :], dtype=torch.float64, requires_grad=True)
        f = Chi2.apply(x, s, om, Dev)
        f.sum().backward()
        assert f.item() <= 1e-28 and torch.max(torch.abs(x.grad)).item() <= 1e-14


@pytest.mark.parametrize("sign", [1, -1])
def test_green_cf_gradcheck(sign):
    rng = np.random.default_rng(9)
    n = 12
    a = torch.tensor(rng.uniform(-2, 2, n + 1), dtype=torch.float64, requires_grad=True)
    b = torch.tensor(rng.uniform(0.1, 1.0, n + 1), dtype=torch.float64, requires_grad=True)
    om = omega_grid(64, 20.0)
    with EDSolver(1, 1, 0.0, 0.0, [0.0, 0.0], [[0.5]], flags=MOED_F_EXACT_CMPLX) as s:
        assert torch.autograd.gradcheck(lambda aa, bb: GreenCF.apply(aa, bb, s, sign, 0.3, om), (a, b), eps=1e-6,
                                        atol=1e-9, rtol=1e-5)
