"""moed_minimize_device_noalloc, the minimiser that allocates no device memory, against
moed_minimize (host buffers) on the same inputs: bit-identical fits on two calls in a row, the
second with fewer fits and fewer frequencies (the shared state array keeps the first call's
contents), in both summation orders; more fits than the state array holds are rejected."""
import ctypes as ct
import os

import numpy as np
import pytest

import moed
from moed import EDSolver, MOED_F_FIT_TREE, MOED_EARG

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")


def _dev(a, torch, dtype=None):
    return torch.from_numpy(np.ascontiguousarray(a)).to(device="cuda:0", dtype=dtype)


def _noalloc(s, X, omega, Dev, torch, ftol=1.0e-7):
    nfit, nx = X.shape
    x = _dev(X, torch)
    om = _dev(omega, torch)
    D = _dev(moed._cplx_flat(Dev), torch)
    fret = torch.full((nfit,), -1.0, dtype=torch.float64, device="cuda:0")
    it = torch.full((nfit,), -1, dtype=torch.int32, device="cuda:0")
    st = torch.full((nfit,), -9, dtype=torch.int32, device="cuda:0")
    torch.cuda.synchronize()  # the inputs are written on torch's stream; the call runs on the handle's
    rc = moed.lib().moed_minimize_device_noalloc(s.h, nfit, nx, x.data_ptr(), len(omega), om.data_ptr(),
                                                 D.data_ptr(), ftol, fret.data_ptr(), it.data_ptr(), st.data_ptr(),
                                                 None)
    return rc, x.cpu().numpy(), fret.cpu().numpy(), it.cpu().numpy(), st.cpu().numpy()


@pytest.mark.parametrize("flags", [0, MOED_F_FIT_TREE])
def test_noalloc_minimize_equals_minimize_on_two_calls(flags):
    torch = pytest.importorskip("torch")
    z = np.load(os.path.join(GOLD, "ref_fit.npz"))
    norb, nbath = int(z["norb"]), int(z["nbath"])
    rng = np.random.default_rng(17)
    X = z["x0"][None, :] + 0.03 * rng.standard_normal((9, len(z["x0"]))) * (z["x0"] != 0)
    X[0] = z["x0"]
    om, Dev = z["omega"], z["Dev"]
    print(f"noalloc: flags {flags}, call 1: {X.shape[0]} fits, {len(om)} frequencies", flush=True)
    with EDSolver(norb, nbath, 0.0, 0.0, np.zeros(norb + nbath), np.zeros((norb, nbath)), flags=flags) as s:
        for nfit, nw in ((9, len(om)), (3, len(om) // 2)):
            print(f"noalloc: {nfit} fits, {nw} frequencies", flush=True)
            x_ref, f_ref, it_ref, st_ref = s.minimization(X[:nfit], om[:nw], Dev[:, :nw])
            rc, x, f, it, st = _noalloc(s, X[:nfit], om[:nw], Dev[:, :nw], torch)
            assert rc == 0
            assert np.array_equal(x, x_ref)
            assert np.array_equal(f, f_ref)
            assert np.array_equal(it, it_ref) and np.array_equal(st, st_ref)
            assert np.all(st == 0) and np.all(it > 0)


def test_noalloc_minimize_rejects_too_many_fits_and_bad_inputs():
    torch = pytest.importorskip("torch")
    z = np.load(os.path.join(GOLD, "ref_fit.npz"))
    norb, nbath = int(z["norb"]), int(z["nbath"])
    om, Dev = z["omega"], z["Dev"]
    with EDSolver(norb, nbath, 0.0, 0.0, np.zeros(norb + nbath), np.zeros((norb, nbath))) as s:
        X = np.repeat(z["x0"][None, :], 1025, axis=0)
        rc = _noalloc(s, X, om, Dev, torch)[0]
        assert rc == MOED_EARG
        buf = ct.create_string_buffer(256)
        moed.lib().moed_last_error(s.h, buf, 256)
        assert b"1024" in buf.value
        bad = om.copy()
        bad[3] = -bad[3]
        assert _noalloc(s, X[:2], bad, Dev, torch)[0] == MOED_EARG
        # a good call after the rejected ones still matches the host-buffer minimiser
        rc, x, f, it, st = _noalloc(s, X[:2], om, Dev, torch)
        x_ref, f_ref, _, _ = s.minimization(X[:2], om, Dev)
        assert rc == 0 and np.array_equal(x, x_ref) and np.array_equal(f, f_ref)
