"""GPU parity of the batched bath fit: FUNC/dfunc (func.f90) in the reference's operation order,
calc_dev, and the CG minimiser (frprmn.f90) against the reference build's trajectory; the
tree-order mode against the oracle within rounding."""
import os

import numpy as np
import pytest

import oracle
from moed import EDSolver, MOED_F_FIT_TREE
from moed.model import config, omega_grid

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")


@pytest.fixture(scope="module")
def fit():
    z = np.load(os.path.join(GOLD, "ref_fit.npz"))
    norb, nbath = int(z["norb"]), int(z["nbath"])
    s = EDSolver(norb, nbath, 0.0, 0.0, np.zeros(norb + nbath), np.zeros((norb, nbath)))
    yield z, s, norb, nbath
    s.close()


def test_func_dfunc_bit_exact_with_reference(fit):
    z, s, norb, nbath = fit
    f, g = s.func_dfunc(z["x0"], z["omega"], z["Dev"])
    assert f[0] == float(z["f"])
    assert np.array_equal(g[0], z["g"])


def test_func_dfunc_batched_equals_oracle(fit):
    z, s, norb, nbath = fit
    rng = np.random.default_rng(1)
    X = z["x0"][None, :] + 0.05 * rng.standard_normal((37, len(z["x0"]))) * (z["x0"] != 0)
    f, g = s.func_dfunc(X, z["omega"], z["Dev"])
    for i in range(len(X)):
        assert f[i] == oracle.func(X[i], z["omega"], z["Dev"], norb, nbath)
        assert np.array_equal(g[i], oracle.dfunc(X[i], z["omega"], z["Dev"], norb, nbath))


def test_calc_dev_bit_exact(fit):
    z, s, norb, nbath = fit
    ef, ek, vk = oracle.x_to_ev(z["x0"], norb, nbath)
    D = s.calc_dev(ek, vk, ef, z["omega"])
    Do = oracle.calc_dev(oracle.Model(norb, nbath, 0, 0, ek, vk), ef, z["omega"])
    assert np.array_equal(D, Do)


def test_minimization_reproduces_reference_trajectory(fit):
    z, s, norb, nbath = fit
    x, fret, it, st = s.minimization(z["x0"], z["omega"], z["Dev"])
    print("fret", fret[0], "ref", float(z["fret"]), "iters", it[0], flush=True)
    assert st[0] == 0
    assert np.max(np.abs(x[0] - z["xmin"])) <= 1e-8
    assert abs(fret[0] - float(z["fret"])) <= 1e-8 * max(1e-12, abs(float(z["fret"]))) + 1e-20
    rc, xo, fo, ito, nfo = oracle.frprmn(z["x0"], z["omega"], z["Dev"], norb, nbath)
    assert np.array_equal(x[0], xo) and fret[0] == fo and it[0] == ito
    assert s.fit_nfev(1)[0] == nfo  # the same line searches, evaluation for evaluation


def test_minimization_batched_independent_starts(fit):
    z, s, norb, nbath = fit
    rng = np.random.default_rng(3)
    X = z["x0"][None, :] + 0.02 * rng.standard_normal((6, len(z["x0"]))) * (z["x0"] != 0)
    X[0] = z["x0"]
    x, fret, it, st = s.minimization(X, z["omega"], z["Dev"])
    for i in range(len(X)):
        rc, xo, fo, ito, _ = oracle.frprmn(X[i], z["omega"], z["Dev"], norb, nbath)
        assert st[i] == rc == 0
        assert it[i] == ito and np.max(np.abs(x[i] - xo)) <= 1e-8


def test_results_do_not_depend_on_the_launch_budget(fit):
    """the minimiser suspends at CG iteration boundaries between launches: one iteration per
    launch and the default give bit-identical fits"""
    z, s, norb, nbath = fit
    rng = np.random.default_rng(5)
    X = z["x0"][None, :] + 0.02 * rng.standard_normal((3, len(z["x0"]))) * (z["x0"] != 0)
    out = []
    for budget in (0, 1, 7):
        s.set_fit_launch(budget)
        out.append(s.minimization(X, z["omega"], z["Dev"]) + (s.fit_nfev(3),))
    s.set_fit_launch(0)
    for r in out[1:]:
        for a, b in zip(r, out[0]):
            assert np.array_equal(a, b)


@pytest.fixture(scope="module")
def fit_tree():
    z = np.load(os.path.join(GOLD, "ref_fit.npz"))
    norb, nbath = int(z["norb"]), int(z["nbath"])
    s = EDSolver(norb, nbath, 0.0, 0.0, np.zeros(norb + nbath), np.zeros((norb, nbath)), flags=MOED_F_FIT_TREE)
    yield z, s, norb, nbath
    s.close()


def test_tree_order_func_dfunc_within_rounding(fit_tree):
    """tree-order sums: chi^2 and gradient agree with the oracle to summation rounding"""
    z, s, norb, nbath = fit_tree
    rng = np.random.default_rng(2)
    X = z["x0"][None, :] + 0.05 * rng.standard_normal((9, len(z["x0"]))) * (z["x0"] != 0)
    f, g = s.func_dfunc(X, z["omega"], z["Dev"])
    for i in range(len(X)):
        fo = oracle.func(X[i], z["omega"], z["Dev"], norb, nbath)
        go = oracle.dfunc(X[i], z["omega"], z["Dev"], norb, nbath)
        assert abs(f[i] - fo) <= 1e-13 * abs(fo)
        assert np.max(np.abs(g[i] - go)) <= 1e-12 * np.abs(go).max()


def test_tree_order_fit_converges_to_the_reference_minimum(fit_tree):
    z, s, norb, nbath = fit_tree
    x, fret, it, st = s.minimization(z["x0"], z["omega"], z["Dev"])
    fr, f0 = float(z["fret"]), float(z["f"])
    print("tree fret", fret[0], "ref", fr, "start", f0, "iters", it[0], flush=True)
    assert st[0] == 0
    # the fixture's target is fitted exactly: both minima are chi^2 ~ 0 on the scale of the start
    assert fret[0] <= 1e-8 * f0 and abs(fret[0] - fr) <= 1e-8 * f0


def test_tree_order_fit_is_independent_of_the_batch(fit_tree):
    z, s, norb, nbath = fit_tree
    rng = np.random.default_rng(8)
    X = z["x0"][None, :] + 0.02 * rng.standard_normal((5, len(z["x0"]))) * (z["x0"] != 0)
    xb, fb, ib, sb = s.minimization(X, z["omega"], z["Dev"])
    for i in (0, 3):
        x1, f1, i1, s1 = s.minimization(X[i], z["omega"], z["Dev"])
        assert np.array_equal(x1[0], xb[i]) and f1[0] == fb[i] and i1[0] == ib[i]


def test_func_dfunc_at_the_15_site_fit_shape():
    """Norb 3, Nbath 12 (nx 51), 4096 frequencies at beta 200: reference order bit-exact with the
    oracle, tree order within rounding"""
    c = config("ns15_fit")
    norb, nbath = c.norb, c.nbath
    om = omega_grid(c.nw, c.beta)
    from moed import dmft
    x0 = dmft.ev_to_x(c.ek, c.vk, c.ek[:norb], norb, nbath)
    rng = np.random.default_rng(15)
    X = x0[None, :] * (1.0 + 0.1 * rng.standard_normal((4, len(x0))))
    ef, ek, vk = oracle.x_to_ev(X[3], norb, nbath)
    Dev = oracle.calc_dev(oracle.Model(norb, nbath, 0, 0, ek, vk), ef, om)
    ref = [(oracle.func(x, om, Dev, norb, nbath), oracle.dfunc(x, om, Dev, norb, nbath)) for x in X]
    # scale of the terms: chi^2 and gradient against a zero target (X[3] fits Dev exactly)
    zero = np.zeros_like(Dev)
    scale = [(oracle.func(x, om, zero, norb, nbath), np.abs(oracle.dfunc(x, om, zero, norb, nbath)).max()) for x in X]
    for flags in (0, MOED_F_FIT_TREE):
        with EDSolver(norb, nbath, 0.0, 0.0, np.zeros(norb + nbath), np.zeros((norb, nbath)), flags=flags) as s:
            f, g = s.func_dfunc(X, om, Dev)
        for i in range(len(X)):
            if flags == 0:
                assert f[i] == ref[i][0] and np.array_equal(g[i], ref[i][1])
            else:
                assert abs(f[i] - ref[i][0]) <= 1e-13 * scale[i][0]
                assert np.max(np.abs(g[i] - ref[i][1])) <= 1e-12 * scale[i][1]


def test_fit_rejects_bad_inputs(fit):
    z, s, norb, nbath = fit
    om = z["omega"].copy()
    with pytest.raises(Exception):
        s.minimization(z["x0"], -om, z["Dev"])
    D = z["Dev"].copy()
    D[0, 0] = np.nan
    with pytest.raises(Exception):
        s.minimization(z["x0"], om, D)
    x = z["x0"].copy()
    x[0] = np.inf
    with pytest.raises(Exception):
        s.func_dfunc(x, om, z["Dev"])


def test_exact_fit_recovers_parameters():
    norb, nbath = 1, 2
    om = omega_grid(128, 20.0)
    xt = np.array([-0.7, 0.9, 0.5, 0.4, 0.1])
    with EDSolver(norb, nbath, 0.0, 0.0, np.zeros(3), np.zeros((1, 2))) as s:
        ef, ek, vk = oracle.x_to_ev(xt, norb, nbath)
        Dev = s.calc_dev(ek, vk, ef, om)
        x, fret, it, st = s.minimization(xt + np.array([0.05, -0.05, 0.04, -0.03, 0.02]), om, Dev)
    assert st[0] == 0 and fret[0] < 1e-12
    assert np.allclose(np.abs(x[0]), np.abs(xt), atol=1e-4)


# ------------------------------------------------------------------ lattice step and the DMFT loop

def test_delta_new_bit_exact_with_reference():
    """device lattice step (ed_delta_new) against the reference build, H(k) from moed_band_tb"""
    import moed
    z = np.load(os.path.join(GOLD, "ref_delta.npz"))
    norb = int(z["norb"])
    hk = moed.band_tb(norb, int(z["nq"]))
    assert np.array_equal(hk, z["hk"])
    with EDSolver(norb, 12, 4.0, 0.6, np.zeros(norb + 12), np.zeros((norb, 12)), beta=200.0) as s:
        gks, dn, nocc = s.delta_new(z["Gr"], z["Dev"], hk, float(z["rmu"]))
    assert np.array_equal(gks, z["gksum"]) and np.array_equal(dn, z["dev_new"])
    assert np.array_equal(nocc, z["nocc"])  # n_from_gksum of the reference build


def test_dmft_loop_steps_match_the_oracle():
    """two iterations of the self-consistency loop on a 6-site, 2-orbital model: every step's
    output against the oracle applied to the same inputs (lattice step bit-exact, fit to the
    reference's CG trajectory within 1e-8, calc_dev bit-exact), and the ground state of the
    second iteration's bath against dense ED"""
    import moed
    from moed import dmft
    from oracle import dense_ed
    rng = np.random.default_rng(12)
    norb, nbath = 2, 4
    ns = norb + nbath
    ek = np.concatenate([np.full(norb, -2.0), rng.uniform(-2, 2, nbath)])
    vk = np.zeros((norb, nbath))
    for b in range(nbath):
        vk[b % norb, b] = rng.uniform(0.3, 0.9)
    U, J, beta, mu = 4.0, 0.5, 10.0, 2.0
    om = omega_grid(128, beta)
    hk = moed.band_tb(norb, 100)
    with EDSolver(norb, nbath, U, J, ek, vk, beta=beta, nstep=40, nev=2, sectors=[(3, 3)]) as s:
        ef = ek[:norb].copy()
        Dev = s.calc_dev(ek, vk, ef, om)
        assert np.array_equal(Dev, oracle.calc_dev(oracle.Model(norb, nbath, U, J, ek, vk), ef, om))
        for it in range(2):
            kept = s.ed_solve()
            sec = dense_ed.Sector(ns, 3, 3)
            w = np.linalg.eigvalsh(dense_ed.hamiltonian(sec, norb, nbath, U, J, ek, vk).toarray())
            assert abs(kept[0][0] - w[0]) <= 1e-10 * abs(w[0])
            G = s.ed_calc_green_ftn(om)
            _, Dn, _ = s.delta_new(G, Dev, hk, mu)
            _, Dn_ref = oracle.delta_new(G, Dev, hk, mu)
            assert np.array_equal(Dn, Dn_ref)
            x0 = dmft.ev_to_x(ek, vk, ef, norb, nbath)
            assert np.array_equal(x0, oracle.ev_to_x(ek, vk.reshape(-1, order="F"), ef, norb, nbath))
            X, fret, _, st = s.minimization(x0, om, Dn)
            rc, xr, fr, _, _ = oracle.frprmn(x0, om, Dn, norb, nbath)
            assert st[0] == 0 and rc == 0
            assert np.max(np.abs(X[0] - xr)) <= 1e-8 * max(1.0, np.abs(xr).max())
            ef, ek, vk = dmft.x_to_ev(X[0], norb, nbath)
            s.set_bath(ek, vk)
            Dev = s.calc_dev(ek, vk, ef, om)


def _sequential(s, t):
    for v in t.tolist():
        s = s + v
    return s


def test_ordered_sum_replay_equals_sequential_loop():
    """The block-wide replay behind the reference-order chi^2 (moed_osum.h) against the plain loop
    s = s + t(i), bit for bit, on inputs built to hit its fallbacks: exact ties (powers of two,
    ulp-sized multiples), sums crossing many binades, jumps past the pass's binades, zeros,
    subnormals, negative, NaN and infinite terms; several chunks of 1024 threads."""
    import ctypes as ct
    from moed import lib
    L = lib()
    f = L.moed_check_ordered_sum
    f.argtypes = [ct.c_void_p, ct.c_int, ct.c_int, ct.c_int, ct.c_double, ct.POINTER(ct.c_double)]
    f.restype = ct.c_int
    rng = np.random.default_rng(7)
    out = ct.c_double()
    ncase = 0
    for kind in range(10):
        for rep in range(12):
            no = 1 + rep % 4
            n = int(rng.integers(1, 3 * 1024 * no + 1))
            i = np.arange(n)
            if kind == 0:
                t = rng.random(n)
            elif kind == 1:
                t = np.ldexp(rng.random(n), -rng.integers(0, 60, n))
            elif kind == 2:
                t = 1.0 / (i + 1.0) ** 3
            elif kind == 3:
                t = np.where(rng.integers(0, 5, n) == 0, 0.0, np.ldexp(1.0, -rng.integers(0, 80, n)))
            elif kind == 4:
                t = np.ldexp(rng.integers(0, 7, n).astype(float), -rng.integers(50, 58, n))
            elif kind == 5:
                t = np.where(rng.integers(0, 50, n) == 0, np.ldexp(rng.random(n), 40), 1e-3 * rng.random(n))
            elif kind == 6:
                t = np.where(rng.integers(0, 200, n) == 0, -rng.random(n), rng.random(n))
            elif kind == 7:
                t = rng.random(n)
                t[n // 3] = np.inf
                if rep % 2:
                    t[n // 2] = np.nan
            elif kind == 8:
                t = np.ldexp(rng.random(n), -1060)
            else:  # chi^2-like: |r|^2 / w_n over Matsubara frequencies, a few orbitals
                w = (2 * (i // no) + 1) * np.pi / 200.0
                t = (rng.random(n) * 1e-3 / (1.0 + w)) ** 2 / w
            t = np.ascontiguousarray(t, dtype=np.float64)
            s0 = 0.0 if rep % 3 == 0 else float(np.ldexp(rng.random() + 0.5, int(rng.integers(-30, 30))))
            if kind == 4:
                s0 = float(2.0 ** (rep % 3))
            nseq = 32 if rep % 2 else 0
            rc = f(t.ctypes.data, n, no, nseq, s0, ct.byref(out))
            assert rc == 0
            ref = _sequential(s0, t)
            ok = (np.isnan(ref) and np.isnan(out.value)) or np.float64(ref).tobytes() == np.float64(out.value).tobytes()
            assert ok, (kind, rep, n, no, nseq, s0, out.value, ref)
            ncase += 1
    assert ncase == 120
