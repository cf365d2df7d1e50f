"""Pins of the CPU oracle (oracle/oracle.c): reference-build golden vectors, the notebook
known-answer matrix, an independent brute-force second quantisation, closed forms and the
reference's own finite-difference recipe.  CPU only."""
import json
import os

import numpy as np
import pytest

import oracle
from oracle import dense_ed
from moed.model import load_configs, omega_grid, start_vector

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def model_of(c):
    return oracle.Model(c.norb, c.nbath, c.U, c.J, c.ek, c.vk, c.tij)


@pytest.fixture(scope="module")
def cfg8():
    return load_configs()[0]


# ------------------------------------------------------------------ basis and combinatorics

def test_icom_matches_binomial_and_reference_edges():
    from math import comb
    for n in range(1, 16):
        for m in range(0, n + 1):
            assert oracle.icom(n, m) == comb(n, m)
    assert oracle.icom(8, -1) == 1  # ed_utils.f90:54-69 returns 1 for m < 0


def test_basis_is_ascending_popcount_enumeration():
    b = oracle.Basis(10, 4, 3)
    assert np.array_equal(b.up, dense_ed.strings(10, 4))
    assert np.array_equal(b.dn, dense_ed.strings(10, 3))
    # colex rank closed form: rank(s) = sum_k C(p_k, k) over set bits p_1 < p_2 < ...
    from math import comb
    for r, s in enumerate(b.up):
        bits = [p for p in range(10) if (s >> p) & 1]
        assert sum(comb(p, k + 1) for k, p in enumerate(bits)) == r


# ------------------------------------------------------------------ known answers

def test_notebook_kat_eigenvalues():
    kat = json.load(open(os.path.join(GOLD, "kat.json")))
    m = oracle.Model(kat["norb"], kat["nbath"], kat["U"], kat["J"], kat["ek"], np.array(kat["vk"]).reshape(2, 1))
    b = oracle.Basis(3, *kat["sector"])
    H = np.array([oracle.multiply_H(m, b, e) for e in np.eye(b.ntot)]).T
    assert np.array_equal(H, H.T)
    w = np.linalg.eigvalsh(H)
    assert np.allclose(w, kat["eigenvalues"], rtol=0, atol=5e-9)


def test_dense_ed_matches_oracle_elements_and_orderings():
    rng = np.random.default_rng(7)
    for norb, nbath, nu, nd, tij in [(2, 4, 3, 2, False), (3, 3, 3, 3, True), (1, 5, 3, 2, False),
                                     (4, 2, 3, 4, True)]:
        ns = norb + nbath
        ek = rng.uniform(-2, 2, ns)
        vk = rng.uniform(0.1, 1.0, (norb, nbath))
        t = None
        if tij:
            t = rng.uniform(-0.3, 0.3, (norb, norb))
            t = t + t.T
            np.fill_diagonal(t, 0.0)
        m = oracle.Model(norb, nbath, 3.0, 0.4, ek, vk, t)
        b = oracle.Basis(ns, nu, nd)
        s = dense_ed.Sector(ns, nu, nd)
        Hd = dense_ed.hamiltonian(s, norb, nbath, 3.0, 0.4, ek, vk, t).toarray()
        Ho = np.zeros_like(Hd)
        for i, (c, v) in enumerate(oracle.h_rows(m, b)):
            Ho[i, c] = v
        off = ~np.eye(len(Hd), dtype=bool)
        assert np.array_equal(Ho[off], Hd[off])  # off-diagonal elements are single terms: exact
        assert np.allclose(np.diag(Ho), np.diag(Hd), rtol=1e-15, atol=1e-14)
        Hi = dense_ed.hamiltonian(s, norb, nbath, 3.0, 0.4, ek, vk, t, mode=dense_ed.interleaved_mode).toarray()
        assert np.allclose(np.linalg.eigvalsh(Hi), np.linalg.eigvalsh(Ho), atol=1e-12)


def test_noninteracting_limit_closed_form():
    rng = np.random.default_rng(3)
    norb, nbath, nu, nd = 2, 4, 3, 2
    ns = norb + nbath
    ek = rng.uniform(-2, 2, ns)
    vk = rng.uniform(0.2, 0.8, (norb, nbath))
    m = oracle.Model(norb, nbath, 0.0, 0.0, ek, vk)
    b = oracle.Basis(ns, nu, nd)
    H = np.array([oracle.multiply_H(m, b, e) for e in np.eye(b.ntot)]).T
    h1 = np.diag(ek)
    h1[:norb, norb:] = vk
    h1[norb:, :norb] = vk.T
    e1 = np.sort(np.linalg.eigvalsh(h1))
    assert np.isclose(np.linalg.eigvalsh(H)[0], e1[:nu].sum() + e1[:nd].sum(), atol=1e-12)


def test_atomic_limit_is_diagonal():
    ek = np.array([-1.0, -0.5, 0.3, 0.7])
    m = oracle.Model(2, 2, 4.0, 0.5, ek, np.zeros((2, 2)))
    b = oracle.Basis(4, 2, 2)
    for i, (c, v) in enumerate(oracle.h_rows(m, b)):
        st = oracle.lib  # noqa: F841
        assert list(c) == [i] or all(abs(x) == 0.5 for x in v[c != i])  # only J terms off the diagonal


# ------------------------------------------------------------------ reference-build golden vectors

def test_hv_matches_reference_build(cfg8):
    z = np.load(os.path.join(GOLD, "ref_ns8.npz"))
    m = model_of(cfg8)
    b = oracle.Basis(cfg8.nsite, *cfg8.sectors[0])
    x = start_vector(int(z["x_seed"]), b.ntot)
    assert np.array_equal(oracle.multiply_H(m, b, x), z["hv_y"])


def test_h_elements_digest_matches_reference_build(cfg8):
    z = np.load(os.path.join(GOLD, "ref_ns8.npz"))
    m = model_of(cfg8)
    b = oracle.Basis(cfg8.nsite, *cfg8.sectors[0])
    assert oracle.h_digest(oracle.h_rows(m, b)) == str(z["h_digest"])
    assert np.array_equal(oracle.diag_elements(m, b), z["h_diag"])


def test_lanczos_matches_reference_build(cfg8):
    z = np.load(os.path.join(GOLD, "ref_ns8.npz"))
    m = model_of(cfg8)
    b = oracle.Basis(cfg8.nsite, *cfg8.sectors[0])
    x = start_vector(int(z["x_seed"]), b.ntot)
    nsc, a, bb = oracle.lanczos(m, b, x, cfg8.nstep)
    assert nsc == int(z["lz_nstep_calc"])
    assert np.array_equal(a, z["lz_a"]) and np.array_equal(bb, z["lz_b"])


def test_apply_c_matches_reference_build(cfg8):
    z = np.load(os.path.join(GOLD, "ref_ns8.npz"))
    b = oracle.Basis(cfg8.nsite, *cfg8.sectors[0])
    x = start_vector(int(z["x_seed"]), b.ntot)
    for pm in (1, 2):
        for iorb in (1, 2):
            for ispin in (1, 2):
                _, out = oracle.apply_c(b, x, pm, iorb, ispin)
                assert np.array_equal(out, z[f"applyc_{pm}{iorb}{ispin}"])


def test_hv_12site_matches_reference_summary():
    c = load_configs()[1]
    z = np.load(os.path.join(GOLD, "ref_ns12.npz"))
    m = model_of(c)
    b = oracle.Basis(c.nsite, *c.sectors[0])
    y = oracle.multiply_H(m, b, start_vector(int(z["x_seed"]), b.ntot))
    assert np.array_equal(y[z["idx"]], z["y_at"])
    import hashlib
    assert hashlib.sha256(y.tobytes()).hexdigest() == str(z["y_sha"])


def test_func_dfunc_frprmn_match_reference_build():
    z = np.load(os.path.join(GOLD, "ref_fit.npz"))
    norb, nbath = int(z["norb"]), int(z["nbath"])
    assert oracle.func(z["x0"], z["omega"], z["Dev"], norb, nbath) == float(z["f"])
    assert np.array_equal(oracle.dfunc(z["x0"], z["omega"], z["Dev"], norb, nbath), z["g"])
    rc, x, fret, it, nf = oracle.frprmn(z["x0"], z["omega"], z["Dev"], norb, nbath)
    assert rc == 0 and fret == float(z["fret"]) and np.array_equal(x, z["xmin"])


def test_delta_new_matches_reference_build():
    """lattice step (band structure, k-sum, new hybridisation, occupations) against the reference build"""
    z = np.load(os.path.join(GOLD, "ref_delta.npz"))
    hk = oracle.band_tb(int(z["norb"]), int(z["nq"]))
    assert np.array_equal(hk, z["hk"])
    gks, dn = oracle.delta_new(z["Gr"], z["Dev"], hk, float(z["rmu"]))
    assert np.array_equal(gks, z["gksum"]) and np.array_equal(dn, z["dev_new"])
    assert np.array_equal(oracle.n_from_gksum(gks, float(z["beta"])), z["nocc"])  # n_from_gksum


def test_delta_new_closed_forms():
    """one k point with Hk = 0: Gksum = 1/(1/G + D + mu), so D_new = 1/G - 1/Gksum + D = -mu up to
    rounding; flat band: the k-sum is the same k point repeated"""
    rng = np.random.default_rng(3)
    G = rng.standard_normal((2, 8)) - 1j * np.abs(rng.standard_normal((2, 8)))
    D = rng.standard_normal((2, 8)) + 1j * rng.standard_normal((2, 8))
    gks, dn = oracle.delta_new(G, D, np.zeros((1, 2)), 0.75)
    assert np.allclose(gks, 1 / (1 / G + D + 0.75), rtol=1e-14)
    assert np.allclose(dn, -0.75, atol=1e-12)
    gks4, dn4 = oracle.delta_new(G, D, np.full((4, 2), 0.2), 0.75)
    gks1, dn1 = oracle.delta_new(G, D, np.full((1, 2), 0.2), 0.75)
    assert np.allclose(gks4, gks1, rtol=1e-14) and np.allclose(dn4, dn1, rtol=1e-13)
    # band structure: cos grid, average zero on a full period
    hk = oracle.band_tb(1, 100)
    assert hk.shape == (100, 1) and abs(hk.mean()) < 1e-14 and hk.min() == -1.0


def test_sort_matches_reference_build():
    z = np.load(os.path.join(GOLD, "ref_sort.npz"))
    assert np.array_equal(oracle.sort_index(z["a"]), z["indx"])


# ------------------------------------------------------------------ Lanczos, Green's function, fit: math pins

def test_lanczos_first_coefficients_closed_form():
    rng = np.random.default_rng(11)
    m = oracle.Model(2, 3, 3.0, 0.5, rng.uniform(-1, 1, 5), rng.uniform(0.2, 0.9, (2, 3)))
    b = oracle.Basis(5, 2, 3)
    v = rng.standard_normal(b.ntot)
    nsc, a, bb = oracle.lanczos(m, b, v, 10)
    vn = v / np.linalg.norm(v)
    Hv = oracle.multiply_H(m, b, vn)
    a0 = vn @ Hv
    r = Hv - a0 * vn
    assert np.isclose(a[0], a0, rtol=1e-14)
    assert np.isclose(bb[1], r @ r, rtol=1e-13)
    assert bb[0] == 0.0


def test_lanczos_spectrum_when_krylov_space_is_exhausted():
    rng = np.random.default_rng(12)
    m = oracle.Model(2, 2, 2.0, 0.3, rng.uniform(-1, 1, 4), rng.uniform(0.3, 0.9, (2, 2)))
    b = oracle.Basis(4, 1, 1)  # D = 16
    H = np.array([oracle.multiply_H(m, b, e) for e in np.eye(b.ntot)]).T
    w, V = np.linalg.eigh(H)
    # without reorthogonalisation the 40-step Ritz values hold every eigenvalue (plus ghost copies)
    nsc, a, bb = oracle.lanczos(m, b, rng.standard_normal(b.ntot), 40)
    T = np.diag(a[:nsc + 1]) + np.diag(np.sqrt(bb[1:nsc + 1]), 1) + np.diag(np.sqrt(bb[1:nsc + 1]), -1)
    ritz = np.linalg.eigvalsh(T)
    for e in w:
        assert np.min(np.abs(ritz - e)) < 1e-8
    # a start vector inside a 3-dimensional invariant subspace breaks down at j = 3 (ed_lanczos.F90:62)
    v = V[:, [0, 5, 9]] @ np.array([0.6, -0.3, 0.5])
    nsc, a, bb = oracle.lanczos(m, b, v, 40)
    assert nsc == 3 and bb[3] < 1e-12 and np.all(a[4:] == 0) and np.all(bb[4:] == 0)
    T = np.diag(a[:3]) + np.diag(np.sqrt(bb[1:3]), 1) + np.diag(np.sqrt(bb[1:3]), -1)
    assert np.allclose(np.linalg.eigvalsh(T), w[[0, 5, 9]], atol=1e-10)


def test_green_diag_equals_lehmann_sum():
    rng = np.random.default_rng(5)
    norb, nbath, nu, nd = 2, 3, 2, 3
    ns = norb + nbath
    ek = rng.uniform(-1.5, 1.5, ns)
    vk = rng.uniform(0.2, 0.8, (norb, nbath))
    U, J = 3.0, 0.4
    m = oracle.Model(norb, nbath, U, J, ek, vk)
    b = oracle.Basis(ns, nu, nd)
    H = np.array([oracle.multiply_H(m, b, e) for e in np.eye(b.ntot)]).T
    w, V = np.linalg.eigh(H)
    om = omega_grid(64, 50.0)
    for iorb in (1, 2):
        for ispin in (1, 2):
            G, (nsp, ap, bp), (nsn, an, bn) = oracle.green_diag(m, b, V[:, 0], w[0], 1.0, 1.0, iorb, ispin, 200,
                                                                om, compat=False)
            assert np.isclose(bp[0] + bn[0], 1.0)  # sum rule of ed_green.F90:174
            E0, GL = dense_ed.lehmann_green(norb, nbath, U, J, ek, vk, nu, nd, iorb - 1, ispin - 1, om)
            assert np.isclose(E0, w[0], atol=1e-12)
            assert np.max(np.abs(G - GL)) < 1e-9 * np.max(np.abs(GL))


def test_green_cf_single_level_closed_form_and_compat_rounding():
    om = omega_grid(16, 10.0)
    a, b = np.array([0.3]), np.array([0.7])
    g = oracle.green_cf(+1, 1.1, 0, a, b, om, compat=False)
    assert np.allclose(g, 0.7 / (1.1 + 1j * om - 0.3), rtol=1e-15)
    gc = oracle.green_cf(+1, 1.1, 0, a, b, om, compat=True)
    z = np.float32(1.1).astype(np.float64) + 1j * om.astype(np.float32).astype(np.float64)
    assert np.allclose(gc, 0.7 / (z - 0.3), rtol=1e-15)
    gh = oracle.green_cf(-1, 1.1, 0, a, b, om, compat=False)
    assert np.allclose(gh, 0.7 / (-1.1 + 1j * om + 0.3), rtol=1e-15)


def test_boltzmann_and_selection_closed_form():
    e = np.array([0.0, 0.01, 0.5, 5.0])
    beta = 10.0
    p = oracle.boltzmann(e, beta)
    Z = np.exp(-beta * e).sum()
    assert np.allclose(p[:3], np.exp(-beta * e[:3]) / Z, rtol=1e-14)
    assert p[3] == 0.0  # -beta*(E_k - E_i) > 10 cuts the state (ed_solver.F90:207)
    kept = oracle.select_states(np.array([0.5, 0.0, 0.01, 5.0]), 2, beta)
    assert [k[2:] for k in kept] == [(1, 2), (2, 1)] or len(kept) == 2


def test_omega_grid_matches_formula():
    om = oracle.omega_grid(100, 37.0)
    assert np.array_equal(om, omega_grid(100, 37.0))
    assert np.allclose(om, (2 * np.arange(100) + 1) * np.pi / 37.0, rtol=1e-15)


def test_dfunc_matches_central_differences():
    z = np.load(os.path.join(GOLD, "ref_fit.npz"))
    norb, nbath = int(z["norb"]), int(z["nbath"])
    x = z["x0"].copy()
    g = oracle.dfunc(x, z["omega"], z["Dev"], norb, nbath, compat=False)
    h = 1e-5  # the reference's own recipe, func.f90:63-70
    for i in range(len(x)):
        xp, xm = x.copy(), x.copy()
        xp[i] += h
        xm[i] -= h
        fd = (oracle.func(xp, z["omega"], z["Dev"], norb, nbath, compat=False) -
              oracle.func(xm, z["omega"], z["Dev"], norb, nbath, compat=False)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1e-3, abs(g[i])) + 1e-9


def test_frprmn_reaches_exact_minimum_of_representable_target():
    norb, nbath = 1, 2
    om = omega_grid(128, 20.0)
    xt = np.array([-0.7, 0.9, 0.5, 0.4, 0.1])
    ef, ek, vk = oracle.x_to_ev(xt, norb, nbath)
    Dev = oracle.calc_dev(oracle.Model(norb, nbath, 0, 0, ek, vk), ef, om, compat=False)
    x0 = xt + np.array([0.05, -0.05, 0.04, -0.03, 0.02])
    rc, x, fret, it, nf = oracle.frprmn(x0, om, Dev, norb, nbath, compat=False)
    assert rc == 0 and fret < 1e-12
    assert np.allclose(np.abs(x), np.abs(xt), atol=1e-4)  # V enters squared


def test_ev_to_x_roundtrip():
    rng = np.random.default_rng(2)
    norb, nbath = 3, 5
    ek = rng.standard_normal(norb + nbath)
    vk = rng.standard_normal((norb, nbath))
    ef = ek[:norb].copy()
    x = oracle.ev_to_x(ek, vk, ef, norb, nbath)
    ef2, ek2, vk2 = oracle.x_to_ev(x, norb, nbath)
    assert np.array_equal(ef2, ef) and np.array_equal(ek2, ek) and np.array_equal(vk2, vk)


# ------------------------------------------------------------------ real-axis spectral function

def test_dos_diag_matches_reference_build():
    """dos_diag (dos.f90:65-106) against the reference build's outputs, bit for bit"""
    z = np.load(os.path.join(GOLD, "ref_dos.npz"))
    for i in range(int(z["ncase"])):
        a = oracle.dos_diag(z["omega"], float(z[f"small_{i}"]), float(z[f"ev_{i}"]), z[f"ap_{i}"], z[f"bp_{i}"],
                            z[f"an_{i}"], z[f"bn_{i}"])
        assert np.array_equal(a, z[f"aff_{i}"]), i


def test_dos_diag_single_level_closed_form():
    """one-step fractions: A(w) = -Im(b+/(w + E + i eta - a+) + b-/(w - E + i eta + a-))/pi, evaluated
    with exact double z (no REAL*4 rounding) against numpy's complex arithmetic"""
    om = np.linspace(-4, 4, 801)
    E, eta, ap, bp, an, bn = -0.75, 0.05, 0.3, 0.6, -0.2, 0.4
    a = oracle.dos_diag(om, eta, E, [ap], [bp], [an], [bn], compat=False)
    g = bp / (om + E + 1j * eta - ap) + bn / (om - E + 1j * eta + an)
    assert np.max(np.abs(a - (-g.imag / np.pi))) <= 1e-14 * np.abs(g.imag / np.pi).max()
    # a Lorentzian of weight bp + bn: its integral over a wide grid
    wide = np.linspace(-400, 400, 800001)
    aw = oracle.dos_diag(wide, eta, E, [ap], [bp], [an], [bn], compat=False)
    assert abs(np.trapezoid(aw, wide) - (bp + bn)) <= 1e-3


def _ref_file_records(z):
    recs = []
    for i in range(int(z["nrec"])):
        nsp, nsn, iev, iorb, even = (int(v) for v in z[f"rec{i}_hdr"])
        recs.append(dict(iev=iev, iorb=iorb, E=z[f"rec{i}_vp"][0], p=z[f"rec{i}_vp"][1], even=bool(even), nsp=nsp,
                         nsn=nsn, ap=z[f"rec{i}_ap"], bp=z[f"rec{i}_bp"], an=z[f"rec{i}_an"], bn=z[f"rec{i}_bn"]))
    return recs


def test_dos_over_records_matches_reference_build():
    """dos (dos.f90:1-63) of the reference build reading an apbpanbn.dat of six records (one state
    with N_up = N_dn, one with both spins, two orbitals) against the oracle's record loop, bit for
    bit; the result is the last record's row per orbital, not an earlier one's"""
    z = np.load(os.path.join(GOLD, "ref_files.npz"))
    recs = _ref_file_records(z)
    om, small = z["dos_omega"], float(z["dos_small"])
    aff = oracle.dos(recs, 2, om, small)
    assert np.array_equal(aff, z["dos_aff"])
    assert not np.allclose(oracle.dos(recs[:4], 2, om, small), z["dos_aff"])


def test_dos_records_overwrite_and_weighted():
    """dos.f90:17-59: every record overwrites its orbital's row; the weighted variant sums
    p * factor * A (factor 1/2 when N_up != N_dn)"""
    om = (np.arange(1, 101) - 50) * 0.01
    rng = np.random.default_rng(6)
    recs = []
    for iev in (1, 2):
        for iorb in (1, 2):
            ap, bp, an, bn = (rng.uniform(0.1, 1, 6) for _ in range(4))
            recs.append(dict(iev=iev, iorb=iorb, E=-1.0 - iev, p=0.5 * iev, even=iev == 1, nsp=3, nsn=5,
                             ap=ap, bp=bp, an=an, bn=bn))
    last = oracle.dos(recs, 2, om, 0.05)
    for o in (0, 1):
        r = recs[2 + o]
        want = oracle.dos_diag(om, 0.05, r["E"], r["ap"][:4], r["bp"][:4], r["an"], r["bn"])
        assert np.array_equal(last[o], want)
    w = oracle.dos(recs, 2, om, 0.05, weighted=True)
    for o in (0, 1):
        acc = np.zeros(len(om))
        for r in recs[o::2]:
            acc = acc + (r["p"] * (1.0 if r["even"] else 0.5)) * oracle.dos_diag(
                om, 0.05, r["E"], r["ap"][:4], r["bp"][:4], r["an"], r["bn"])
        assert np.array_equal(w[o], acc)
