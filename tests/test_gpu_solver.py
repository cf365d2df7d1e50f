"""GPU parity of lanczos_iteration, ed_solve, ed_calc_green_ftn and the continued fraction
against the oracle, dense ED and the committed golden energies."""
import json
import os

import numpy as np
import pytest

import oracle
from oracle import dense_ed
import moed
from moed import EDSolver, MOED_F_EXACT_CMPLX
from moed.model import load_configs, omega_grid, start_vector

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")


def omodel(c):
    return oracle.Model(c.norb, c.nbath, c.U, c.J, c.ek, c.vk, c.tij)


def test_lanczos_matches_reference_recurrence(configs):
    c = configs["ns8_plain"]
    z = np.load(os.path.join(GOLD, "ref_ns8.npz"))
    nu, nd = c.sectors[0]
    with EDSolver.from_config(c) as s:
        x = start_vector(int(z["x_seed"]), s.dim(nu, nd))
        nsc, a, b = s.lanczos_iteration(nu, nd, x, c.nstep)
    assert nsc == int(z["lz_nstep_calc"])
    # the recurrence without reorthogonalisation amplifies rounding once Ritz values converge:
    # the early coefficients agree to rounding, and the continued fractions agree to 1e-8
    assert np.allclose(a[:25], z["lz_a"][:25], rtol=1e-10, atol=1e-12)
    assert np.allclose(b[1:25], z["lz_b"][1:25], rtol=1e-10)
    assert b[0] == 0.0
    om = omega_grid(c.nw, c.beta)
    bb = b.copy()
    bb[0] = 1.0
    br = z["lz_b"].copy()
    br[0] = 1.0
    # z = E + i w with E below the spectrum (E0 = -17.97), where the fraction is well conditioned
    g = oracle.green_cf(+1, -30.0, nsc, a, bb, om)
    gr = oracle.green_cf(+1, -30.0, int(z["lz_nstep_calc"]), z["lz_a"], br, om)
    assert np.max(np.abs(g - gr)) <= 1e-8 * np.max(np.abs(gr))


def test_lanczos_breakdown_in_small_krylov_space():
    rng = np.random.default_rng(12)
    ek, vk = rng.uniform(-1, 1, 4), rng.uniform(0.3, 0.9, (2, 2))
    m = oracle.Model(2, 2, 2.0, 0.3, ek, vk)
    b = oracle.Basis(4, 1, 1)
    H = np.array([oracle.multiply_H(m, b, e) for e in np.eye(b.ntot)]).T
    w, V = np.linalg.eigh(H)
    v = V[:, [0, 5, 9]] @ np.array([0.6, -0.3, 0.5])
    with EDSolver(2, 2, 2.0, 0.3, ek, vk) as s:
        nsc, a, bb = s.lanczos_iteration(1, 1, v, 40)
    nsco, ao, bo = oracle.lanczos(m, b, v, 40)
    assert nsc == nsco == 3
    assert np.all(a[4:] == 0) and np.all(bb[4:] == 0)
    assert np.allclose(a[:3], ao[:3], rtol=1e-12)


def test_solve_ground_state_8site_against_dense(configs):
    c = configs["ns8_plain"]
    e0 = json.load(open(os.path.join(GOLD, "e0.json")))["ns8_plain"]
    with EDSolver.from_config(c, nev=2) as s:
        kept = s.ed_solve()
        ev = s.sector_eigenvalues(1)
        vec = s.eigvec(1, 1)
    E0 = float.fromhex(e0["E0"])
    assert abs(ev[0] - E0) <= 1e-10 * abs(E0)
    assert abs(ev[1] - float.fromhex(e0["E1"])) <= 1e-10 * abs(E0)
    assert kept[0][0] == ev[0] and kept[0][2:] == (1, 1)
    # Ritz vector residual
    m = omodel(c)
    r = oracle.multiply_H(m, oracle.Basis(c.nsite, *c.sectors[0]), vec) - ev[0] * vec
    assert np.linalg.norm(r) < 1e-9 and abs(np.linalg.norm(vec) - 1) < 1e-12


def test_solve_ground_state_12site_against_arpack(configs):
    c = configs["ns12_plain"]
    e0 = json.load(open(os.path.join(GOLD, "e0.json"))).get("ns12_plain")
    if e0 is None:
        pytest.skip("no ARPACK golden energy for the 12-site sector")
    with EDSolver.from_config(c) as s:
        s.ed_solve()
        ev = s.sector_eigenvalues(1)
    E0 = float.fromhex(e0["E0"])
    print("E0 gpu", repr(ev[0]), "arpack", repr(E0), flush=True)
    assert abs(ev[0] - E0) <= 1e-10 * abs(E0)


def test_solve_and_green_14site(configs):
    """the 14-site workload (sector (7, 7), D = 11,778,624): E0 against SciPy's ARPACK (tol = 0)
    within 1e-10; the Green's functions on 2048 frequencies: particle and hole weights of every
    orbital add up to 1 (b+_0 + b-_0, ed_green.F90:174) and Im G(i w_n) < 0 at every frequency"""
    c = configs["ns14_roofline"]
    e0 = json.load(open(os.path.join(GOLD, "e0.json"))).get("ns14_roofline")
    if e0 is None:
        pytest.skip("no ARPACK golden energy for the 14-site sector")
    om = omega_grid(c.nw, c.beta)
    with EDSolver.from_config(c) as s:
        kept = s.ed_solve()
        ev = s.sector_eigenvalues(1)
        G = s.ed_calc_green_ftn(om)
        recs = s.green_records()
    E0 = float.fromhex(e0["E0"])
    print("E0 gpu", repr(ev[0]), "arpack", repr(E0), flush=True)
    assert abs(ev[0] - E0) <= 1e-10 * abs(E0)
    assert kept[0][0] == ev[0]
    assert len(recs) == c.norb
    for r in recs:
        assert abs(r["bp"][0] + r["bn"][0] - 1.0) < 1e-12
    assert np.all(G.imag < 0.0)


@pytest.mark.parametrize("flags", [0, moed.MOED_F_BATCHED, moed.MOED_F_FULL_REORTH,
                                   moed.MOED_F_BATCHED | moed.MOED_F_FULL_REORTH])
def test_solve_all_sectors_small_model_and_selection(flags):
    """every sector of a 5-site model (dims 1..100): eigenvalues vs dense ED, state selection
    vs the oracle's restatement of ed_solve (SORT + boltzmann_factor); sector by sector and with
    all sectors in one multi-problem run"""
    rng = np.random.default_rng(21)
    norb, nbath = 2, 3
    ns = norb + nbath
    ek, vk = rng.uniform(-1.5, 1.5, ns), rng.uniform(0.2, 0.8, (norb, nbath))
    U, J, beta, nev = 3.0, 0.4, 5.0, 4
    sectors = [(a, b) for a in range(ns + 1) for b in range(ns + 1)]
    with EDSolver(norb, nbath, U, J, ek, vk, beta=beta, nev=nev, nstep=30, sectors=sectors, flags=flags) as s:
        kept = s.ed_solve()
        evs = [s.sector_eigenvalues(i + 1) for i in range(len(sectors))]
    allv = []
    for (nu, nd), ev in zip(sectors, evs):
        sec = dense_ed.Sector(ns, nu, nd)
        w = np.linalg.eigvalsh(dense_ed.hamiltonian(sec, norb, nbath, U, J, ek, vk).toarray())
        k = min(nev, len(w))
        assert np.allclose(ev[:k], w[:k], rtol=0, atol=1e-10 * max(1, abs(w[0])))
        assert np.all(ev[k:] == np.finfo(float).max)
        allv.extend(ev)
    ref = oracle.select_states(np.array(allv), nev, beta)
    assert len(kept) == len(ref)
    for k, r in zip(kept, ref):
        assert k[2:] == r[2:] and abs(k[0] - r[0]) < 1e-10 and abs(k[1] - r[1]) < 1e-9


def test_four_orbitals_with_tij_against_dense_ed():
    """Norb = 4 (the largest) with impurity hopping t_ij and two bath sites: the lowest eigenvalues
    of four sectors against dense ED, and G(iw) of every orbital against the Lehmann sum"""
    rng = np.random.default_rng(44)
    norb, nbath = 4, 2
    ns = norb + nbath
    ek = rng.uniform(-1.5, 1.5, ns)
    vk = np.zeros((norb, nbath))
    vk[0, 0], vk[1, 1], vk[2, 0], vk[3, 1] = rng.uniform(0.2, 0.8, 4)
    t = rng.uniform(-0.3, 0.3, (norb, norb))
    t = t + t.T
    np.fill_diagonal(t, 0.0)
    U, J = 2.5, 0.3
    sectors = [(3, 3), (2, 3), (3, 2), (4, 2)]
    om = omega_grid(64, 40.0)
    with EDSolver(norb, nbath, U, J, ek, vk, beta=1000.0, nev=3, nstep=120, sectors=sectors, tij=t) as s:
        s.ed_solve()
        evs = [s.sector_eigenvalues(i + 1) for i in range(len(sectors))]
    for (nu, nd), ev in zip(sectors, evs):
        w = np.linalg.eigvalsh(dense_ed.hamiltonian(dense_ed.Sector(ns, nu, nd), norb, nbath, U, J, ek, vk, t).toarray())
        assert np.allclose(ev[:3], w[:3], rtol=0, atol=1e-10 * max(1, abs(w[0])))
    # G of the (3, 3) ground state (spin multiplets span sectors, so one sector keeps one state)
    with EDSolver(norb, nbath, U, J, ek, vk, beta=1000.0, nev=1, nstep=120, sectors=[(3, 3)], tij=t,
                  flags=MOED_F_EXACT_CMPLX) as s:
        kept = s.ed_solve()
        G = s.ed_calc_green_ftn(om)
    assert len(kept) == 1
    for iorb in range(1, norb + 1):
        _, g = dense_ed.lehmann_green(norb, nbath, U, J, ek, vk, 3, 3, iorb - 1, 0, om, tij=t)
        assert np.max(np.abs(G[iorb - 1] - g)) <= 1e-9 * np.max(np.abs(g)), iorb


def test_batched_solve_and_green_equal_sector_by_sector():
    """7-site model, all 64 sectors (dims up to 1225): the multi-problem run (persistent work-queue
    kernel) gives the same kept states as the sector-by-sector run, eigenvalues within 1e-12 and
    G(iw) within 1e-12"""
    rng = np.random.default_rng(77)
    norb, nbath = 2, 5
    ns = norb + nbath
    ek = rng.uniform(-2, 2, ns)
    vk = np.zeros((norb, nbath))
    for b in range(nbath):
        vk[b % norb, b] = rng.uniform(0.2, 0.9)
    sectors = [(a, b) for a in range(ns + 1) for b in range(ns + 1)]
    om = moed.omega_grid(256, 20.0)
    out = []
    for flags in (0, moed.MOED_F_BATCHED):
        with EDSolver(norb, nbath, 3.0, 0.5, ek, vk, beta=20.0, nev=4, nstep=40, sectors=sectors,
                      flags=flags) as s:
            kept = s.ed_solve()
            evs = np.concatenate([s.sector_eigenvalues(i + 1) for i in range(len(sectors))])
            out.append((kept, evs, s.ed_calc_green_ftn(om)))
    (k0, e0, g0), (k1, e1, g1) = out
    # the batched solve runs the small sectors in the persistent work-queue kernel (its own
    # reductions, so rounded differently); the same states are kept
    assert [k[2:] for k in k0] == [k[2:] for k in k1]
    fin = e0 < np.finfo(float).max
    assert np.array_equal(fin, e1 < np.finfo(float).max)
    assert np.max(np.abs(e0[fin] - e1[fin])) <= 1e-12 * np.max(np.abs(e0[fin]))
    assert np.allclose([k[0] for k in k0], [k[0] for k in k1], rtol=1e-12, atol=0)
    assert np.allclose([k[1] for k in k0], [k[1] for k in k1], rtol=1e-10, atol=1e-12)
    # the batched Green's-function chains reduce per 1024-element item (rounded differently from
    # the chain-by-chain sums, deterministic either way)
    assert np.max(np.abs(g0 - g1)) <= 1e-12 * np.max(np.abs(g0))
    w = np.linalg.eigvalsh(dense_ed.hamiltonian(dense_ed.Sector(ns, 3, 4), norb, nbath, 3.0, 0.5, ek, vk).toarray())
    isec = sectors.index((3, 4))
    assert np.allclose(e0[4 * isec:4 * isec + 4], w[:4], rtol=0, atol=1e-10 * abs(w[0]))


def _solve_green(s, om, nsec):
    kept = s.ed_solve()
    evs = np.concatenate([s.sector_eigenvalues(i + 1) for i in range(nsec)])
    return kept, evs, s.eigvec(1, 1), s.ed_calc_green_ftn(om)


def _same(a, b):
    assert a[0] == b[0]
    for u, v in zip(a[1:], b[1:]):
        assert np.array_equal(u, v)


def test_krylov_basis_chunks_12site(configs):
    """the eigensolver's Krylov basis allocated one and three vectors at a time (kernels address the
    vectors through a table) gives the same eigenvalues, eigenvector and G(iw), bit for bit, as the
    default chunk; so does changing the chunk size between two solves on one handle"""
    c = configs["ns12_plain"]
    om = omega_grid(c.nw, c.beta)
    with EDSolver.from_config(c) as s:
        ref = _solve_green(s, om, 1)
        s.set_eig_chunk(1)
        again = _solve_green(s, om, 1)
    with EDSolver.from_config(c) as s:
        s.set_eig_chunk(3)
        chunked = _solve_green(s, om, 1)
        again3 = _solve_green(s, om, 1)  # the kept chunks serve the second solve
    print("12-site E0", repr(ref[1][0]), flush=True)
    _same(ref, again)
    _same(ref, chunked)
    _same(ref, again3)
    with EDSolver.from_config(c) as s:
        with pytest.raises(moed.MoedError):
            s.set_eig_chunk(-1)


@pytest.mark.parametrize("flags", [0, moed.MOED_F_BATCHED])
def test_krylov_basis_chunks_all_sectors(flags):
    """7-site model, all 64 sectors, sector by sector and in one multi-problem run (sectors of
    different sizes reuse one handle's chunks): a basis allocated two vectors at a time gives the
    default chunk's results bit for bit"""
    rng = np.random.default_rng(78)
    norb, nbath = 2, 5
    ns = norb + nbath
    ek = rng.uniform(-2, 2, ns)
    vk = np.zeros((norb, nbath))
    for b in range(nbath):
        vk[b % norb, b] = rng.uniform(0.2, 0.9)
    sectors = [(a, b) for a in range(ns + 1) for b in range(ns + 1)]
    om = moed.omega_grid(256, 20.0)
    out = []
    for chunk in (0, 2):
        with EDSolver(norb, nbath, 3.0, 0.5, ek, vk, beta=20.0, nev=4, nstep=40, sectors=sectors,
                      flags=flags) as s:
            s.set_eig_chunk(chunk)
            out.append(_solve_green(s, om, len(sectors)))
    _same(out[0], out[1])


def test_green_8site_against_oracle(configs):
    c = configs["ns8_plain"]
    nu, nd = c.sectors[0]
    om = omega_grid(c.nw, c.beta)
    with EDSolver.from_config(c) as s:
        kept = s.ed_solve()
        G = s.ed_calc_green_ftn(om)
        vec = s.eigvec(1, 1)
        recs = s.green_records()
    m = omodel(c)
    b = oracle.Basis(c.nsite, nu, nd)
    Gr = np.zeros((c.norb, c.nw), complex)
    for (E, p, isec, ilev) in kept:
        for iorb in range(1, c.norb + 1):  # N_up == N_dn: spin up only, factor 1 (ed_green.F90:87-94)
            g, _, _ = oracle.green_diag(m, b, vec, E, p, 1.0, iorb, 1, c.nstep, om)
            Gr[iorb - 1] += g
    err = np.max(np.abs(G - Gr)) / np.max(np.abs(Gr))
    print("G rel err", err, flush=True)
    assert err <= 1e-8
    assert len(recs) == c.norb and all(abs(r["bp"][0] + r["bn"][0] - 1) < 1e-12 for r in recs)


def test_green_12site_against_oracle(configs):
    """the 12-site workload (sector (6, 6), D = 853,776, 150-step continued fractions on 1024
    frequencies): ground state, apply_c, the reference Lanczos recurrence and the continued
    fraction on the GPU against the oracle's green_diag (OpenMP over the host cores) on the same
    eigenvector, within 1e-8"""
    import os as _os
    c = configs["ns12_plain"]
    nu, nd = c.sectors[0]
    om = omega_grid(c.nw, c.beta)
    with EDSolver.from_config(c) as s:
        kept = s.ed_solve()
        G = s.ed_calc_green_ftn(om)
        vec = s.eigvec(1, 1)
    m = omodel(c)
    b = oracle.Basis(c.nsite, nu, nd)
    Gr = np.zeros((c.norb, c.nw), complex)
    nt = min(16, _os.cpu_count() or 1)
    for (E, p, isec, ilev) in kept:
        assert (isec, ilev) == (1, 1)
        for iorb in range(1, c.norb + 1):  # N_up == N_dn: spin up only, factor 1
            print("oracle green_diag orbital", iorb, flush=True)
            g, _, _ = oracle.green_diag(m, b, vec, E, p, 1.0, iorb, 1, c.nstep, om, nthreads=nt)
            Gr[iorb - 1] += g
    err = np.max(np.abs(G - Gr)) / np.max(np.abs(Gr))
    print("G rel err", err, flush=True)
    assert err <= 1e-8


@pytest.mark.parametrize("name, sector", [("ns14_roofline", (3, 3)), ("ns14_roofline", (3, 2)),
                                          ("ns15_fit", (2, 2)), ("ns15_fit", (3, 2))])
def test_green_odd_rows_against_oracle(configs, name, sector):
    """chains in sectors whose up-string count is odd: at 14 sites (two orbitals; the targets have
    C(14, 4) = 1001 and C(14, 2) = 91 up strings) they run on rows padded to an even stride, at 15
    sites (three orbitals, every C(15, k) odd) unpadded; G equals the oracle's green_diag on the same
    eigenvector within 1e-8, both spins when N_up != N_dn"""
    import os as _os
    c = configs[name]
    nu, nd = sector
    om = omega_grid(256, c.beta)
    with EDSolver.from_config(c, sectors=[sector], nev=1, nstep=60, nw=len(om)) as s:
        kept = s.ed_solve()
        G = s.ed_calc_green_ftn(om)
        vec = s.eigvec(1, 1)
    m = omodel(c)
    b = oracle.Basis(c.nsite, nu, nd)
    Gr = np.zeros((c.norb, len(om)), complex)
    nt = min(16, _os.cpu_count() or 1)
    for (E, p, isec, ilev) in kept:
        spins, factor = ((1,), 1.0) if nu == nd else ((1, 2), 0.5)
        for ispin in spins:
            for iorb in range(1, c.norb + 1):
                g, _, _ = oracle.green_diag(m, b, vec, E, p, factor, iorb, ispin, 60, om, nthreads=nt)
                Gr[iorb - 1] += g
    err = np.max(np.abs(G - Gr)) / np.max(np.abs(Gr))
    print("G rel err", sector, err, flush=True)
    assert err <= 1e-8


def test_green_exact_against_lehmann():
    rng = np.random.default_rng(5)
    norb, nbath, nu, nd = 2, 3, 2, 3
    ns = norb + nbath
    ek, vk = rng.uniform(-1.5, 1.5, ns), rng.uniform(0.2, 0.8, (norb, nbath))
    U, J = 3.0, 0.4
    om = omega_grid(64, 50.0)
    with EDSolver(norb, nbath, U, J, ek, vk, beta=1000.0, nev=1, nstep=200, sectors=[(nu, nd)],
                  flags=MOED_F_EXACT_CMPLX) as s:
        s.ed_solve()
        G = s.ed_calc_green_ftn(om)
    for iorb in (1, 2):
        GL = np.zeros(len(om), complex)
        for ispin in (1, 2):  # N_up != N_dn: both spins with factor 1/2
            _, g = dense_ed.lehmann_green(norb, nbath, U, J, ek, vk, nu, nd, iorb - 1, ispin - 1, om)
            GL += 0.5 * g
        assert np.max(np.abs(G[iorb - 1] - GL)) <= 1e-9 * np.max(np.abs(GL))


def test_green_cf_kernel_matches_oracle():
    rng = np.random.default_rng(4)
    n = 60
    a = rng.uniform(-2, 2, n + 1)
    b = rng.uniform(0.1, 1.0, n + 1)
    om = omega_grid(300, 40.0)
    with EDSolver(1, 1, 0.0, 0.0, [0.0, 0.0], [[0.5]]) as s:
        for sign in (1, -1):
            g = s.green_cf(sign, 0.37, n, a, b, om)
            go = oracle.green_cf(sign, 0.37, n, a, b, om)
            assert np.array_equal(g, go)  # same operations, no contraction: bit-identical


def test_green_cf_adjoint_against_finite_differences():
    rng = np.random.default_rng(8)
    n = 20
    a = rng.uniform(-2, 2, n + 1)
    b = rng.uniform(0.1, 1.0, n + 1)
    om = omega_grid(128, 30.0)
    gbar = rng.standard_normal(len(om)) + 1j * rng.standard_normal(len(om))
    with EDSolver(1, 1, 0.0, 0.0, [0.0, 0.0], [[0.5]], flags=MOED_F_EXACT_CMPLX) as s:
        for sign in (1, -1):
            ga, gb = s.green_cf_adjoint(sign, 0.2, n, a, b, om, gbar)

            def L(aa, bb):
                g = s.green_cf(sign, 0.2, n, aa, bb, om)
                return np.sum(gbar.real * g.real + gbar.imag * g.imag)

            h = 1e-6
            for j in range(0, n + 1, 3):
                ap, am = a.copy(), a.copy()
                ap[j] += h
                am[j] -= h
                fd = (L(ap, b) - L(am, b)) / (2 * h)
                assert abs(fd - ga[j]) <= 1e-6 * max(1.0, abs(fd))
                bp, bm = b.copy(), b.copy()
                bp[j] += h
                bm[j] -= h
                fd = (L(a, bp) - L(a, bm)) / (2 * h)
                assert abs(fd - gb[j]) <= 1e-6 * max(1.0, abs(fd))


def test_full_reorthogonalisation_14site_against_arpack(configs):
    """full reorthogonalisation (CGS2 against the whole basis every step, contractions and Ritz
    product on the FP64 matrix cores) at the 14-site sector: E0 within 1e-10 of SciPy's ARPACK, the
    Ritz vector's residual, and the default solver's E0 within 1e-10"""
    c = configs["ns14_roofline"]
    E0 = float.fromhex(json.load(open(os.path.join(GOLD, "e0.json")))["ns14_roofline"]["E0"])
    nu, nd = c.sectors[0]
    with EDSolver.from_config(c, flags=moed.MOED_F_FULL_REORTH) as s:
        s.ed_solve()
        ev = s.sector_eigenvalues(1)
        vec = s.eigvec(1, 1)
        y = s.multiply_H(nu, nd, vec)
    with EDSolver.from_config(c) as s:
        s.ed_solve()
        ev0 = s.sector_eigenvalues(1)
    print("E0 full", repr(ev[0]), "partial", repr(ev0[0]), "arpack", repr(E0), flush=True)
    assert abs(ev[0] - E0) <= 1e-10 * abs(E0)
    assert abs(ev[0] - ev0[0]) <= 1e-10 * abs(E0)
    assert abs(np.linalg.norm(vec) - 1.0) < 1e-12
    assert np.linalg.norm(y - ev[0] * vec) < 1e-8 * abs(E0)


@pytest.mark.parametrize("beta", [2.0, 40.0])
def test_green_chains_batched_match_sequential(beta):
    """ed_calc_green_ftn with every chain in one multi-problem run (MOED_F_BATCHED) against the
    chain-by-chain path: every sector of a 5-site model, so kept states include N_up != N_dn (both
    spins, factor 1/2), odd sector dimensions and empty or full orbitals (c+|psi> = 0 or no
    target sector); G within 1e-12 of the sequential sum, the same records (step counts, b(0),
    coefficients within 1e-10)"""
    rng = np.random.default_rng(31)
    norb, nbath = 2, 3
    ns = norb + nbath
    ek, vk = rng.uniform(-1.5, 1.5, ns), rng.uniform(0.2, 0.8, (norb, nbath))
    sectors = [(a, b) for a in range(ns + 1) for b in range(ns + 1)]
    om = omega_grid(128, beta)
    out = {}
    for flags in (0, moed.MOED_F_BATCHED):
        with EDSolver(norb, nbath, 3.0, 0.4, ek, vk, beta=beta, nev=3, nstep=20, sectors=sectors, flags=flags) as s:
            kept = s.ed_solve()
            G = s.ed_calc_green_ftn(om)
            out[flags] = (kept, G, s.green_records())
    (k0, G0, r0), (k1, G1, r1) = out[0], out[moed.MOED_F_BATCHED]
    print(f"beta {beta}: {len(k0)} kept states, {len(r0)} records", flush=True)
    assert len(k0) > 1 and [k[2:] for k in k0] == [k[2:] for k in k1]
    assert np.max(np.abs(G1 - G0)) <= 1e-12 * np.max(np.abs(G0))
    assert len(r0) == len(r1)
    for a, b in zip(r0, r1):
        for key in ("iev", "iorb", "ispin", "even", "nsp", "nsn"):
            assert a[key] == b[key], key
        assert abs(a["bp"][0] - b["bp"][0]) <= 1e-14 and abs(a["bn"][0] - b["bn"][0]) <= 1e-14
        # the recurrence without reorthogonalisation amplifies rounding once Ritz values converge
        # (small sectors: within a few steps), so the coefficients are compared over the first steps
        # and the continued fractions through G above
        for key, n in (("ap", a["nsp"]), ("bp", a["nsp"]), ("an", a["nsn"]), ("bn", a["nsn"])):
            m = min(n, 8)
            assert np.allclose(a[key][:m], b[key][:m], rtol=1e-10, atol=1e-12), key


def test_lanczos_graph_replays_are_bit_identical(configs):
    """lanczos_iteration's steps run as a HIP graph captured on the first call for a sector and
    replayed afterwards: calls on two sectors, interleaved, and a fresh handle give the same a, b
    bit for bit"""
    c = configs["ns8_plain"]
    out = {}
    with EDSolver.from_config(c) as s:
        for (nu, nd) in ((4, 4), (5, 4), (4, 4), (5, 4), (4, 4)):
            x = start_vector(99 + nu, s.dim(nu, nd))
            out.setdefault((nu, nd), []).append(s.lanczos_iteration(nu, nd, x, 60))
    with EDSolver.from_config(c) as s:
        fresh = s.lanczos_iteration(5, 4, start_vector(104, s.dim(5, 4)), 60)
    for runs in out.values():
        for r in runs[1:]:
            assert r[0] == runs[0][0] and np.array_equal(r[1], runs[0][1]) and np.array_equal(r[2], runs[0][2])
    assert fresh[0] == out[(5, 4)][0][0] and np.array_equal(fresh[1], out[(5, 4)][0][1])
    m = omodel(c)
    b54 = oracle.Basis(c.nsite, 5, 4)
    nsco, ao, bo = oracle.lanczos(m, b54, start_vector(104, b54.ntot), 60)
    assert np.allclose(out[(5, 4)][0][1][:20], ao[:20], rtol=1e-10, atol=1e-12)


def _degenerate_model():
    """two orbitals with identical baths: every level of an odd-N sector is (at least) doubly
    degenerate"""
    norb, nbath = 2, 4
    eb, vb = np.array([-0.9, 0.6]), np.array([0.45, 0.3])
    ek = np.concatenate([np.full(norb, -1.2), eb, eb])
    vk = np.zeros((norb, nbath))
    vk[0, :2] = vb
    vk[1, 2:] = vb
    return norb, nbath, 3.0, 0.4, ek, vk


@pytest.mark.parametrize("flags", [0, moed.MOED_F_BATCHED])
def test_degenerate_levels_found_with_multiplicity(flags):
    """identical orbitals, every sector of the model, nev = 4: the eigenvalues of every sector solved
    by Lanczos (D > 64) equal dense ED's lowest four with their multiplicities (the deflated
    restarts of the degeneracy check find the copies a single start vector misses)"""
    norb, nbath, U, J, ek, vk = _degenerate_model()
    ns = norb + nbath
    sectors = [(a, b) for a in range(ns + 1) for b in range(ns + 1)]
    nev = 4
    with EDSolver(norb, nbath, U, J, ek, vk, beta=50.0, nev=nev, nstep=40, sectors=sectors, flags=flags) as s:
        s.ed_solve()
        evs = [s.sector_eigenvalues(i + 1) for i in range(len(sectors))]
    ndeg = 0
    for (nu, nd), ev in zip(sectors, evs):
        sec = dense_ed.Sector(ns, nu, nd)
        w = np.linalg.eigvalsh(dense_ed.hamiltonian(sec, norb, nbath, U, J, ek, vk).toarray())
        k = min(nev, len(w))
        assert np.allclose(ev[:k], w[:k], rtol=0, atol=1e-9 * max(1, abs(w[0]))), (nu, nd, ev[:k], w[:k])
        if sec.dim > 64 and abs(w[1] - w[0]) < 1e-9:
            ndeg += 1
    assert ndeg >= 4  # the model does have degenerate ground levels in Lanczos-sized sectors


def test_degenerate_ground_pair_green_against_lehmann():
    """sector (2, 1) of the identical-orbital model (D = 90): its degenerate ground pair is kept with
    p = 1/2 each, and G(iw) of both orbitals (both spins, factor 1/2, ed_green.F90:87-94) equals the
    Lehmann sum averaged over the pair, so the two orbitals' G agree"""
    norb, nbath, U, J, ek, vk = _degenerate_model()
    ns = norb + nbath
    om = omega_grid(64, 1000.0)
    with EDSolver(norb, nbath, U, J, ek, vk, beta=1000.0, nev=4, nstep=200, sectors=[(2, 1)],
                  flags=MOED_F_EXACT_CMPLX) as s:
        kept = s.ed_solve()
        G = s.ed_calc_green_ftn(om)
    s0 = dense_ed.Sector(ns, 2, 1)
    w, V = np.linalg.eigh(dense_ed.hamiltonian(s0, norb, nbath, U, J, ek, vk).toarray())
    assert abs(w[1] - w[0]) < 1e-10 and w[2] - w[1] > 0.05
    assert len(kept) == 2 and all(abs(k[1] - 0.5) < 1e-9 for k in kept)
    z = 1j * om
    for orb in range(norb):
        g = np.zeros(len(om), complex)
        for st in range(2):
            psi = V[:, st]
            for spin in (0, 1):
                du, dd = (1, 0) if spin == 0 else (0, 1)
                sp_ = dense_ed.Sector(ns, 2 + du, 1 + dd)
                ep, vp = np.linalg.eigh(dense_ed.hamiltonian(sp_, norb, nbath, U, J, ek, vk).toarray())
                amp = vp.T @ (dense_ed.c_matrix(s0, sp_, orb, spin, True) @ psi)
                g += 0.5 * 0.5 * np.sum((amp ** 2)[:, None] / (z[None, :] - (ep - w[0])[:, None]), axis=0)
                sm = dense_ed.Sector(ns, 2 - du, 1 - dd)
                em, vm = np.linalg.eigh(dense_ed.hamiltonian(sm, norb, nbath, U, J, ek, vk).toarray())
                amp = vm.T @ (dense_ed.c_matrix(s0, sm, orb, spin, False) @ psi)
                g += 0.5 * 0.5 * np.sum((amp ** 2)[:, None] / (z[None, :] + (em - w[0])[:, None]), axis=0)
        assert np.max(np.abs(G[orb] - g)) <= 1e-8 * np.max(np.abs(g)), orb
    assert np.max(np.abs(G[0] - G[1])) <= 1e-8 * np.max(np.abs(G[0]))
