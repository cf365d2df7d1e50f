"""Fixtures for the three large workloads, written in the build container from the oracle, SciPy's
ARPACK and (for the bath fit) the reference build; the GPU tests read them back.

  ns14_green.npz      14-site sector (7, 7): ground state by eigsh (ARPACK dsaupd, which=SA, tol=0)
                      on the oracle's H v; G(i w_n) of both orbitals on the workload's 2048
                      frequencies at beta = 200 from the oracle's green_diag (ed_green.F90:136-185)
                      on that ground state, 200 Lanczos steps per recurrence
  ns10_batched.npz    10-site model, all 121 sectors: the lowest 8 eigenvalues of every sector
                      (dense eigh for D <= 5000, eigsh tol = 0 above), the oracle's state selection
                      (ed_solver.F90:65-90) at beta = 50, and G(i w_n) on 1024 frequencies summed
                      over the kept states by the oracle's green_diag, 100 steps
  ns15_gs.npz         15-site three-orbital sector (7, 8) with t_ij (D = 41,409,225): E0 and E1 by
                      eigsh, and G+ of orbital 1, spin up (the particle recurrence of green_diag,
                      150 steps, 4096 frequencies at beta = 200) on that ground state
  ns15_fit.npz        bath fit at the 15-site shape (Norb 3, Nbath 12, nx 51, 4096 frequencies,
                      weight 1/w_n): the reference build's minimization (func.f90:144-163 ->
                      frprmn.f90) from the workload's bath to a target hybridisation with six bath
                      sites per orbital (not representable by four), plus the oracle's iteration
                      and FUNC-evaluation counts for the same run

Run:  python tests/golden/make_large.py ns14_green|ns10_batched|ns15_gs|ns15_fit [--threads N]
"""
import argparse
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [ROOT, os.path.join(ROOT, "mo-dmft_amd")]

import oracle  # noqa: E402
from moed.model import config, omega_grid, start_vector  # noqa: E402


def model_of(c):
    return oracle.Model(c.norb, c.nbath, c.U, c.J, c.ek, c.vk, c.tij)


def ground_state(m, b, nthreads, k=2, seed=4242):
    import scipy.sparse.linalg as sla
    op = sla.LinearOperator((b.ntot, b.ntot), matvec=lambda v: oracle.multiply_H(m, b, v, nthreads),
                            dtype=np.float64)
    t = time.time()
    w, V = sla.eigsh(op, k=k, which="SA", tol=0, v0=start_vector(seed, b.ntot))
    o = np.argsort(w)
    w, V = w[o], V[:, o]
    r = oracle.multiply_H(m, b, V[:, 0], nthreads) - w[0] * V[:, 0]
    print(f"  eigsh D={b.ntot}: E0={w[0]!r} E1={w[1]!r} |r|={np.linalg.norm(r):.2e} ({time.time() - t:.0f}s)",
          flush=True)
    return w, V[:, 0].copy()


def gen_ns14_green(nthreads):
    c = config("ns14_roofline")
    m = model_of(c)
    nu, nd = c.sectors[0]
    b = oracle.Basis(c.nsite, nu, nd)
    w, psi = ground_state(m, b, nthreads)
    om = omega_grid(c.nw, c.beta)
    G = np.zeros((c.norb, c.nw), complex)
    recs = {}
    for iorb in range(1, c.norb + 1):  # N_up = N_dn: spin up only, factor 1 (ed_green.F90:87-94)
        t = time.time()
        g, (nsp, ap, bp), (nsn, an, bn) = oracle.green_diag(m, b, psi, w[0], 1.0, 1.0, iorb, 1, c.nstep, om,
                                                            nthreads=nthreads)
        G[iorb - 1] = g
        recs.update({f"nsp{iorb}": nsp, f"nsn{iorb}": nsn, f"bp0_{iorb}": bp[0], f"bn0_{iorb}": bn[0]})
        print(f"  green_diag orbital {iorb}: nsp {nsp} nsn {nsn} ({time.time() - t:.0f}s)", flush=True)
    np.savez_compressed(os.path.join(HERE, "ns14_green.npz"), E0=w[0], E1=w[1], omega=om, G=G, nstep=c.nstep,
                        **recs)


def gen_ns10_batched(nthreads):
    import scipy.sparse.linalg as sla
    c = config("ns10_batched")
    m = model_of(c)
    nev = c.nev
    ev = np.full((len(c.sectors), nev), np.finfo(float).max)
    vecs = {}
    t0 = time.time()
    for isec, (nu, nd) in enumerate(c.sectors):
        b = oracle.Basis(c.nsite, nu, nd)
        D = b.ntot
        if D <= 5000:
            H = np.array([oracle.multiply_H(m, b, e, 1) for e in np.eye(D)]).T
            w, V = np.linalg.eigh(0.5 * (H + H.T))
            k = min(nev, D)
            ev[isec, :k] = w[:k]
            vecs[isec] = V[:, :k]
        else:
            op = sla.LinearOperator((D, D), matvec=lambda v: oracle.multiply_H(m, b, v, nthreads), dtype=np.float64)
            w, V = sla.eigsh(op, k=nev, which="SA", tol=0, v0=start_vector(4242 + isec, D))
            o = np.argsort(w)
            ev[isec] = w[o]
            vecs[isec] = V[:, o]
    print(f"  121 sectors solved ({time.time() - t0:.0f}s)", flush=True)
    kept = oracle.select_states(ev.reshape(-1), nev, c.beta)
    om = omega_grid(c.nw, c.beta)
    G = np.zeros(2 * c.norb * c.nw)
    Gc = np.zeros((c.norb, c.nw), complex)
    for (E, p, isec1, ilev1) in kept:
        nu, nd = c.sectors[isec1 - 1]
        b = oracle.Basis(c.nsite, nu, nd)
        psi = vecs[isec1 - 1][:, ilev1 - 1]
        even = nu == nd
        factor = 1.0 if even else 0.5
        for ispin in range(1, 2 if even else 3):
            for iorb in range(1, c.norb + 1):
                g, _, _ = oracle.green_diag(m, b, psi, E, p, factor, iorb, ispin, c.nstep, om, Gr=Gc[iorb - 1],
                                            nthreads=nthreads)
                Gc[iorb - 1] = g
    print(f"  {len(kept)} kept states, G done ({time.time() - t0:.0f}s)", flush=True)
    del G
    np.savez_compressed(os.path.join(HERE, "ns10_batched.npz"), ev=ev, kept=np.array(kept, dtype=float),
                        omega=om, G=Gc, beta=c.beta, nstep=c.nstep)


def gen_ns15_gs(nthreads):
    c = config("ns15_fit")
    m = model_of(c)
    nu, nd = c.sectors[0]
    b = oracle.Basis(c.nsite, nu, nd)
    w, psi = ground_state(m, b, nthreads)
    om = omega_grid(c.nw, c.beta)
    # G+ of orbital 1, spin up: the particle half of green_diag (c+ psi into (nu + 1, nd))
    bo, v = oracle.apply_c(b, psi, 1, 1, 1)
    n2 = float(np.dot(v, v))
    t = time.time()
    nsc, a, bb = oracle.lanczos(m, bo, v, c.nstep, nthreads)
    bb = bb.copy()
    bb[0] = n2
    gp = oracle.green_cf(+1, w[0], nsc, a, bb, om)
    print(f"  G+ orbital 1 spin up: nstep_calc {nsc} ({time.time() - t:.0f}s)", flush=True)
    np.savez_compressed(os.path.join(HERE, "ns15_gs.npz"), E0=w[0], E1=w[1], omega=om, gp=gp, normsq=n2,
                        nstep=c.nstep, nsc=nsc)


def gen_ns15_fit(nthreads):
    c = config("ns15_fit")
    norb, nbath = c.norb, c.nbath
    om = omega_grid(c.nw, c.beta)
    # target: six bath sites per orbital, e ~ U[-3, 3], V ~ U[0.1, 1.0], seed 1505
    rng = np.random.default_rng(1505)
    nb6 = 6
    et = rng.uniform(-3.0, 3.0, norb * nb6)
    vt = rng.uniform(0.1, 1.0, norb * nb6)
    ekt = np.concatenate([c.ek[:norb], et])
    vkt = np.zeros((norb, norb * nb6))
    for k in range(norb * nb6):
        vkt[k // nb6, k] = vt[k]
    Dev = oracle.calc_dev(oracle.Model(norb, norb * nb6, c.U, c.J, ekt, vkt), c.ek[:norb], om)
    x0 = oracle.ev_to_x(c.ek, c.vk, c.ek[:norb].copy(), norb, nbath)
    m = model_of(c)
    dflat = np.stack([Dev.T.reshape(-1).real, Dev.T.reshape(-1).imag], 1).reshape(-1)
    t = time.time()
    raw = oracle.run_ref(5, m, i1=len(om), payload=[x0, om, dflat])
    fret_ref = float(np.frombuffer(raw, "<f8", 1)[0])
    xmin_ref = np.frombuffer(raw[8:], "<f8").copy()
    print(f"  reference minimization: fret {fret_ref:.6e} ({time.time() - t:.0f}s)", flush=True)
    t = time.time()
    rc, xo, fo, it, nf = oracle.frprmn(x0, om, Dev, norb, nbath)
    print(f"  oracle frprmn: rc {rc} fret {fo:.6e} iter {it} nfev {nf} ({time.time() - t:.0f}s); "
          f"x equal to the reference build's: {np.array_equal(xo, xmin_ref)}", flush=True)
    assert rc == 0 and np.array_equal(xo, xmin_ref) and fo == fret_ref
    np.savez_compressed(os.path.join(HERE, "ns15_fit.npz"), norb=norb, nbath=nbath, omega=om, Dev=Dev, x0=x0,
                        xmin=xmin_ref, fret=fret_ref, iter=it, nfev=nf)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("what", choices=("ns14_green", "ns10_batched", "ns15_gs", "ns15_fit"))
    ap.add_argument("--threads", type=int, default=0)
    a = ap.parse_args()
    globals()["gen_" + a.what](a.threads)
    print("done", a.what)


if __name__ == "__main__":
    main()
