"""The DMFT self-consistency loop of the reference (main.f90:59-127), driven through the C ABI.

Per iteration: ed_solve -> ed_calc_green_ftn -> ed_delta_new (+ n_from_gksum) -> minimization of
the hybridisation misfit -> new bath (x_to_ev) -> calc_dev.  The reference fits from one starting
point (the current bath); here ``nfit`` fits run in one batched launch: the current bath and
``nfit - 1`` seeded perturbations of it, and the lowest chi^2 wins (ties: the lowest index, so
``nfit = 1`` is the reference's own step).
"""
import time

import numpy as np


def ev_to_x(ek, vk, ef, norb, nbath):
    """x = [bath energies, V(orb 1, :), ..., V(orb Norb, :), ef]  (func.f90:165-191)"""
    vk = np.asarray(vk, np.float64).reshape(norb, nbath)
    return np.concatenate([np.asarray(ek, np.float64)[norb:norb + nbath], vk.reshape(-1), np.asarray(ef, np.float64)])


def x_to_ev(x, norb, nbath):
    """(ef, ek, vk) from x, with ek(1:Norb) = ef (func.f90:193-216)"""
    x = np.asarray(x, np.float64)
    ef = x[len(x) - norb:].copy()
    ek = np.concatenate([ef, x[:nbath]])
    vk = x[nbath:nbath + norb * nbath].reshape(norb, nbath).copy()
    return ef, ek, vk


def start_points(x0, nfit, seed, spread):
    """the current bath, then nfit - 1 relative perturbations of its nonzero entries
    (splitmix64 counter hash, as the start vectors), so zero couplings stay zero"""
    from .model import start_vector
    x0 = np.asarray(x0, np.float64)
    X = np.repeat(x0[None, :], nfit, axis=0)
    if nfit > 1:
        r = start_vector(seed, (nfit - 1) * len(x0)).reshape(nfit - 1, len(x0))
        X[1:] = x0 * (1.0 + spread * r)
    return X


def dmft_loop(solver, hk, mu, omega, niter, nfit=1, seed=1, spread=0.1, ftol=1.0e-7, scf_tol=1.0e-4,
              on_iter=None):
    """Run up to ``niter`` iterations; returns the per-iteration records and the final bath.
    ``on_iter(record)`` is called after each iteration (progress reporting)."""
    norb, nbath = solver.norb, solver.nbath
    ek = solver._ek.copy()
    vk = solver._vk.reshape(norb, nbath, order="F").copy()
    ef = ek[:norb].copy()
    Dev = solver.calc_dev(ek, vk, ef, omega)  # ed_green_init (ed_green.F90:57)
    Gprev = np.zeros((norb, len(omega)), np.complex128)
    out = []
    for it in range(niter):
        t0 = time.perf_counter()
        kept = solver.ed_solve()
        t1 = time.perf_counter()
        G = solver.ed_calc_green_ftn(omega)
        t2 = time.perf_counter()
        _, Dev, nocc = solver.delta_new(G, Dev, hk, mu)
        t3 = time.perf_counter()
        X = start_points(ev_to_x(ek, vk, ef, norb, nbath), nfit, seed + it, spread)
        X, fret, iters, status = solver.minimization(X, omega, Dev, ftol)
        best = int(np.argmin(fret))
        t4 = time.perf_counter()
        ef, ek, vk = x_to_ev(X[best], norb, nbath)
        tol = float(np.mean(np.sum(np.abs(Gprev - G), axis=1) / len(omega)))  # main.f90:94-99
        Gprev = G
        solver.set_bath(ek, vk)
        Dev = solver.calc_dev(ek, vk, ef, omega)
        t5 = time.perf_counter()
        out.append(dict(iteration=it, E0=kept[0][0], nocc=nocc.tolist(), fret=float(fret[best]), best=best,
                        fit_iters=int(iters[best]), scf_diff=tol, t_solve=t1 - t0, t_green=t2 - t1,
                        t_delta=t3 - t2, t_fit=t4 - t3, t_total=t5 - t0))
        if on_iter is not None:
            on_iter(out[-1])
        if tol < scf_tol:
            break
    return out, (ef, ek, vk)
