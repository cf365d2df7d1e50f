"""Benchmark: FP64 sector H v on one MI355X at the 14-site half-filled sector (N_up = N_dn = 7,
D = 3432^2 = 11,778,624), the reference's multiply_H (ed_hamiltonian.F90:88-191).

One step = one H v over the whole sector, input and output resident in HBM.  Prints ONE JSON
line with the algorithmic throughput 16*D bytes per H v (read x once, write y once) against the
8 TB/s HBM peak, and the reference's own multiply_H (the reference build, oracle/_ref) timed on
the host cores over MPI ranks as the CPU baseline.

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--no-cpu] [--config ns14_roofline]

The same line also carries the other workloads' wall times, each timed in the same run: the
10-site batched-versus-plain solve (with one DMFT iteration's lattice step and bath fits), the
14-site solves, and the 8-iteration DMFT loop at the 15-site fit workload in both fit orders
(about four minutes; --no-dmft skips the loops, --no-solves the loops and the 14-site solves).

With --dmft it instead runs the DMFT self-consistency loop of main.f90 on the 15-site, three-orbital
sector (7, 8) (D = 41,409,225), 4096 frequencies at beta = 200, the tight-binding lattice step over
Nq = 10^4 k-points and --nfit bath fits per iteration in one launch, and prints one JSON line with
the wall time of each stage of each iteration:

    python bench.py --dmft [--dmft-iters 8] [--nfit 256] [--fit-order tree|reference]

With --solve CONFIG it times one impurity solve (ed_solve, then ed_calc_green_ftn) of a named
workload, all of its sectors in one multi-problem run with --batched:

    python bench.py --solve ns14_roofline | ns10_batched [--batched] [--full-reorth]
"""
import argparse
import json
import math
import os
import struct
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path[:0] = [ROOT, os.path.join(ROOT, "mo-dmft_amd")]

HBM_PEAK_GBS = 8000.0
METRIC = "H·v GB/s & % HBM peak at Ns=14 FP64; batched vs plain full-solve wall s"


def cpu_reference_hv(cfg, nu, nd, calls=2):
    """the reference's multiply_H (reference build) over MPI ranks on this host's cores"""
    import numpy as np
    drv = os.path.join(ROOT, "oracle", "_ref", "ref_driver")
    mpiexec = "/opt/conda/bin/mpiexec"
    if not (os.path.exists(drv) and os.path.exists(mpiexec)):
        return None
    ncores = min(16, os.cpu_count() or 1)
    with tempfile.TemporaryDirectory() as td:
        fin, fout = os.path.join(td, "in.bin"), os.path.join(td, "out.bin")
        with open(fin, "wb") as f:
            f.write(struct.pack("<8i", 8, cfg.norb, cfg.nbath, nu, nd, calls, 0, 0))
            f.write(struct.pack("<2d", cfg.U, cfg.J))
            f.write(np.asarray(cfg.ek, "<f8").tobytes())
            f.write(np.asarray(cfg.vk, "<f8").reshape(-1, order="F").tobytes())
        r = subprocess.run([mpiexec, "-n", str(ncores), drv, fin, fout], capture_output=True, text=True,
                           timeout=600, cwd=td)
        if r.returncode != 0 or not os.path.exists(fout):
            return None
        raw = open(fout, "rb").read()
        sec = struct.unpack("<d", raw[:8])[0]
        ranks = struct.unpack("<i", raw[8:12])[0]
    D = cfg.dim(nu, nd)
    return {"value": 16.0 * D / sec / 1e9, "unit": "GB/s", "cores": ranks, "kind": "reference",
            "sample": f"{calls} calls of the reference's multiply_H (ed_hamiltonian.F90:88, amdflang -O2 build of "
                      f"the reference sources) on the same sector, {ranks} MPI ranks, one per host core; "
                      f"{sec:.3f} s per H v", "seconds_per_hv": sec}


def _ref_run(task, cfg, nu, nd, i1, payload=(), ranks=None, timeout=300):
    """run one task of the reference build over MPI ranks on this host; raw output bytes or None"""
    import numpy as np
    drv = os.path.join(ROOT, "oracle", "_ref", "ref_driver")
    mpiexec = "/opt/conda/bin/mpiexec"
    if not (os.path.exists(drv) and os.path.exists(mpiexec)):
        return None
    n = ranks or min(16, os.cpu_count() or 1)
    with tempfile.TemporaryDirectory() as td:
        fin, fout = os.path.join(td, "in.bin"), os.path.join(td, "out.bin")
        with open(fin, "wb") as f:
            f.write(struct.pack("<8i", task, cfg.norb, cfg.nbath, nu, nd, i1, 0, 0))
            f.write(struct.pack("<2d", cfg.U, cfg.J))
            f.write(np.asarray(cfg.ek, "<f8").tobytes())
            f.write(np.asarray(cfg.vk, "<f8").reshape(-1, order="F").tobytes())
            for p in payload:
                f.write(np.ascontiguousarray(p).tobytes())
        r = subprocess.run([mpiexec, "-n", str(n), drv, fin, fout], capture_output=True, text=True,
                           timeout=timeout, cwd=td)
        if r.returncode != 0 or not os.path.exists(fout):
            return None
        return open(fout, "rb").read()


def cpu_extra_baselines(gpu):
    """CPU baselines beyond H v, each beside the GPU time of the same work:
    lanczos_iteration steps at the 14-site sector and one minimization at the 15-site fit shape
    (reference build, MPI ranks on the host cores), and ground-state solves with the oracle's C
    restatement of multiply_H under SciPy's ARPACK (the reference's diag calls PARPACK, absent here)"""
    import numpy as np
    out = {}
    c3 = config_of("ns14_roofline")
    nu, nd = c3.sectors[0]
    raw = _ref_run(14, c3, nu, nd, 4)
    if raw:
        sec = struct.unpack("<d", raw[:8])[0]
        out["lanczos_step_ns14"] = {
            "kind": "reference", "cores": struct.unpack("<i", raw[8:12])[0], "cpu_s_per_step": round(sec, 4),
            "gpu_s_per_step": gpu.get("lanczos_step_ns14"),
            "sample": "lanczos_iteration (ed_lanczos.F90:14-68) with Nstep = 4 on the 14-site sector (5 products), "
                      "rows split over the MPI ranks"}
    z = np.load(os.path.join(ROOT, "tests", "golden", "ns15_fit.npz"))
    c5 = config_of("ns15_fit")
    dflat = np.stack([z["Dev"].T.reshape(-1).real, z["Dev"].T.reshape(-1).imag], 1).reshape(-1)
    raw = _ref_run(15, c5, 0, 0, len(z["omega"]), [z["x0"], z["omega"], dflat])
    if raw:
        sec = struct.unpack("<d", raw[:8])[0]
        out["minimization_ns15_shape"] = {
            "kind": "reference", "cores": struct.unpack("<i", raw[8:12])[0], "cpu_s_one_fit": round(sec, 3),
            "cpu_s_256_fits_extrapolated": round(256 * sec, 1), "gpu_s_one_fit": gpu.get("fit_one_ns15"),
            "gpu_s_256_fits": gpu.get("fit_256_ns15"),
            "sample": "one minimization (func.f90:144 -> frprmn) at nx 51, 4096 frequencies (tests/golden/"
                      "ns15_fit.npz), frequencies split over the MPI ranks as ed_green_init splits them; the "
                      "256-fit figure is that time x 256 (extrapolated, not run)"}
    import oracle
    import scipy.sparse.linalg as sla
    from moed.model import start_vector
    nt = min(16, os.cpu_count() or 1)
    for name in ("ns8_plain", "ns12_plain"):
        c = config_of(name)
        su, sd = c.sectors[0]
        m = oracle.Model(c.norb, c.nbath, c.U, c.J, c.ek, c.vk, c.tij)
        b = oracle.Basis(c.nsite, su, sd)
        op = sla.LinearOperator((b.ntot, b.ntot), matvec=lambda v: oracle.multiply_H(m, b, v, nt), dtype=np.float64)
        t = time.perf_counter()
        w = sla.eigsh(op, k=1, which="SA", tol=1e-12, v0=start_vector(4242, b.ntot), return_eigenvectors=False)
        sec = time.perf_counter() - t
        out[f"ground_state_{name}"] = {
            "kind": "port", "cores": nt, "cpu_s": round(sec, 3), "gpu_s": gpu.get(f"ground_state_{name}"),
            "E0_cpu": float(w[0]), "E0_gpu": gpu.get(f"E0_{name}"),
            "sample": f"sector {su, sd} (D = {b.ntot}): SciPy eigsh (ARPACK dsaupd, which=SA, tol 1e-12) on "
                      f"the oracle's C restatement of multiply_H with {nt} OpenMP threads"}
    return out


def gpu_extra_timings():
    """GPU times of the work the extra CPU baselines time"""
    import numpy as np
    import moed
    from moed.model import config, start_vector
    out = {}
    c3 = config("ns14_roofline")
    nu, nd = c3.sectors[0]
    with moed.EDSolver.from_config(c3) as s:
        x = start_vector(5, s.dim(nu, nd))
        s.lanczos_iteration(nu, nd, x, 4)  # captures the step graph
        t = time.perf_counter()
        s.lanczos_iteration(nu, nd, x, 200)
        out["lanczos_step_ns14"] = round((time.perf_counter() - t) / 201, 6)
    z = np.load(os.path.join(ROOT, "tests", "golden", "ns15_fit.npz"))
    norb, nbath = int(z["norb"]), int(z["nbath"])
    with moed.EDSolver(norb, nbath, 0.0, 0.0, np.zeros(norb + nbath), np.zeros((norb, nbath))) as s:
        t = time.perf_counter()
        s.minimization(z["x0"], z["omega"], z["Dev"])
        out["fit_one_ns15"] = round(time.perf_counter() - t, 3)
        rng = np.random.default_rng(11)
        X = z["x0"][None, :] * (1 + 0.1 * rng.uniform(-1, 1, (256, len(z["x0"]))))
        t = time.perf_counter()
        s.minimization(X, z["omega"], z["Dev"])
        out["fit_256_ns15"] = round(time.perf_counter() - t, 3)
    for name in ("ns8_plain", "ns12_plain"):
        c = config(name)
        with moed.EDSolver.from_config(c) as s:
            s.ed_solve()  # tables
        with moed.EDSolver.from_config(c) as s:
            t = time.perf_counter()
            kept = s.ed_solve()
            out[f"ground_state_{name}"] = round(time.perf_counter() - t, 4)
            out[f"E0_{name}"] = kept[0][0]
    return out


def config_solves():
    """wall times of the 14-site sector's solve and Green's functions, timed here so that the line
    carries them: partial reorthogonalisation, and the solve with full reorthogonalisation on the
    FP64 matrix cores"""
    import moed
    from moed.model import config, omega_grid
    out = {}
    c3 = config("ns14_roofline")
    om = omega_grid(c3.nw, c3.beta)
    for name, flags in (("ns14_partial_reorth", 0), ("ns14_full_reorth_mfma", moed.MOED_F_FULL_REORTH)):
        with moed.EDSolver.from_config(c3, flags=flags) as s:
            t0 = time.perf_counter()
            kept = s.ed_solve()
            t1 = time.perf_counter()
            s.ed_calc_green_ftn(om)
            t2 = time.perf_counter()
        out[name] = {"t_solve_s": round(t1 - t0, 4), "t_green_s": round(t2 - t1, 4), "E0": kept[0][0]}
    return out


def dmft_loops(niter=8, nfit=256):
    """the DMFT loop (main.f90:59-127) at the 15-site fit workload, niter iterations of solve, G,
    lattice step and nfit bath fits in one launch, once with each fit order: reference order (every
    fit's CG trajectory bit-identical to the reference's minimization for its inputs) and tree
    order (faster sums, rounded differently); progress per iteration on stderr"""
    import numpy as np
    import moed
    from moed import dmft
    from moed.model import config, omega_grid
    c5 = config("ns15_fit")
    om5 = omega_grid(c5.nw, c5.beta)
    hk = moed.band_tb(c5.norb, 10000)
    keys = ("t_solve", "t_green", "t_delta", "t_fit", "t_total")
    out, baths = {}, {}
    for order, flags in (("reference", 0), ("tree", moed.MOED_F_FIT_TREE)):
        def report(r):
            print(f"bench.py: DMFT loop ({order} order) iteration {r['iteration']}: {r['t_total']:.2f} s "
                  f"(fit {r['t_fit']:.2f}), chi2 {r['fret']:.4e}", file=sys.stderr, flush=True)

        with moed.EDSolver.from_config(c5, flags=flags) as s:
            recs, bath = dmft.dmft_loop(s, hk, c5.mu, om5, niter, nfit=nfit, scf_tol=0.0, on_iter=report)
        baths[order] = bath
        out[order] = {"wall_s": round(sum(r["t_total"] for r in recs), 3),
                      "per_stage_s": {k: round(sum(r[k] for r in recs), 3) for k in keys},
                      "fit_s_per_iteration": [round(r["t_fit"], 3) for r in recs],
                      "final_chi2": float(f"{recs[-1]['fret']:.6g}"),
                      "final_scf_diff": float(f"{recs[-1]['scf_diff']:.4g}")}
    # The fitted hybridisation depends on V(o,k)^2 and not on the order of an orbital's bath sites,
    # so the two orders' final baths are compared in a canonical form (per orbital its sites sorted
    # by energy, |V|, then ef) and through Delta_o(i w_n) itself.
    def canonical(b):
        ef, ek, vk = b
        parts = []
        for o in range(c5.norb):
            ks = np.nonzero(vk[o])[0]
            parts.append(np.array(sorted(zip(ek[c5.norb + ks], np.abs(vk[o, ks])))).ravel())
        return np.concatenate(parts + [ef])

    def delta(b):
        ef, ek, vk = b
        z = 1j * om5[None, :, None] - ek[c5.norb:][None, None, :]
        return np.sum(vk[:, None, :] ** 2 / z, axis=2)

    dt, dr = delta(baths["tree"]), delta(baths["reference"])
    return {"workload": f"DMFT loop, {c5.nsite}-site sector (7, 8), D = 41409225, t_ij on, {c5.nw} frequencies, "
                        f"Nq = 10^4, {nfit} bath fits per iteration in one launch, {niter} iterations",
            "iterations": niter, "nfit": nfit, **out,
            "final_bath_tree_vs_reference": {
                "canonical_max_abs_diff": float(f"{np.max(np.abs(canonical(baths['tree']) - canonical(baths['reference']))):.3g}"),
                "delta_max_rel_diff": float(f"{np.max(np.abs(dt - dr)) / np.max(np.abs(dr)):.3g}")}}


def config_of(name):
    from moed.model import config
    return config(name)


def cpu_port_hv(cfg, nu, nd):
    """fallback: the oracle's C restatement with OpenMP over rows"""
    import numpy as np
    import oracle
    m = oracle.Model(cfg.norb, cfg.nbath, cfg.U, cfg.J, cfg.ek, cfg.vk, cfg.tij)
    b = oracle.Basis(cfg.nsite, nu, nd)
    x = np.random.default_rng(0).standard_normal(b.ntot)
    nt = min(16, os.cpu_count() or 1)
    t = time.time()
    oracle.multiply_H(m, b, x, nt)
    sec = time.time() - t
    return {"value": 16.0 * b.ntot / sec / 1e9, "unit": "GB/s", "cores": nt, "kind": "port",
            "sample": f"1 call of the C restatement of multiply_H, {nt} OpenMP threads; {sec:.3f} s per H v",
            "seconds_per_hv": sec}


def batched_vs_plain(cfg_name="ns10_batched", reps=1, nfit=256, nq=10000):
    """full solve (every sector, nev eigenpairs each, state selection) plus the Green's functions,
    sector by sector and with all sectors in one multi-problem run; wall seconds of each.  Each leg
    then also runs the rest of one DMFT iteration on its own G -- the lattice step (ed_delta_new)
    and nfit bath fits in one launch (tree order) -- and the totals with that step are reported
    beside the solve + G figures."""
    import numpy as np
    import moed
    from moed import dmft
    from moed.model import config, omega_grid
    c = config(cfg_name)
    om = omega_grid(c.nw, c.beta)
    hk = moed.band_tb(c.norb, nq)
    out = {}
    for name, flags in (("plain", 0), ("batched", moed.MOED_F_BATCHED)):
        best = best_all = None
        for _ in range(reps + 1):  # the first run of each form also builds the string tables
            s = moed.EDSolver.from_config(c, flags=flags | moed.MOED_F_FIT_TREE)
            ek = s._ek.copy()
            vk = s._vk.reshape(c.norb, c.nbath, order="F").copy()
            ef = ek[:c.norb].copy()
            Dev = s.calc_dev(ek, vk, ef, om)
            t0 = time.perf_counter()
            kept = s.ed_solve()
            G = s.ed_calc_green_ftn(om)
            t1 = time.perf_counter()
            _, Dev, _ = s.delta_new(G, Dev, hk, c.mu)
            X = dmft.start_points(dmft.ev_to_x(ek, vk, ef, c.norb, c.nbath), nfit, 1, 0.1)
            X, fret, _, _ = s.minimization(X, om, Dev, 1.0e-7)
            t2 = time.perf_counter()
            s.close()
            best = t1 - t0 if best is None else min(best, t1 - t0)
            best_all = t2 - t0 if best_all is None else min(best_all, t2 - t0)
        out[name] = best
        out[name + "_all"] = best_all
        out[name + "_E0"] = kept[0][0]
        out[name + "_chi2"] = float(np.min(fret))
    return {"workload": f"{c.nsite}-site model, all {len(c.sectors)} sectors, {c.nev} eigenpairs each, "
                        f"G(iw) on {c.nw} frequencies at beta={c.beta:g}",
            "plain_s": round(out["plain"], 4), "batched_s": round(out["batched"], 4),
            "speedup": round(out["plain"] / out["batched"], 2),
            "E0_equal": bool(out["plain_E0"] == out["batched_E0"]),
            "with_fit": {"workload": f"the same, then the lattice step over Nq={nq} and {nfit} bath fits "
                                    "(one launch, tree order) on each leg's own G",
                         "plain_s": round(out["plain_all"], 4), "batched_s": round(out["batched_all"], 4),
                         "speedup": round(out["plain_all"] / out["batched_all"], 2),
                         "chi2_plain": float(f"{out['plain_chi2']:.6g}"),
                         "chi2_batched": float(f"{out['batched_chi2']:.6g}")}}


def hbm_traffic(path):
    """HBM bytes per H v from the committed counter summary (line 'hbm_bytes_per_launch: N')"""
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("hbm_bytes_per_launch:"):
            return int(line.split(":")[1])
    return None


def dmft_bench(args):
    """the DMFT loop (main.f90:59-127) at the 15-site fit workload; progress on stderr"""
    import numpy as np
    import moed
    from moed import dmft
    from moed.model import config, omega_grid

    cfg = config("ns15_fit")
    flags = moed.MOED_F_FIT_TREE if args.fit_order == "tree" else 0
    nq = 10000
    om = omega_grid(cfg.nw, cfg.beta)
    hk = moed.band_tb(cfg.norb, nq)
    t0 = time.perf_counter()
    with moed.EDSolver.from_config(cfg, flags=flags) as s:
        nu, nd = cfg.sectors[0]

        def report(r):
            print(f"bench.py --dmft: iteration {r['iteration']}: {r['t_total']:.2f} s (solve {r['t_solve']:.2f}, "
                  f"green {r['t_green']:.2f}, lattice {r['t_delta']:.3f}, fit {r['t_fit']:.2f}), "
                  f"chi2 {r['fret']:.4e}, scf diff {r['scf_diff']:.3e}", file=sys.stderr, flush=True)

        recs, _ = dmft.dmft_loop(s, hk, cfg.mu, om, args.dmft_iters, nfit=args.nfit, scf_tol=0.0, on_iter=report)
        D = s.dim(nu, nd)
    wall = time.perf_counter() - t0
    keys = ("t_solve", "t_green", "t_delta", "t_fit", "t_total")
    out = {
        "metric": "DMFT loop wall s, 15-site 3-orbital sector, bath fits per iteration in one launch",
        "value": round(sum(r["t_total"] for r in recs), 3), "unit": "s", "n_gpus": 1,
        "steps": len(recs), "higher_is_better": False, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"DMFT loop, {cfg.nsite}-site sector N_up={nu} N_dn={nd}, D={D}, t_ij on",
                   "norb": cfg.norb, "nbath": cfg.nbath, "U": cfg.U, "J": cfg.J, "nw": cfg.nw, "beta": cfg.beta,
                   "nq": nq, "mu": cfg.mu, "nfit": args.nfit, "fit_order": args.fit_order,
                   "nstep": cfg.nstep},
        "per_stage_s": {k: round(sum(r[k] for r in recs), 3) for k in keys},
        "iterations": [{k: (float(f"{v:.6g}") if isinstance(v, float) else v) for k, v in r.items()} for r in recs],
        "wall_s_with_setup": round(wall, 3),
    }
    print(json.dumps(out))


def solve_bench(args):
    """one ed_solve + ed_calc_green_ftn of a named workload"""
    import moed
    from moed.model import config, omega_grid

    cfg = config(args.solve)
    flags = (moed.MOED_F_BATCHED if args.batched else 0) | (moed.MOED_F_FULL_REORTH if args.full_reorth else 0)
    om = omega_grid(cfg.nw, cfg.beta)
    with moed.EDSolver.from_config(cfg, flags=flags) as s:
        if args.persistent is not None:
            s.set_persistent(args.persistent)
        t0 = time.perf_counter()
        kept = s.ed_solve()
        t1 = time.perf_counter()
        s.ed_calc_green_ftn(om)
        t2 = time.perf_counter()
    out = {"metric": "impurity solve wall s (ed_solve + ed_calc_green_ftn)", "value": round(t2 - t0, 4), "unit": "s",
           "n_gpus": 1, "higher_is_better": False, "dtype": "f64", "data": "synthetic",
           "config": {"workload": f"{cfg.name}: {cfg.nsite} sites, {len(cfg.sectors)} sector(s), nev {cfg.nev}, "
                                  f"nstep {cfg.nstep}, {cfg.nw} frequencies", "batched": bool(args.batched),
                      "reorthogonalisation": "full (CGS2, FP64 MFMA)" if args.full_reorth else "partial"},
           "t_solve_s": round(t1 - t0, 4), "t_green_s": round(t2 - t1, 4), "E0": kept[0][0], "states": len(kept)}
    print(json.dumps(out))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--config", default="ns14_roofline")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-solves", action="store_true",
                    help="skip the 14-site solves and the 15-site DMFT loops")
    ap.add_argument("--no-dmft", action="store_true", help="skip the 15-site DMFT loops (both fit orders)")
    ap.add_argument("--threads", type=int, default=0)
    ap.add_argument("--strip", type=int, default=0)
    ap.add_argument("--copy-check", type=int, default=0,
                    help="after the timed region, N plain device copies of the same vector (a kernel of known "
                         "traffic, 16*D bytes, that calibrates the FETCH_SIZE / WRITE_SIZE counters)")
    ap.add_argument("--dmft", action="store_true", help="time the DMFT loop at the 15-site fit workload instead")
    ap.add_argument("--dmft-iters", type=int, default=8)
    ap.add_argument("--nfit", type=int, default=256)
    ap.add_argument("--fit-order", choices=("tree", "reference"), default="tree")
    ap.add_argument("--solve", default=None, help="time one impurity solve of this workload instead")
    ap.add_argument("--batched", action="store_true", help="with --solve: all sectors in one multi-problem run")
    ap.add_argument("--persistent", type=int, default=None,
                    help="with --solve --batched: largest sector of the persistent launch (0: none)")
    ap.add_argument("--full-reorth", action="store_true",
                    help="with --solve: full reorthogonalisation (CGS2 on the FP64 matrix cores) in the eigensolver")
    args = ap.parse_args()
    if args.dmft:
        return dmft_bench(args)
    if args.solve:
        return solve_bench(args)
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: the H v benchmark runs one sector on one GPU; a run on more than one GPU is not "
              "supported", file=sys.stderr)
        sys.exit(2)

    import torch
    import moed
    from moed.model import config

    cfg = config(args.config)
    nu, nd = cfg.sectors[0]
    solver = moed.EDSolver.from_config(cfg)
    if args.threads or args.strip:
        solver.set_hv_shape(args.threads, args.strip)
    D = solver.dim(nu, nd)
    dev = torch.device("cuda:0")
    x = torch.empty(D, dtype=torch.float64, device=dev)
    y = torch.empty(D, dtype=torch.float64, device=dev)
    # a dedicated (default-priority) stream: the kernels and the timing events share it
    ts = torch.cuda.Stream(device=dev)
    stream = ts.cuda_stream
    rc = moed.lib().moed_fill_start_device(solver.h, 12345, D, x.data_ptr(), stream)
    assert rc == 0
    torch.cuda.synchronize()
    t0 = torch.cuda.Event(enable_timing=True)
    t1 = torch.cuda.Event(enable_timing=True)
    t2 = torch.cuda.Event(enable_timing=True)
    t3 = torch.cuda.Event(enable_timing=True)

    def hv_step():
        solver.multiply_H_device(nu, nd, x.data_ptr(), y.data_ptr(), stream)

    for _ in range(args.warmup):
        hv_step()
    # the K steps captured once in a HIP graph (no host launch gaps between the kernels)
    g = None
    try:
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g, stream=ts):
            for _ in range(args.steps):
                hv_step()
    except Exception as e:  # capture unsupported: report the eager figure
        print(f"bench.py: graph capture failed ({e}); timing the eager loop", file=sys.stderr)
        g = None
    torch.cuda.synchronize()
    # Clock ramp: kernel traces of consecutive K-step loops (eager and graph alike) run 125-135 us
    # per H v in the first few milliseconds after the GPU was idle and settle near 110 us after
    # ~40 ms of sustained load.  So untimed H v steps keep the GPU loaded for at least WARM_MS
    # before the timed steps, which are queued directly behind them with no host synchronisation
    # in between; events bracket exactly the K timed steps of each form.
    WARM_MS = 150.0
    n_warm = max(1, math.ceil(WARM_MS / (0.12 * args.steps)))
    wall0 = time.perf_counter()
    with torch.cuda.stream(ts):
        for _ in range(n_warm):
            if g is not None:
                g.replay()
            else:
                for _ in range(args.steps):
                    hv_step()
        t0.record(ts)
        for _ in range(args.steps):  # eager: one library call per step
            hv_step()
        t1.record(ts)
        if g is not None:
            t2.record(ts)
            g.replay()
            t3.record(ts)
    torch.cuda.synchronize()
    wall = time.perf_counter() - wall0  # warm-up load plus both timed forms
    ms_eager = t0.elapsed_time(t1) / args.steps
    ms_graph = t2.elapsed_time(t3) / args.steps if g is not None else None
    ms = ms_graph if ms_graph is not None else ms_eager
    with torch.cuda.stream(ts):
        for _ in range(args.copy_check):
            y.copy_(x)
    torch.cuda.synchronize()
    alg_bytes = 16.0 * D
    gbs = alg_bytes / (ms * 1e-3) / 1e9
    traffic = hbm_traffic(os.path.join(ROOT, "profiles", f"hv_{args.config}_traffic.md"))
    batched = None if args.no_batched else batched_vs_plain()
    solves = None if (args.no_batched or args.no_solves) else config_solves()
    loops = None if (args.no_batched or args.no_solves or args.no_dmft) else dmft_loops()
    cpu = cpu_extra = None
    if not args.no_cpu:
        cpu = cpu_reference_hv(cfg, nu, nd) or cpu_port_hv(cfg, nu, nd)
        cpu_extra = cpu_extra_baselines(gpu_extra_timings())
    out = {
        "metric": METRIC, "value": round(gbs, 2), "unit": "GB/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms, 6), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"H v, {cfg.nsite}-site sector N_up={nu} N_dn={nd}, D={D}",
                   "norb": cfg.norb, "nbath": cfg.nbath, "U": cfg.U, "J": cfg.J, "sector": [nu, nd], "dim": D,
                   "params": "seeded bath, tests/golden/configs.json"},
        "roofline": {"bound": "hbm", "achieved": round(gbs, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(gbs / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "alg_bytes_per_launch": alg_bytes,
                     "kernel": "one H v = row pass (hv_rows_persist_kernel at this sector) + hv_strip_kernel on one stream"},
        "cpu_baseline": cpu,
        "cpu_baseline_more": cpu_extra,
        "batched_vs_plain": batched,
        "config_solves": solves,
        "dmft_loop": loops,
        "wall_s_warm_and_timed": round(wall, 4),
        "untimed_load_steps": n_warm * args.steps,
        "ms_per_step_eager": round(ms_eager, 6),
        "timing": "HIP graph of the K steps" if ms_graph is not None else "eager loop",
    }
    if cpu:
        out["gpu_over_cpu"] = round(gbs / cpu["value"], 1)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
