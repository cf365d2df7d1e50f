"""Autotuner for the launch shapes of libmoed: times the candidates on the GPU it runs on and writes
the winners to ``tuning.json`` beside this file, which ``EDSolver.from_config`` applies.

Tuned per workload (every candidate gives bit-identical results, so tuning changes only speed):

  * the sector H v (multiply_H, ed_hamiltonian.F90:88-191) of the workload's largest sector: row
    pass shape (0: the default, resident blocks where the sector allows them; 256, 512, 1024: one
    block of that size per row) x strip width in columns (64, 128), timed over HIP-graph replays
    of the product on resident vectors;
  * for workloads with many sectors, the largest sector that the batched solve runs in its
    persistent work-queue launch (none, 2048, 8192, 32768, all), timed over whole solves.

    python -m moed.autotune [--workloads ns14_roofline,ns15_fit,...] [--reps 3]
"""
from __future__ import annotations

import argparse
import json
import os
import time

import numpy as np

from . import MOED_F_BATCHED, EDSolver, lib
from .model import config

HERE = os.path.dirname(os.path.abspath(__file__))
TABLE = os.path.join(HERE, "tuning.json")
HV_SHAPES = [(t, w) for t in (0, 256, 512, 1024) for w in (64, 128)]
PERSIST = [0, 2048, 8192, 32768, 1 << 20]


def load_table(path: str = TABLE) -> dict:
    if not os.path.exists(path):
        return {}
    with open(path) as f:
        return json.load(f).get("workloads", {})


def apply(solver: EDSolver, workload: str, table: dict | None = None) -> dict:
    """set the tuned shapes of `workload` on the solver (nothing when the table has no entry)"""
    t = (load_table() if table is None else table).get(workload)
    if not t:
        return {}
    if "hv_threads" in t:
        solver.set_hv_shape(t["hv_threads"], t["hv_strip"])
    if "persistent_max_dim" in t:
        solver.set_persistent(t["persistent_max_dim"])
    return t


def _time_hv(s: EDSolver, nu: int, nd: int, steps: int, reps: int) -> float:
    import torch
    D = s.dim(nu, nd)
    dev = torch.device("cuda:0")
    x = torch.empty(D, dtype=torch.float64, device=dev)
    y = torch.empty(D, dtype=torch.float64, device=dev)
    ts = torch.cuda.Stream(device=dev)
    st = ts.cuda_stream
    assert lib().moed_fill_start_device(s.h, 7, D, x.data_ptr(), st) == 0
    for _ in range(3):
        s.multiply_H_device(nu, nd, x.data_ptr(), y.data_ptr(), st)
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=ts):
        for _ in range(steps):
            s.multiply_H_device(nu, nd, x.data_ptr(), y.data_ptr(), st)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    best = None
    with torch.cuda.stream(ts):
        g.replay()  # load before the timed replays
        for _ in range(reps):
            e0.record(ts)
            g.replay()
            e1.record(ts)
            torch.cuda.synchronize()
            t = e0.elapsed_time(e1) / steps * 1e-3
            best = t if best is None else min(best, t)
    return best


def tune_hv(workload: str, steps: int = 20, reps: int = 3) -> dict:
    c = config(workload)
    nu, nd = max(c.sectors, key=lambda s: c.dim(*s))
    res = {}
    with EDSolver.from_config(c) as s:
        for thr, strip in HV_SHAPES:
            s.set_hv_shape(thr, strip)
            res[f"{thr}x{strip}"] = _time_hv(s, nu, nd, steps, reps)
    best = min(res, key=res.get)
    thr, strip = (int(v) for v in best.split("x"))
    return {"hv_threads": thr, "hv_strip": strip, "hv_sector": [nu, nd],
            "hv_us": {k: round(v * 1e6, 2) for k, v in res.items()}}


def tune_persistent(workload: str, reps: int = 2) -> dict:
    c = config(workload)
    res = {}
    for pmax in PERSIST:
        best = None
        for _ in range(reps + 1):  # the first run also builds the string tables
            with EDSolver.from_config(c, flags=MOED_F_BATCHED) as s:
                s.set_persistent(pmax)
                t0 = time.perf_counter()
                s.ed_solve()
                t = time.perf_counter() - t0
            best = t if best is None else min(best, t)
        res[str(pmax)] = best
    pick = int(min(res, key=res.get))
    return {"persistent_max_dim": pick, "solve_s": {k: round(v, 4) for k, v in res.items()}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--workloads", default="ns8_plain,ns12_plain,ns14_roofline,ns10_batched,ns15_fit")
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--out", default=TABLE)
    a = ap.parse_args()
    import torch
    out = {"device": torch.cuda.get_device_name(0), "workloads": {}}
    for w in a.workloads.split(","):
        t = tune_hv(w, reps=a.reps)
        if len(config(w).sectors) > 1:
            t.update(tune_persistent(w))
        out["workloads"][w] = t
        print(w, json.dumps(t), flush=True)
    with open(a.out, "w") as f:
        json.dump(out, f, indent=1)
    print("wrote", a.out)


if __name__ == "__main__":
    main()
