"""Workload definitions and small host-side helpers shared by tests and bench.py.

``Config`` carries the reference's configuration globals (ed_config.f90:11-37):
Norb, Nbath (total bath sites), Nsite, U, Jex, beta, ek(1:Nsite), vk(1:Norb, Norb+1:Nsite),
the sectors as (N_up, N_dn), Nstep, Nw, Nev -- plus the impurity-hopping extension t_ij.
"""
from __future__ import annotations

import json
import math
import os
from dataclasses import dataclass, field

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
CONFIGS_JSON = os.path.join(ROOT, "tests", "golden", "configs.json")


@dataclass
class Config:
    name: str
    norb: int
    nbath: int
    nsite: int
    U: float
    J: float
    mu: float
    beta: float
    ek: np.ndarray
    vk: np.ndarray  # (norb, nbath)
    tij: np.ndarray | None
    sectors: list  # [(nu, nd)]
    nstep: int
    nw: int
    nev: int
    extra: dict = field(default_factory=dict)

    def dim(self, nu, nd):
        return math.comb(self.nsite, nu) * math.comb(self.nsite, nd)


def _h(v):
    return float.fromhex(v)


def load_configs(path: str = CONFIGS_JSON) -> list[Config]:
    with open(path) as f:
        raw = json.load(f)
    out = []
    for c in raw:
        norb, nbath = c["norb"], c["nbath"]
        tij = None
        if c.get("tij_colmajor"):
            tij = np.array([_h(v) for v in c["tij_colmajor"]]).reshape(norb, norb, order="F")
        out.append(Config(
            name=c["name"], norb=norb, nbath=nbath, nsite=c["nsite"], U=_h(c["U"]), J=_h(c["J"]),
            mu=_h(c["mu"]), beta=_h(c["beta"]), ek=np.array([_h(v) for v in c["ek"]]),
            vk=np.array([_h(v) for v in c["vk_colmajor"]]).reshape(norb, nbath, order="F"), tij=tij,
            sectors=[(nup, nel - nup) for nel, nup in c["sectors_nelec_nup"]], nstep=c["nstep"], nw=c["nw"],
            nev=c["nev"], extra={k: c[k] for k in ("nfit", "ndmft", "nb_per_orb", "seed") if k in c}))
    return out


def config(name: str) -> Config:
    for c in load_configs():
        if c.name == name:
            return c
    raise KeyError(name)


_M64 = (1 << 64) - 1


def start_vector(seed: int, n: int) -> np.ndarray:
    """Counter-based start vector: x_i = 2*u_i - 1, u_i = splitmix64(seed ^ i) >> 11 * 2^-53.
    The device kernel moed_fill_start computes the same bits in place."""
    i = np.arange(n, dtype=np.uint64)
    z = (np.uint64(seed) ^ i) + np.uint64(0x9E3779B97F4A7C15)
    z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
    z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    z = z ^ (z >> np.uint64(31))
    u = (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
    return 2.0 * u - 1.0


def omega_grid(nw: int, beta: float) -> np.ndarray:
    """Matsubara grid of ed_green_init at one rank (ed_green.F90:45-51):
    w_n = ((2*float(n) + 1) * pi) / beta, pi = 4*atan(1d0)."""
    n = np.arange(nw, dtype=np.float32).astype(np.float64)
    return (2.0 * n + 1) * (4.0 * math.atan(1.0)) / beta
