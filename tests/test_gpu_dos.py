"""GPU parity of the real-axis spectral function (dos.f90): dos_diag against the reference build's
outputs, and dos over the Green's-function records of a solve against the oracle."""
import os

import numpy as np
import pytest

import oracle
from moed import EDSolver
from moed.model import omega_grid

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")


def test_dos_diag_bit_exact_with_reference():
    z = np.load(os.path.join(GOLD, "ref_dos.npz"))
    with EDSolver(1, 1, 0.0, 0.0, np.zeros(2), np.zeros((1, 1))) as s:
        for i in range(int(z["ncase"])):
            a = s.dos_diag(z["omega"], float(z[f"small_{i}"]), float(z[f"ev_{i}"]), z[f"ap_{i}"], z[f"bp_{i}"],
                           z[f"an_{i}"], z[f"bn_{i}"])
            assert np.array_equal(a, z[f"aff_{i}"]), i


@pytest.mark.parametrize("sectors", [[(3, 3)], [(3, 2), (2, 3)]])
def test_dos_after_green_matches_oracle(sectors):
    """6-site, 2-orbital model: dos over the records of ed_calc_green_ftn, reference (overwrite) and
    weighted modes, against the oracle applied to the same records; N_up != N_dn sectors give
    records of both spins with factor 1/2"""
    print("sectors", sectors, flush=True)
    rng = np.random.default_rng(21)
    norb, nbath = 2, 4
    ek = np.concatenate([np.full(norb, -1.5), rng.uniform(-2, 2, nbath)])
    vk = np.zeros((norb, nbath))
    for b in range(nbath):
        vk[b % norb, b] = rng.uniform(0.3, 0.9)
    om = (np.arange(1, 601) - 300) * 0.01
    with EDSolver(norb, nbath, 3.0, 0.4, ek, vk, beta=20.0, nstep=30, nev=2, sectors=sectors) as s:
        s.ed_solve()
        s.ed_calc_green_ftn(omega_grid(64, 20.0))
        recs = s.green_records()
        assert len(recs) > 0
        for weighted in (False, True):
            a = s.dos(om, 0.05, weighted=weighted)
            ref = oracle.dos(recs, norb, om, 0.05, weighted=weighted)
            assert np.array_equal(a, ref), weighted
