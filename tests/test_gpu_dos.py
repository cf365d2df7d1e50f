"""GPU parity of the real-axis spectral function (dos.f90): dos_diag against the reference build's
outputs, and dos over the Green's-function records of a solve against the oracle; the on-disk
records (apbpanbn.dat, eigenvector files) against the reference build's bytes and in round trips."""
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


def _model():
    rng = np.random.default_rng(21)
    norb, nbath = 2, 4
    ek = np.concatenate([np.full(norb, -1.5), rng.uniform(-2, 2, nbath)])
    vk = np.zeros((norb, nbath))
    for b in range(nbath):
        vk[b % norb, b] = rng.uniform(0.3, 0.9)
    return norb, nbath, ek, vk


def test_apbpanbn_import_and_export_match_the_reference_bytes(tmp_path):
    """a file written with ed_calc_green_ftn's record statements (reference build) reads back as
    its records, and moed_export_green writes the same bytes"""
    z = np.load(os.path.join(GOLD, "ref_files.npz"))
    path = tmp_path / "apbpanbn.dat"
    path.write_bytes(z["ab_bytes"].tobytes())
    norb, nbath, ek, vk = _model()
    with EDSolver(norb, nbath, 3.0, 0.4, ek, vk, nstep=10) as s:
        s.import_green(str(path))
        recs = s.green_records()
        assert len(recs) == int(z["nrec"])
        for i, r in enumerate(recs):
            nsp, nsn, iev, iorb, even = (int(v) for v in z[f"rec{i}_hdr"])
            assert (r["nsp"], r["nsn"], r["iev"], r["iorb"], r["even"]) == (nsp, nsn, iev, iorb, bool(even))
            assert (r["E"], r["p"]) == tuple(z[f"rec{i}_vp"])
            assert r["ispin"] == (1 if even or i in (2, 3) else 2)
            for k, n in (("ap", nsp), ("bp", nsp), ("an", nsn), ("bn", nsn)):
                assert np.array_equal(r[k][:n + 1], z[f"rec{i}_{k}"])
                # records shorter than the handle's nstep come back zero-padded to nstep + 1
                assert n < 10 and np.all(r[k][n + 1:] == 0.0)
        out = tmp_path / "again.dat"
        s.export_green(str(out))
        assert out.read_bytes() == z["ab_bytes"].tobytes()
    with EDSolver(norb, nbath, 3.0, 0.4, ek, vk, nstep=5) as s:  # records of up to 7 steps do not fit
        with pytest.raises(Exception):
            s.import_green(str(path))
    bad = tmp_path / "truncated.dat"
    bad.write_bytes(z["ab_bytes"].tobytes()[:-3])
    with EDSolver(norb, nbath, 3.0, 0.4, ek, vk, nstep=10) as s:
        with pytest.raises(Exception):
            s.import_green(str(bad))


def test_dos_of_imported_records_matches_reference_build(tmp_path):
    """moed_import_green of the reference-written apbpanbn.dat, then moed_dos: Aff equals the
    reference build's dos (dos.f90:1-63) over the same file, bit for bit"""
    z = np.load(os.path.join(GOLD, "ref_files.npz"))
    path = tmp_path / "apbpanbn.dat"
    path.write_bytes(z["ab_bytes"].tobytes())
    norb, nbath, ek, vk = _model()
    with EDSolver(norb, nbath, 3.0, 0.4, ek, vk, nstep=10) as s:
        s.import_green(str(path))
        a = s.dos(z["dos_omega"], float(z["dos_small"]))
        assert np.array_equal(a, z["dos_aff"])


def test_records_and_eigenvectors_round_trip_after_a_solve(tmp_path):
    """export after ed_solve / ed_calc_green_ftn, import into a fresh solver: the same records and
    the same spectral function; every eigenvector file reads back as the device's eigenvector"""
    import moed
    norb, nbath, ek, vk = _model()
    om = (np.arange(1, 201) - 100) * 0.01
    args = dict(beta=20.0, nstep=30, nev=2, sectors=[(3, 2), (2, 3), (3, 3)])
    with EDSolver(norb, nbath, 3.0, 0.4, ek, vk, **args) as s:
        s.ed_solve()
        s.ed_calc_green_ftn(omega_grid(64, 20.0))
        s.export_green(str(tmp_path / "apbpanbn.dat"))
        s.export_eigvecs(str(tmp_path))
        recs, aff = s.green_records(), s.dos(om)
        vecs = {(i, l): s.eigvec(i, l) for i in (1, 2, 3) for l in (1, 2)}
    with EDSolver(norb, nbath, 3.0, 0.4, ek, vk, **args) as t:
        t.import_green(str(tmp_path / "apbpanbn.dat"))
        recs2 = t.green_records()
        assert np.array_equal(t.dos(om), aff)
    assert len(recs2) == len(recs)
    for a, b in zip(recs, recs2):
        assert all(np.array_equal(a[k], b[k]) for k in a)
    for (i, l), v in vecs.items():
        assert np.array_equal(moed.import_eigvec(str(tmp_path), i, l), v)
