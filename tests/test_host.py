"""Host-side logic of the package (no GPU): the DMFT loop's parameter packing and start points and
the band structure against the oracle's restatements of the reference; the eigenvector files
against the reference build's bytes."""
import numpy as np
import pytest

import oracle
from moed import dmft


def test_ev_x_packing_matches_reference_layout():
    rng = np.random.default_rng(2)
    norb, nbath = 3, 12
    ek = rng.standard_normal(norb + nbath)
    vk = rng.standard_normal((norb, nbath))
    ef = rng.standard_normal(norb)
    x = dmft.ev_to_x(ek, vk, ef, norb, nbath)
    assert np.array_equal(x, oracle.ev_to_x(ek, vk.reshape(-1, order="F"), ef, norb, nbath))
    ef2, ek2, vk2 = dmft.x_to_ev(x, norb, nbath)
    ef3, ek3, vk3 = oracle.x_to_ev(x, norb, nbath)
    assert np.array_equal(ef2, ef3) and np.array_equal(ek2, ek3) and np.array_equal(vk2, vk3)
    assert np.array_equal(ek2[:norb], ef) and np.array_equal(vk2, vk)


def test_start_points_keep_zero_couplings_and_the_current_bath():
    x0 = np.array([0.5, -1.0, 0.0, 0.7, 0.0, 2.0])
    X = dmft.start_points(x0, 8, 3, 0.1)
    assert np.array_equal(X[0], x0)
    assert np.all(X[:, x0 == 0] == 0)
    assert np.all(np.abs(X[1:][:, x0 != 0] / x0[x0 != 0] - 1) <= 0.1)


def test_band_structure_host_entry_point():
    import moed
    assert np.array_equal(moed.band_tb(3, 400), oracle.band_tb(3, 400))


# ------------------------------------------------------------------ on-disk records (host only)

def _files():
    import os
    return np.load(os.path.join(os.path.dirname(__file__), "golden", "ref_files.npz"))


def test_eigvec_file_is_byte_identical_to_the_reference(tmp_path):
    """export_eigvec (ed_io.F90:13-33): the reference build's file, byte for byte"""
    import moed
    z = _files()
    moed.export_eigvec(str(tmp_path), 7, 3, z["vec"])
    name = moed.eigvec_filename(7, 3)
    assert name == str(z["eig_name"])
    assert (tmp_path / name).read_bytes() == z["eig_bytes"].tobytes()


def test_eigvec_filename_overflows_like_fortran():
    """I2.2 cannot hold sector 121: the reference prints asterisks (ed_io.F90:145)"""
    import moed
    assert moed.eigvec_filename(121, 12) == str(_files()["eig_name_121"])
    assert moed.eigvec_filename(1, 1, 0) == "eigvec-01-001-0000"


def test_import_eigvec_reads_the_reference_file(tmp_path):
    import moed
    z = _files()
    (tmp_path / str(z["eig_name"])).write_bytes(z["eig_bytes"].tobytes())
    assert np.array_equal(moed.import_eigvec(str(tmp_path), 7, 3), z["vec"])


def test_import_eigvec_rejects_missing_and_malformed_files(tmp_path):
    import moed
    z = _files()
    with pytest.raises(Exception):
        moed.import_eigvec(str(tmp_path), 7, 3)
    (tmp_path / str(z["eig_name"])).write_bytes(z["eig_bytes"].tobytes()[:-5])
    with pytest.raises(Exception):
        moed.import_eigvec(str(tmp_path), 7, 3)


def test_tuning_table_names_workloads_and_valid_shapes():
    """the autotuner's committed table (moed/tuning.json): every entry names a workload of
    tests/golden/configs.json and shapes that moed_set_hv_shape / moed_set_persistent accept"""
    from moed import autotune
    from moed.model import load_configs
    names = {c.name for c in load_configs()}
    table = autotune.load_table()
    assert table and set(table) <= names
    for w, t in table.items():
        assert (t["hv_threads"], t["hv_strip"]) in autotune.HV_SHAPES
        assert t["hv_us"][f"{t['hv_threads']}x{t['hv_strip']}"] == min(t["hv_us"].values())
        if "persistent_max_dim" in t:
            assert t["persistent_max_dim"] in autotune.PERSIST
