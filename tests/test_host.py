"""Host-side logic of the package (no GPU): the DMFT loop's parameter packing and start points,
and the band structure, against the oracle's restatements of the reference."""
import numpy as np

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
