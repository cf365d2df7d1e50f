"""GPU parity of the sector H v and apply_c kernels against the oracle (bit-exact matrix elements,
product within rounding) and size-independent properties at the 14-site sector."""
import os

import numpy as np
import pytest

import oracle
import moed
from moed import EDSolver
from moed.model import load_configs, start_vector

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")


def omodel(norb, nbath, U, J, ek, vk, tij=None):
    return oracle.Model(norb, nbath, U, J, ek, vk, tij)


def configs_by_name():
    return {c.name: c for c in load_configs()}


def random_model(rng, norb, nbath, tij=False, dense_v=False):
    ns = norb + nbath
    ek = rng.uniform(-2, 2, ns)
    vk = rng.uniform(0.1, 1.0, (norb, nbath))
    if not dense_v:  # block-diagonal coupling as in the workloads
        nb = max(1, nbath // norb)
        mask = np.zeros((norb, nbath), bool)
        for k in range(nbath):
            mask[min(k // nb, norb - 1), k] = True
        vk = np.where(mask, vk, 0.0)
    t = None
    if tij:
        t = rng.uniform(-0.3, 0.3, (norb, norb))
        t = t + t.T
        np.fill_diagonal(t, 0.0)
    return ek, vk, t


def gpu_columns(s, nu, nd, D):
    cols = []
    for j in range(D):
        e = np.zeros(D)
        e[j] = 1.0
        y = s.multiply_H(nu, nd, e)
        nz = np.nonzero(y)[0]
        cols.append((nz.astype(np.int64), y[nz]))
    return cols


CASES = [
    # norb, nbath, nu, nd, tij, dense_v
    (2, 4, 3, 2, False, False),
    (1, 5, 3, 3, False, False),
    (3, 3, 3, 4, True, False),
    (4, 2, 3, 3, True, False),
    (2, 3, 2, 2, False, True),
]


@pytest.mark.parametrize("case", CASES)
def test_matrix_elements_bit_exact(case):
    norb, nbath, nu, nd, tij, dense_v = case
    print("case", case, flush=True)
    rng = np.random.default_rng(hash(case) % 2**32)
    ek, vk, t = random_model(rng, norb, nbath, tij, dense_v)
    m = omodel(norb, nbath, 3.5, 0.45, ek, vk, t)
    b = oracle.Basis(norb + nbath, nu, nd)
    with EDSolver(norb, nbath, 3.5, 0.45, ek, vk, tij=t) as s:
        up, dn = s.generate_basis(nu, nd)
        assert np.array_equal(up, b.up) and np.array_equal(dn, b.dn)
        cols = gpu_columns(s, nu, nd, b.ntot)
    ref = oracle.h_rows(m, b)  # H is exactly symmetric: row j = column j
    assert oracle.h_digest(cols) == oracle.h_digest(ref)


def test_matrix_elements_match_reference_build_8site(configs):
    c = configs["ns8_plain"]
    z = np.load(os.path.join(GOLD, "ref_ns8.npz"))
    nu, nd = c.sectors[0]
    with EDSolver.from_config(c) as s:
        D = s.dim(nu, nd)
        cols = gpu_columns(s, nu, nd, D)
    assert oracle.h_digest(cols) == str(z["h_digest"])
    diag = np.array([v[r == j][0] if np.any(r == j) else 0.0 for j, (r, v) in enumerate(cols)])
    assert np.array_equal(diag, z["h_diag"])


def test_hv_product_8site_all_launch_shapes(configs):
    c = configs["ns8_plain"]
    z = np.load(os.path.join(GOLD, "ref_ns8.npz"))
    nu, nd = c.sectors[0]
    with EDSolver.from_config(c) as s:
        x = start_vector(int(z["x_seed"]), s.dim(nu, nd))
        ys = []
        for threads in (0, 256, 512, 1024):
            for strip in (0, 64, 128):
                s.set_hv_shape(threads, strip)
                ys.append(s.multiply_H(nu, nd, x))
    for y in ys[1:]:
        assert np.array_equal(y, ys[0])  # every element's terms are summed in the same order
    ref = z["hv_y"]
    scale = np.abs(ref).max()
    assert np.max(np.abs(ys[0] - ref)) <= 1e-14 * scale


def test_hv_12site_against_oracle_and_reference(configs):
    c = configs["ns12_plain"]
    z = np.load(os.path.join(GOLD, "ref_ns12.npz"))
    nu, nd = c.sectors[0]
    with EDSolver.from_config(c) as s:
        x = start_vector(int(z["x_seed"]), s.dim(nu, nd))
        y = s.multiply_H(nu, nd, x)
    m = omodel(c.norb, c.nbath, c.U, c.J, c.ek, c.vk)
    yo = oracle.multiply_H(m, oracle.Basis(c.nsite, nu, nd), x)
    scale = np.abs(yo).max()
    assert np.max(np.abs(y - yo)) <= 1e-14 * scale
    assert np.max(np.abs(y[z["idx"]] - z["y_at"])) <= 1e-14 * scale


def test_hv_14site_properties_and_oracle(configs):
    """full-size sector (D = 11,778,624): symmetry, linearity and the oracle's product"""
    c = configs["ns14_roofline"]
    nu, nd = c.sectors[0]
    with EDSolver.from_config(c) as s:
        D = s.dim(nu, nd)
        x = start_vector(1, D)
        y = start_vector(2, D)
        Hx = s.multiply_H(nu, nd, x)
        Hy = s.multiply_H(nu, nd, y)
        Hxy = s.multiply_H(nu, nd, 0.5 * x - 2.0 * y)
    assert abs(np.dot(y, Hx) - np.dot(Hy, x)) <= 1e-12 * np.linalg.norm(Hx) * np.linalg.norm(y)
    assert np.max(np.abs(Hxy - (0.5 * Hx - 2.0 * Hy))) <= 1e-12 * np.abs(Hx).max()
    m = omodel(c.norb, c.nbath, c.U, c.J, c.ek, c.vk)
    yo = oracle.multiply_H(m, oracle.Basis(c.nsite, nu, nd), x)
    assert np.max(np.abs(Hx - yo)) <= 1e-14 * np.abs(yo).max()


@pytest.mark.parametrize("sector", [(7, 7), (6, 7)])
def test_hv_14site_resident_row_pass_equals_per_row_blocks(configs, sector):
    """the default row pass at 14 sites (resident blocks with the records in registers; nU = 3432
    stages 16-byte pairs only, nU = 3003 also its last element alone) against one block per row of
    1024 and 512 threads: bit-identical; (6, 7) also against the oracle"""
    c = configs["ns14_roofline"]
    nu, nd = sector
    with EDSolver.from_config(c, tuned=False) as s:
        D = s.dim(nu, nd)
        x = start_vector(nu * 7 + nd, D)
        ys = []
        for threads in (0, 1024, 512):
            s.set_hv_shape(threads, 0)
            ys.append(s.multiply_H(nu, nd, x))
    for y in ys[1:]:
        assert np.array_equal(y, ys[0])
    if sector == (6, 7):
        m = omodel(c.norb, c.nbath, c.U, c.J, c.ek, c.vk)
        yo = oracle.multiply_H(m, oracle.Basis(c.nsite, nu, nd), x, min(16, os.cpu_count() or 1))
        assert np.max(np.abs(ys[0] - yo)) <= 1e-14 * np.abs(yo).max()


def test_hv_15site_tij_rows_and_properties(configs):
    """the largest sector: 15 sites, 3 orbitals with impurity hopping, (7, 8), D = 41,409,225 (odd
    nU = 6435): 400 rows against the oracle's rows of H (first and last rows, every row-block edge,
    and a seeded sample), symmetry and linearity of the whole product"""
    c = configs["ns15_fit"]
    nu, nd = c.sectors[0]
    with EDSolver.from_config(c) as s:
        D = s.dim(nu, nd)
        x = start_vector(3, D)
        y = start_vector(4, D)
        Hx = s.multiply_H(nu, nd, x)
        Hy = s.multiply_H(nu, nd, y)
        Hxy = s.multiply_H(nu, nd, 2.0 * x + 0.25 * y)
    assert abs(np.dot(y, Hx) - np.dot(Hy, x)) <= 1e-12 * np.linalg.norm(Hx) * np.linalg.norm(y)
    assert np.max(np.abs(Hxy - (2.0 * Hx + 0.25 * Hy))) <= 1e-12 * np.abs(Hx).max()
    m = omodel(c.norb, c.nbath, c.U, c.J, c.ek, c.vk, c.tij)
    b = oracle.Basis(c.nsite, nu, nd)
    nU = b.ntot // oracle.icom(c.nsite, nd)
    rng = np.random.default_rng(15)
    rows = np.unique(np.concatenate([[0, 1, D - 2, D - 1], np.arange(0, D, nU)[::97], np.arange(nU - 1, D, nU)[::97],
                                     rng.integers(0, D, 400)]))[:400]
    scale = np.abs(Hx).max()
    for r, (cols, vals) in zip(rows, oracle.h_rows(m, b, rows=rows)):
        assert abs(Hx[r] - np.dot(vals, x[cols])) <= 1e-14 * scale, r


# move-slot counts whose records once had a stride shorter than the dwords the row pass reads
# (2, 3, 10, 11, 28, 29 slots), small sectors on both sides of the byte-offset limit nU <= 4095,
# and sectors with more than 8 down electrons (15 down-site adds)
SLOT_CASES = [
    # norb, nbath, dense V, t_ij, V entries kept (None: all), sectors, slots
    (1, 2, True, False, None, [(1, 1), (2, 1)], 2),
    (1, 3, True, False, None, [(2, 2), (1, 3)], 3),
    (2, 8, False, False, None, [(3, 4)], 8),
    (1, 11, True, False, None, [(3, 3), (6, 6)], 11),
    (2, 13, True, True, None, [(1, 10), (7, 0), (2, 2), (7, 15)], 28),
    (3, 12, True, True, 23, [(2, 2), (1, 9), (7, 15), (8, 1)], 29),
]


@pytest.mark.parametrize("case", SLOT_CASES)
def test_hv_slot_counts_and_sector_shapes(case):
    norb, nbath, dense, tij, keep, sectors, nslot = case
    print("case", case, flush=True)
    rng = np.random.default_rng(100 + nslot)
    ek, vk, t = random_model(rng, norb, nbath, tij, dense)
    if keep is not None:  # keep the first `keep` couplings of a seeded order
        idx = rng.permutation(vk.size)[keep:]
        vk.reshape(-1)[idx] = 0.0
    assert np.count_nonzero(vk) + (0 if t is None else np.count_nonzero(t)) == nslot
    m = omodel(norb, nbath, 2.5, 0.4, ek, vk, t)
    with EDSolver(norb, nbath, 2.5, 0.4, ek, vk, tij=t) as s:
        for nu, nd in sectors:
            D = s.dim(nu, nd)
            x = start_vector(nu * 17 + nd, D)
            y = s.multiply_H(nu, nd, x)
            yo = oracle.multiply_H(m, oracle.Basis(norb + nbath, nu, nd), x)
            assert np.isfinite(y).all(), (nu, nd)
            assert np.max(np.abs(y - yo)) <= 1e-14 * np.abs(yo).max(), (nu, nd)


@pytest.mark.parametrize("name, sector", [("ns15_fit", (2, 2)), ("ns15_fit", (8, 1)), ("ns15_fit", (1, 14)),
                                          ("ns15_fit", (6, 5)), ("ns12_plain", (6, 6)), ("ns12_plain", (5, 6)),
                                          ("ns14_roofline", (7, 7))])
def test_hv_device_vectors_at_either_8_byte_phase(name, sector):
    """moed_hv_device with x and y starting at either 8-byte phase (odd nU at 15 sites, where every
    C(15, k) is odd, and even nU at 12 and 14 sites; the row pass stages and the 128-column strips
    read 16-byte pairs at either phase; at 14 sites the row pass is the resident one): y is
    bit-identical to the host-buffer product and to the 64-column (8-byte load) strips in every
    case, and nothing is written outside y.  The device buffers come from the HIP runtime the
    library uses."""
    import ctypes as ct
    hip = ct.CDLL("libamdhip64.so")
    hip.hipMalloc.argtypes = [ct.POINTER(ct.c_void_p), ct.c_size_t]
    hip.hipMemcpy.argtypes = [ct.c_void_p, ct.c_void_p, ct.c_size_t, ct.c_int]
    hip.hipFree.argtypes = [ct.c_void_p]
    H2D, D2H = 1, 2
    c = configs_by_name()[name]
    nu, nd = sector
    with EDSolver.from_config(c) as s:
        D = s.dim(nu, nd)
        x = start_vector(nu * 31 + nd, D)
        ref = s.multiply_H(nu, nd, x)
        nbytes = 8 * (D + 2)
        bufs = [ct.c_void_p(), ct.c_void_p()]
        for b in bufs:
            assert hip.hipMalloc(ct.byref(b), nbytes) == 0
        try:
            for strip in (128, 64):
                s.set_hv_shape(0, strip)
                for xoff in (0, 1):
                    for yoff in (0, 1):
                        print("case", name, sector, strip, xoff, yoff, flush=True)
                        xh = np.zeros(D + 2)
                        xh[xoff:xoff + D] = x
                        assert hip.hipMemcpy(bufs[0], xh.ctypes.data, nbytes, H2D) == 0
                        assert hip.hipMemcpy(bufs[1], np.zeros(D + 2).ctypes.data, nbytes, H2D) == 0
                        s.multiply_H_device(nu, nd, bufs[0].value + 8 * xoff, bufs[1].value + 8 * yoff)
                        assert hip.hipDeviceSynchronize() == 0
                        y = np.empty(D + 2)
                        assert hip.hipMemcpy(y.ctypes.data, bufs[1], nbytes, D2H) == 0
                        assert np.array_equal(y[yoff:yoff + D], ref)
                        assert y[:yoff].sum() == 0.0 and y[yoff + D:].sum() == 0.0  # nothing written outside y
        finally:
            for b in bufs:
                hip.hipFree(b)


@pytest.mark.parametrize("case", [(2, 6, 4, 4), (3, 3, 3, 2), (1, 4, 0, 2), (2, 4, 6, 5)])
def test_apply_c_bit_exact(case):
    norb, nbath, nu, nd = case
    print("case", case, flush=True)
    rng = np.random.default_rng(sum(case))
    ek, vk, _ = random_model(rng, norb, nbath)
    ns = norb + nbath
    b = oracle.Basis(ns, nu, nd)
    x = rng.standard_normal(b.ntot)
    with EDSolver(norb, nbath, 2.0, 0.3, ek, vk) as s:
        for pm in (1, 2):
            for iorb in range(1, norb + 1):
                for ispin in (1, 2):
                    du, dd = (1, 0) if ispin == 1 else (0, 1)
                    sg = 1 if pm == 1 else -1
                    if not (0 <= nu + sg * du <= ns and 0 <= nd + sg * dd <= ns):
                        with pytest.raises(Exception):
                            s.apply_c(nu, nd, x, pm, iorb, ispin)
                        continue
                    _, out = s.apply_c(nu, nd, x, pm, iorb, ispin)
                    _, ref = oracle.apply_c(b, x, pm, iorb, ispin)
                    assert np.array_equal(out, ref), (pm, iorb, ispin)


def test_apply_c_matches_reference_build(configs):
    c = configs["ns8_plain"]
    z = np.load(os.path.join(GOLD, "ref_ns8.npz"))
    nu, nd = c.sectors[0]
    with EDSolver.from_config(c) as s:
        x = start_vector(int(z["x_seed"]), s.dim(nu, nd))
        for pm in (1, 2):
            for iorb in (1, 2):
                for ispin in (1, 2):
                    _, out = s.apply_c(nu, nd, x, pm, iorb, ispin)
                    assert np.array_equal(out, z[f"applyc_{pm}{iorb}{ispin}"])


def test_bad_arguments_raise():
    """each error class comes back as its MOED_E* code (the reference aborts through die(), sys.F:5-41):
    a sector out of range, a bad operator, a zero start vector, results asked for before
    moed_solve, a non-finite bath"""
    rng = np.random.default_rng(0)
    ek, vk, _ = random_model(rng, 2, 2)

    def code(f, *a):
        with pytest.raises(moed.MoedError) as ei:
            f(*a)
        return ei.value.code

    with EDSolver(2, 2, 1.0, 0.1, ek, vk, sectors=[(2, 2)], nev=1) as s:
        assert code(s.multiply_H, 5, 1, np.zeros(4)) == moed.MOED_ESECTOR
        assert code(s.apply_c, 2, 2, np.zeros(36), 3, 1, 1) == moed.MOED_EARG
        assert code(s.lanczos_iteration, 2, 2, np.zeros(36), 10) == moed.MOED_EZERO
        assert code(s.ed_calc_green_ftn, moed.omega_grid(8, 10.0)) == moed.MOED_ESTATE
        assert code(s.sector_eigenvalues, 1) == moed.MOED_ESTATE
        s.ed_solve()
        assert code(s.sector_eigenvalues, 2) == moed.MOED_EARG
    with pytest.raises(moed.MoedError):
        EDSolver(2, 2, 1.0, 0.1, np.array([0, 0, np.nan, 0]), vk)
