"""CPU oracle for the ED impurity solver -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s cpu_baseline leg may
import this package, and only as the checker.  The product library never calls it.

Contents
  * ``liboracle.so`` (oracle.c): a plain-C restatement of the reference hot path
    (ed_basis.F90, ed_operators.F90, ed_hamiltonian.F90:29-85,88-191, ed_lanczos.F90:14-68,
    ed_solver.F90:193-219, ed_green.F90:25-253,260-273, func.f90, frprmn.f90, dos.f90:65-106),
    wrapped here with ctypes; ``dos`` is the record loop of dos.f90:1-63 in Python.
    Every function is pinned by the tests in tests/test_oracle.py (reference-build outputs,
    dense ED, closed forms); PARPACK, which the reference's diag calls, is absent and is not
    restated (E0 is pinned by dense ED and SciPy's ARPACK instead).
  * ``dense_ed`` : an independent brute-force second-quantisation build of the same
    Hamiltonian (numpy/scipy), used to pin the restatement.
  * ``run_ref``  : runs the reference build ``oracle/_ref/ref_driver`` (the reference's own
    Fortran modules compiled from its sources) where it has been built.
"""
from __future__ import annotations

import ctypes as ct
import os
import struct
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_DRIVER = os.path.join(HERE, "_ref", "ref_driver")

_dp = ct.POINTER(ct.c_double)
_ip = ct.POINTER(ct.c_int)
_i32p = ct.POINTER(ct.c_int32)


class _Model(ct.Structure):
    _fields_ = [("norb", ct.c_int32), ("nbath", ct.c_int32), ("nsite", ct.c_int32),
                ("U", ct.c_double), ("J", ct.c_double),
                ("ek", _dp), ("vk", _dp), ("tij", _dp)]


class _Basis(ct.Structure):
    _fields_ = [("nsite", ct.c_int32), ("ne_up", ct.c_int32), ("ne_dn", ct.c_int32),
                ("nup", ct.c_int64), ("ndn", ct.c_int64), ("ntot", ct.c_int64),
                ("up", _i32p), ("dn", _i32p), ("idx_up", _i32p), ("idx_dn", _i32p)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle library not built: {LIB_PATH} (run __graft_entry__.build())")
        L = ct.CDLL(LIB_PATH)
        L.or_band_tb.argtypes = [ct.c_int, ct.c_int, _dp]
        L.or_delta_new.argtypes = [ct.c_int, ct.c_int, ct.c_int, ct.c_double, _dp, _dp, _dp, _dp]
        L.or_n_from_gksum.argtypes = [ct.c_int, ct.c_int, ct.c_double, _dp, _dp]
        L.or_icom.restype = ct.c_int32
        L.or_icom.argtypes = [ct.c_int32, ct.c_int32]
        L.or_basis_make.argtypes = [ct.c_int32, ct.c_int32, ct.c_int32, ct.POINTER(_Basis)]
        L.or_basis_free.argtypes = [ct.POINTER(_Basis)]
        L.or_multiply_H.argtypes = [ct.POINTER(_Model), ct.POINTER(_Basis), _dp, _dp, ct.c_int]
        L.or_diag_element.restype = ct.c_double
        L.or_diag_element.argtypes = [ct.POINTER(_Model), ct.POINTER(_Basis), ct.c_int64]
        L.or_apply_c.argtypes = [ct.POINTER(_Basis), _dp, ct.c_int, ct.c_int, ct.c_int, ct.POINTER(_Basis), _dp]
        L.or_lanczos_iteration.argtypes = [ct.POINTER(_Model), ct.POINTER(_Basis), _dp, ct.c_int, _ip, _dp, _dp,
                                           ct.c_int]
        L.or_boltzmann_factor.argtypes = [_dp, ct.c_int, ct.c_double, _dp]
        L.or_sort.argtypes = [ct.c_int, _dp, _ip]
        L.or_omega_grid.argtypes = [ct.c_int, ct.c_double, _dp]
        L.or_green_cf.argtypes = [ct.c_int, ct.c_double, ct.c_int, _dp, _dp, ct.c_int, _dp, ct.c_int, _dp]
        L.or_dos_diag.argtypes = [ct.c_int, _dp, ct.c_double, ct.c_double, ct.c_int, ct.c_int, _dp, _dp, _dp, _dp,
                                  ct.c_int, _dp]
        L.or_green_diag.argtypes = [ct.POINTER(_Model), ct.POINTER(_Basis), _dp, ct.c_double, ct.c_double,
                                    ct.c_double, ct.c_int, ct.c_int, ct.c_int, ct.c_int, _dp, ct.c_int, _dp,
                                    _ip, _ip, _dp, _dp, _dp, _dp, ct.c_int]
        L.or_calc_dev.argtypes = [ct.POINTER(_Model), _dp, ct.c_int, _dp, ct.c_int, _dp]
        L.or_x_to_ev.argtypes = [_dp, ct.c_int, ct.c_int, ct.c_int, ct.c_int, _dp, _dp, _dp]
        L.or_ev_to_x.argtypes = [_dp, _dp, _dp, ct.c_int, ct.c_int, ct.c_int, _dp]
        L.or_func.restype = ct.c_double
        L.or_func.argtypes = [_dp, ct.c_int, ct.c_int, ct.c_int, _dp, _dp, ct.c_int, ct.c_int, ct.c_int]
        L.or_dfunc.argtypes = [_dp, _dp, ct.c_int, ct.c_int, ct.c_int, _dp, _dp, ct.c_int, ct.c_int, ct.c_int]
        L.or_frprmn.argtypes = [_dp, ct.c_int, ct.c_double, _ip, _dp, ct.c_int, ct.c_int, _dp, _dp, ct.c_int,
                                ct.c_int, ct.c_int, _ip]
        _lib = L
    return _lib


def _d(a):
    return a.ctypes.data_as(_dp)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


class Model:
    """Impurity model: norb orbitals + nbath bath sites (reference globals of ed_config.f90:11-26)."""

    def __init__(self, norb, nbath, U, J, ek, vk, tij=None):
        self.norb, self.nbath, self.nsite = int(norb), int(nbath), int(norb) + int(nbath)
        self.U, self.J = float(U), float(J)
        self.ek = _f64(ek).reshape(self.nsite)
        self.vk = np.asfortranarray(np.asarray(vk, dtype=np.float64).reshape(self.norb, self.nbath, order="F"))
        self.tij = None if tij is None else np.asfortranarray(np.asarray(tij, np.float64).reshape(norb, norb))
        self._c = _Model(self.norb, self.nbath, self.nsite, self.U, self.J, _d(self.ek),
                         self.vk.ctypes.data_as(_dp),
                         None if self.tij is None else self.tij.ctypes.data_as(_dp))

    @property
    def c(self):
        return ct.byref(self._c)


class Basis:
    """Sector basis (generate_basis, ed_basis.F90:35-107)."""

    def __init__(self, nsite, ne_up, ne_dn):
        self._c = _Basis()
        if lib().or_basis_make(nsite, ne_up, ne_dn, ct.byref(self._c)):
            raise ValueError("bad sector")
        self.nsite, self.ne_up, self.ne_dn = nsite, ne_up, ne_dn
        self.nup, self.ndn, self.ntot = self._c.nup, self._c.ndn, self._c.ntot
        self.up = np.ctypeslib.as_array(self._c.up, (self.nup,)).copy()
        self.dn = np.ctypeslib.as_array(self._c.dn, (self.ndn,)).copy()

    @property
    def c(self):
        return ct.byref(self._c)

    def __del__(self):
        try:
            lib().or_basis_free(ct.byref(self._c))
        except Exception:
            pass


def icom(n, m):
    return lib().or_icom(n, m)


def multiply_H(model, basis, x, nthreads=0):
    x = _f64(x)
    y = np.empty_like(x)
    lib().or_multiply_H(model.c, basis.c, _d(x), _d(y), nthreads)
    return y


def diag_elements(model, basis, rows=None):
    rows = range(basis.ntot) if rows is None else rows
    return np.array([lib().or_diag_element(model.c, basis.c, int(i)) for i in rows])


def apply_c(basis, vec, pm, iorb, ispin):
    dup, ddn = (1, 0) if ispin == 1 else (0, 1)
    s = 1 if pm == 1 else -1
    bout = Basis(basis.nsite, basis.ne_up + s * dup, basis.ne_dn + s * ddn)
    vec = _f64(vec)
    out = np.empty(bout.ntot)
    lib().or_apply_c(basis.c, _d(vec), pm, iorb, ispin, bout.c, _d(out))
    return bout, out


def lanczos(model, basis, vec, nstep, nthreads=0):
    vec = _f64(vec)
    a = np.zeros(nstep + 1)
    b = np.zeros(nstep + 1)
    nc = ct.c_int(0)
    lib().or_lanczos_iteration(model.c, basis.c, _d(vec), nstep, ct.byref(nc), _d(a), _d(b), nthreads)
    return nc.value, a, b


def boltzmann(e, beta):
    e = _f64(e)
    p = np.empty_like(e)
    lib().or_boltzmann_factor(_d(e), len(e), beta, _d(p))
    return p


def sort_index(a):
    """NR heapsort index (func.f90:301-353); returns 1-based indices."""
    a = _f64(a)
    ind = np.zeros(len(a), dtype=np.intc)
    lib().or_sort(len(a), _d(a), ind.ctypes.data_as(_ip))
    return ind


def select_states(eigval_all, nev, beta, threshold=0.001):
    """ed_solve state selection (ed_solver.F90:65-90): sort, Boltzmann, keep p>1e-3 among
    the first nev, renormalise.  Returns list of (val, prob, sector(1-based), level(1-based))."""
    ind = sort_index(eigval_all)
    pev = boltzmann(eigval_all, beta)
    nev_calc = sum(1 for i in range(nev) if pev[ind[i] - 1] > threshold)
    out = []
    Z = 0.0
    for i in range(nev_calc):
        k = ind[i] - 1
        out.append([eigval_all[k], pev[k], k // nev + 1, k % nev + 1])
        Z = Z + pev[k]
    for o in out:
        o[1] = o[1] / Z
    return [tuple(o) for o in out]


def omega_grid(nw, beta):
    om = np.empty(nw)
    lib().or_omega_grid(nw, beta, _d(om))
    return om


def green_cf(sign, E, nsc, a, b, omega, compat=True):
    a, b, omega = _f64(a), _f64(b), _f64(omega)
    g = np.empty(2 * len(omega))
    lib().or_green_cf(sign, E, nsc, _d(a), _d(b), len(omega), _d(omega), int(compat), _d(g))
    return g[0::2] + 1j * g[1::2]


def green_diag(model, basis, eigvec, E, pev, factor, iorb, ispin, nstep, omega, Gr=None, compat=True,
               nthreads=0):
    """green_diag (ed_green.F90:136-185); returns (Gr, (nsp, ap, bp), (nsn, an, bn))."""
    nw = len(omega)
    G = np.zeros(2 * nw) if Gr is None else np.stack([Gr.real, Gr.imag], 1).reshape(-1).copy()
    ap, bp, an, bn = (np.zeros(nstep + 1) for _ in range(4))
    nsp, nsn = ct.c_int(0), ct.c_int(0)
    rc = lib().or_green_diag(model.c, basis.c, _d(_f64(eigvec)), E, pev, factor, iorb, ispin, nstep, nw,
                             _d(_f64(omega)), int(compat), _d(G), ct.byref(nsp), ct.byref(nsn),
                             _d(ap), _d(bp), _d(an), _d(bn), nthreads)
    if rc:
        raise ValueError("green_diag: bad sector")
    return G[0::2] + 1j * G[1::2], (nsp.value, ap, bp), (nsn.value, an, bn)


def dos_diag(omega_real, small, ev, ap, bp, an, bn, compat=True):
    """dos_diag (dos.f90:65-106) for one record: A(w) = -Im(G+ + G-)/pi at z = (ev +- w, small)."""
    om = _f64(omega_real)
    ap, bp, an, bn = _f64(ap), _f64(bp), _f64(an), _f64(bn)
    out = np.empty(len(om))
    lib().or_dos_diag(len(om), _d(om), float(small), float(ev), len(ap) - 1, len(an) - 1, _d(ap), _d(bp), _d(an),
                      _d(bn), int(compat), _d(out))
    return out


def dos(records, norb, omega_real, small, weighted=False, compat=True):
    """dos (dos.f90:1-63) over Green's-function records in file order: each record overwrites its
    orbital's row (the reference applies no probability or spin weight, dos.f90:102).  weighted=True
    instead sums p * factor * A over the records (factor 1 when even, else 1/2, as ed_green.F90:87-94)."""
    aff = np.zeros((norb, len(omega_real)))
    for r in records:
        nsp, nsn = r["nsp"], r["nsn"]
        a = dos_diag(omega_real, small, r["E"], r["ap"][:nsp + 1], r["bp"][:nsp + 1], r["an"][:nsn + 1],
                     r["bn"][:nsn + 1], compat)
        if weighted:
            aff[r["iorb"] - 1] += (r["p"] * (1.0 if r["even"] else 0.5)) * a
        else:
            aff[r["iorb"] - 1] = a
    return aff


def calc_dev(model, ef, omega, compat=True):
    nw = len(omega)
    D = np.empty(2 * model.norb * nw)
    lib().or_calc_dev(model.c, _d(_f64(ef)), nw, _d(_f64(omega)), int(compat), _d(D))
    D = D[0::2] + 1j * D[1::2]
    return D.reshape(model.norb, nw, order="F")


def band_tb(norb, nq):
    """Hk(o,o,q) of ed_set_band_structure (ed_hamiltonian.F90:29-85), shape (nq, norb)"""
    hk = np.empty(nq * norb)
    lib().or_band_tb(norb, nq, _d(hk))
    return hk.reshape(nq, norb)


def delta_new(Gr, Dev, hk, rmu):
    """ed_delta_new (ed_green.F90:187-215): returns (Gksum, D_ev_new), complex (norb, nw)"""
    norb, nw = Gr.shape
    nq = hk.shape[0]
    d = _dev_flat(Dev)
    gks = np.empty(2 * norb * nw)
    lib().or_delta_new(norb, nw, nq, float(rmu), _d(_f64(hk.reshape(-1))), _d(_dev_flat(Gr)), _d(d), _d(gks))
    tocx = lambda v: (v[0::2] + 1j * v[1::2]).reshape(norb, nw, order="F")
    return tocx(gks), tocx(d)


def n_from_gksum(Gksum, beta):
    norb, nw = Gksum.shape
    out = np.empty(norb)
    lib().or_n_from_gksum(norb, nw, float(beta), _d(_dev_flat(Gksum)), _d(out))
    return out


def _dev_flat(Dev):
    """complex (norb, nw) -> interleaved re/im in column-major order (the Fortran layout)."""
    z = np.asarray(Dev, dtype=np.complex128).reshape(-1, order="F")
    out = np.empty(2 * len(z))
    out[0::2], out[1::2] = z.real, z.imag
    return out


def x_to_ev(x, norb, nbath):
    x = _f64(x)
    ns = norb + nbath
    ef, ek, vk = np.empty(norb), np.empty(ns), np.empty(norb * nbath)
    lib().or_x_to_ev(_d(x), len(x), ns, nbath, norb, _d(ef), _d(ek), _d(vk))
    return ef, ek, vk.reshape(norb, nbath, order="F")


def ev_to_x(ek, vk, ef, norb, nbath):
    ns = norb + nbath
    nx = nbath + nbath * norb + norb
    x = np.empty(nx)
    vkf = np.asfortranarray(np.asarray(vk, np.float64).reshape(norb, nbath, order="F"))
    lib().or_ev_to_x(_d(_f64(ek)), vkf.ctypes.data_as(_dp), _d(_f64(ef)), ns, nbath, nx, _d(x))
    return x


def func(x, omega, Dev, norb, nbath, compat=True):
    """FUNC (func.f90:1-45); Dev is complex (norb, nw)."""
    x = _f64(x)
    D = _dev_flat(Dev)
    return lib().or_func(_d(x), norb + nbath, len(omega), len(x), _d(_f64(omega)), _d(D), nbath, norb,
                         int(compat))


def dfunc(x, omega, Dev, norb, nbath, compat=True):
    x = _f64(x)
    D = _dev_flat(Dev)
    g = np.empty_like(x)
    lib().or_dfunc(_d(x), _d(g), len(omega), norb + nbath, len(x), _d(_f64(omega)), _d(D), nbath, norb,
                   int(compat))
    return g


def frprmn(x, omega, Dev, norb, nbath, ftol=1.0e-7, compat=True):
    """minimization -> frprmn (func.f90:144-163, frprmn.f90); returns (rc, x, fret, iter, nfev)."""
    p = _f64(x).copy()
    D = _dev_flat(Dev)
    it, nf = ct.c_int(0), ct.c_int(0)
    fret = ct.c_double(0.0)
    rc = lib().or_frprmn(_d(p), len(p), ftol, ct.byref(it), ct.byref(fret), len(omega), norb + nbath,
                         _d(_f64(omega)), _d(D), nbath, norb, int(compat), ct.byref(nf))
    return rc, p, fret.value, it.value, nf.value


# ------------------------------------------------------------------ reference build

def ref_available():
    return os.path.exists(REF_DRIVER)


def run_ref(task, model, ne_up=0, ne_dn=0, i1=0, i2=0, i3=0, payload=()):
    """Run the reference build on one task (see oracle/ref_driver.f90).  Returns the raw
    output bytes.  Only usable where oracle/_ref has been built."""
    if not ref_available():
        raise RuntimeError("reference build not available")
    with tempfile.TemporaryDirectory() as td:
        fin, fout = os.path.join(td, "in.bin"), os.path.join(td, "out.bin")
        with open(fin, "wb") as f:
            f.write(struct.pack("<8i", task, model.norb, model.nbath, ne_up, ne_dn, i1, i2, i3))
            f.write(struct.pack("<2d", model.U, model.J))
            f.write(model.ek.astype("<f8").tobytes())
            f.write(np.asfortranarray(model.vk).T.reshape(-1).astype("<f8").tobytes())
            for p in payload:
                f.write(np.ascontiguousarray(p).tobytes())
        subprocess.run([REF_DRIVER, fin, fout], check=True, cwd=td, stdout=subprocess.DEVNULL,
                       stderr=subprocess.DEVNULL)
        with open(fout, "rb") as f:
            return f.read()


def h_rows(model, basis, rows=None, cap=512):
    """nonzero elements of the rows of H (or_h_row): list of (cols, vals) with cols ascending."""
    L = lib()
    L.or_h_row.restype = ct.c_int
    L.or_h_row.argtypes = [ct.POINTER(_Model), ct.POINTER(_Basis), ct.c_int64, ct.POINTER(ct.c_int64), _dp, ct.c_int]
    cols = np.zeros(cap, np.int64)
    vals = np.zeros(cap)
    out = []
    for i in (range(basis.ntot) if rows is None else rows):
        n = L.or_h_row(model.c, basis.c, int(i), cols.ctypes.data_as(ct.POINTER(ct.c_int64)), _d(vals), cap)
        if n > cap:
            raise RuntimeError("row capacity exceeded")
        o = np.argsort(cols[:n], kind="stable")
        out.append((cols[:n][o].copy(), vals[:n][o].copy()))
    return out


def h_digest(columns):
    """sha256 over (j, nnz, rows, value bits) of each column's exact nonzeros, the format of
    tests/golden/make_golden.py's reference digest."""
    import hashlib
    h = hashlib.sha256()
    for j, (rows, vals) in enumerate(columns):
        keep = vals != 0.0
        rows, vals = rows[keep], vals[keep]
        h.update(np.int64(j).tobytes())
        h.update(np.int64(len(rows)).tobytes())
        h.update(rows.astype(np.int64).tobytes())
        h.update(vals.astype(np.float64).tobytes())
    return h.hexdigest()
