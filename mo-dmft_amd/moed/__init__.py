"""moed: Python view of libmoed.so, the MI355X exact-diagonalisation impurity solver.

The class ``EDSolver`` mirrors the reference's ED call surface (yw-choi/MO-DMFT):

  reference (Fortran)                                   here
  generate_basis(ne_up, ne_down)   ed_basis.F90:35      EDSolver.generate_basis(nu, nd)
  multiply_H(basis, A, B)          ed_hamiltonian.F90:88 EDSolver.multiply_H(nu, nd, A)
  apply_c(basis,vec,pm,iorb,ispin) ed_operators.F90:10  EDSolver.apply_c(nu, nd, vec, pm, iorb, ispin)
  lanczos_iteration(basis,vec,..)  ed_lanczos.F90:14    EDSolver.lanczos_iteration(nu, nd, vec, nstep)
  ed_solve                         ed_solver.F90:28     EDSolver.ed_solve()
  ed_calc_green_ftn(nev_calc)      ed_green.F90:60      EDSolver.ed_calc_green_ftn(omega)
  calc_dev                         ed_green.F90:260     EDSolver.calc_dev(ek, vk, ef, omega)
  FUNC / dfunc                     func.f90:1,47        EDSolver.func_dfunc(x, omega, Dev)  (batched)
  minimization(x, ...)             func.f90:144         EDSolver.minimization(x, omega, Dev) (batched)

All compute runs in libmoed.so on the GPU; there is no CPU fallback.  Errors raise MoedError
with the library's message (the reference calls die() at the same points, sys.F:5-41).
"""
from __future__ import annotations

import ctypes as ct
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmoed.so")

MOED_F_EXACT_CMPLX = 1
MOED_F_BATCHED = 8
MOED_F_FIT_TREE = 16
MOED_F_FULL_REORTH = 32
MOED_F_NO_DEGENERACY_CHECK = 64

MOED_EARG, MOED_ESECTOR, MOED_EZERO, MOED_EHIP, MOED_ENOMEM, MOED_ESTATE, MOED_ECONV = -1, -2, -3, -4, -5, -6, -7
_ERR = {-1: "EARG", -2: "ESECTOR", -3: "EZERO", -4: "EHIP", -5: "ENOMEM", -6: "ESTATE", -7: "ECONV"}


class MoedError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"moed error {code} ({_ERR.get(code, '?')}): {msg}")
        self.code = code


_dp = ct.POINTER(ct.c_double)
_ip = ct.POINTER(ct.c_int32)


class Params(ct.Structure):
    _fields_ = [("norb", ct.c_int32), ("nsite", ct.c_int32), ("U", ct.c_double), ("J", ct.c_double),
                ("beta", ct.c_double), ("nstep", ct.c_int32), ("nev", ct.c_int32), ("nw", ct.c_int32),
                ("ek", _dp), ("vk", _dp), ("tij", _dp), ("nsector", ct.c_int32), ("nelec", _ip),
                ("nup", _ip), ("flags", ct.c_uint32), ("gs_tol", ct.c_double), ("gs_maxvec", ct.c_int32),
                ("device", ct.c_int32)]


EXPORTS = [
    "moed_create", "moed_destroy", "moed_last_error", "moed_set_bath", "moed_sector_dim", "moed_basis",
    "moed_hv", "moed_hv_device", "moed_apply_c", "moed_lanczos", "moed_solve", "moed_sector_eigenvalues",
    "moed_eigvec", "moed_green", "moed_green_nrecords", "moed_green_record", "moed_green_cf",
    "moed_green_cf_adjoint", "moed_calc_dev", "moed_chi2_grad", "moed_minimize", "moed_omega_grid",
    "moed_fill_start_device", "moed_set_hv_shape", "moed_set_eig_chunk", "moed_stream", "moed_band_tb",
    "moed_delta_new", "moed_set_fit_launch", "moed_fit_nfev", "moed_dos_diag", "moed_dos", "moed_eigvec_filename",
    "moed_export_eigvec", "moed_import_eigvec", "moed_export_green", "moed_import_green",
    "moed_minimize_device_noalloc", "moed_set_persistent",
]

_lib = None


def lib():
    """Load libmoed.so (raises if it has not been built: the product has no fallback)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"libmoed.so not built at {LIB_PATH}; run __graft_entry__.build()")
        L = ct.CDLL(LIB_PATH)
        P = ct.POINTER
        L.moed_create.argtypes = [P(Params), P(ct.c_void_p)]
        L.moed_destroy.argtypes = [ct.c_void_p]
        L.moed_last_error.argtypes = [ct.c_void_p, ct.c_char_p, ct.c_int32]
        L.moed_set_bath.argtypes = [ct.c_void_p, _dp, _dp]
        L.moed_sector_dim.argtypes = [ct.c_void_p, ct.c_int32, ct.c_int32, P(ct.c_int64)]
        L.moed_basis.argtypes = [ct.c_void_p, ct.c_int32, ct.c_int32, _ip, _ip]
        L.moed_hv.argtypes = [ct.c_void_p, ct.c_int32, ct.c_int32, _dp, _dp]
        L.moed_hv_device.argtypes = [ct.c_void_p, ct.c_int32, ct.c_int32, ct.c_void_p, ct.c_void_p, ct.c_void_p]
        L.moed_apply_c.argtypes = [ct.c_void_p] + [ct.c_int32] * 5 + [_dp, _dp]
        L.moed_lanczos.argtypes = [ct.c_void_p, ct.c_int32, ct.c_int32, _dp, ct.c_int32, _ip, _dp, _dp]
        L.moed_solve.argtypes = [ct.c_void_p, _ip, _dp, _dp, _ip, _ip]
        L.moed_sector_eigenvalues.argtypes = [ct.c_void_p, ct.c_int32, _dp]
        L.moed_eigvec.argtypes = [ct.c_void_p, ct.c_int32, ct.c_int32, _dp]
        L.moed_green.argtypes = [ct.c_void_p, ct.c_int32, _dp, _dp]
        L.moed_green_nrecords.argtypes = [ct.c_void_p, _ip]
        L.moed_green_record.argtypes = [ct.c_void_p, ct.c_int32, _ip, _ip, _ip, _dp, _dp, _ip, _ip, _ip, _dp, _dp,
                                        _dp, _dp]
        L.moed_green_cf.argtypes = [ct.c_void_p, ct.c_int32, ct.c_double, ct.c_int32, _dp, _dp, ct.c_int32, _dp, _dp]
        L.moed_green_cf_adjoint.argtypes = [ct.c_void_p, ct.c_int32, ct.c_double, ct.c_int32, _dp, _dp, ct.c_int32,
                                            _dp, _dp, _dp, _dp]
        L.moed_calc_dev.argtypes = [ct.c_void_p, _dp, _dp, _dp, ct.c_int32, _dp, _dp]
        L.moed_chi2_grad.argtypes = [ct.c_void_p, ct.c_int32, ct.c_int32, _dp, ct.c_int32, _dp, _dp, _dp, _dp]
        L.moed_minimize.argtypes = [ct.c_void_p, ct.c_int32, ct.c_int32, _dp, ct.c_int32, _dp, _dp, ct.c_double,
                                    _dp, _ip, _ip]
        L.moed_minimize_device_noalloc.argtypes = [ct.c_void_p, ct.c_int32, ct.c_int32, ct.c_void_p, ct.c_int32,
                                                   ct.c_void_p, ct.c_void_p, ct.c_double, ct.c_void_p, ct.c_void_p,
                                                   ct.c_void_p, ct.c_void_p]
        L.moed_omega_grid.argtypes = [ct.c_int32, ct.c_double, _dp]
        L.moed_fill_start_device.argtypes = [ct.c_void_p, ct.c_uint64, ct.c_int64, ct.c_void_p, ct.c_void_p]
        L.moed_set_hv_shape.argtypes = [ct.c_void_p, ct.c_int32, ct.c_int32]
        L.moed_set_eig_chunk.argtypes = [ct.c_void_p, ct.c_int32]
        L.moed_set_persistent.argtypes = [ct.c_void_p, ct.c_int64]
        L.moed_set_fit_launch.argtypes = [ct.c_void_p, ct.c_int32]
        L.moed_dos_diag.argtypes = [ct.c_void_p, ct.c_int32, _dp, ct.c_double, ct.c_double, ct.c_int32, ct.c_int32,
                                    _dp, _dp, _dp, _dp, _dp]
        L.moed_dos.argtypes = [ct.c_void_p, ct.c_int32, _dp, ct.c_double, ct.c_int32, _dp]
        L.moed_eigvec_filename.argtypes = [ct.c_int32, ct.c_int32, ct.c_int32, ct.c_char_p, ct.c_int32]
        L.moed_export_eigvec.argtypes = [ct.c_char_p, ct.c_int32, ct.c_int32, ct.c_int32, ct.c_int64, _dp]
        L.moed_import_eigvec.argtypes = [ct.c_char_p, ct.c_int32, ct.c_int32, ct.c_int32, ct.POINTER(ct.c_int64),
                                         _dp]
        L.moed_export_green.argtypes = [ct.c_void_p, ct.c_char_p]
        L.moed_import_green.argtypes = [ct.c_void_p, ct.c_char_p]
        L.moed_fit_nfev.argtypes = [ct.c_void_p, ct.c_int32, _ip]
        L.moed_band_tb.argtypes = [ct.c_int32, ct.c_int32, _dp]
        L.moed_delta_new.argtypes = [ct.c_void_p, ct.c_int32, ct.c_int32, ct.c_double, _dp, _dp, _dp, _dp, _dp]
        L.moed_stream.argtypes = [ct.c_void_p]
        L.moed_stream.restype = ct.c_void_p
        _lib = L
    return _lib


def _d(a):
    return a.ctypes.data_as(_dp)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _cplx_flat(z):
    """complex (norb, nw) -> interleaved (re, im) in Fortran column-major order."""
    z = np.asarray(z, dtype=np.complex128).reshape(-1, order="F")
    out = np.empty(2 * len(z))
    out[0::2], out[1::2] = z.real, z.imag
    return out


def omega_grid(nw, beta):
    om = np.empty(nw)
    rc = lib().moed_omega_grid(nw, beta, _d(om))
    if rc:
        raise MoedError(rc, "omega grid")
    return om


def eigvec_filename(isector, iev, node=0):
    """eigvec_filename (ed_io.F90:141-147)"""
    buf = ct.create_string_buffer(64)
    rc = lib().moed_eigvec_filename(isector, iev, node, buf, 64)
    if rc:
        raise MoedError(rc, "eigvec_filename")
    return buf.value.decode()


def export_eigvec(directory, isector, iev, vec, node=0):
    """export_eigvec (ed_io.F90:13-33): the reference's unformatted eigenvector file"""
    v = _f64(vec)
    rc = lib().moed_export_eigvec(os.fsencode(directory), isector, iev, node, len(v), _d(v))
    if rc:
        raise MoedError(rc, "export_eigvec")


def import_eigvec(directory, isector, iev, node=0):
    """import_eigvec (ed_io.F90:105-139)"""
    n = ct.c_int64()
    rc = lib().moed_import_eigvec(os.fsencode(directory), isector, iev, node, ct.byref(n), None)
    if rc:
        raise MoedError(rc, "import_eigvec: missing or malformed file")
    v = np.empty(n.value)
    rc = lib().moed_import_eigvec(os.fsencode(directory), isector, iev, node, ct.byref(n), _d(v))
    if rc:
        raise MoedError(rc, "import_eigvec")
    return v


def band_tb(norb, nq):
    """diagonal of H(k) on the nk x nk grid (ed_set_band_structure), shape (nq, norb)"""
    hk = np.empty(nq * norb)
    rc = lib().moed_band_tb(norb, nq, _d(hk))
    if rc:
        raise MoedError(rc, "band_tb: nq must be a square")
    return hk.reshape(nq, norb)


class EDSolver:
    """One handle of libmoed: model parameters, sectors and device-resident state."""

    def __init__(self, norb, nbath, U, J, ek, vk, beta=100.0, nstep=50, nev=1, nw=0, sectors=(), tij=None,
                 flags=0, gs_tol=0.0, gs_maxvec=0, device=0):
        self.norb, self.nbath, self.nsite = int(norb), int(nbath), int(norb) + int(nbath)
        self.nstep, self.nev, self.nw, self.beta = int(nstep), int(nev), int(nw), float(beta)
        self._ek = _f64(ek).reshape(self.nsite).copy()
        self._vk = np.asarray(vk, np.float64).reshape(self.norb, self.nbath, order="F").reshape(-1, order="F").copy()
        self._tij = None if tij is None else np.asarray(tij, np.float64).reshape(-1, order="F").copy()
        self.sectors = [(int(a), int(b)) for a, b in sectors]
        self._nelec = np.array([a + b for a, b in self.sectors] or [0], dtype=np.int32)
        self._nup = np.array([a for a, b in self.sectors] or [0], dtype=np.int32)
        p = Params(self.norb, self.nsite, float(U), float(J), self.beta, self.nstep, self.nev, self.nw,
                   _d(self._ek), _d(self._vk), None if self._tij is None else _d(self._tij),
                   len(self.sectors), self._nelec.ctypes.data_as(_ip), self._nup.ctypes.data_as(_ip),
                   int(flags), float(gs_tol), int(gs_maxvec), int(device))
        h = ct.c_void_p()
        rc = lib().moed_create(ct.byref(p), ct.byref(h))
        if rc:
            raise MoedError(rc, "moed_create failed")
        self.h = h

    @classmethod
    def from_config(cls, cfg, tuned=True, **kw):
        """a solver for a named workload; with tuned, the launch shapes the autotuner picked for it
        (moed/tuning.json: results are bit-identical for every shape, only the speed differs)"""
        args = dict(beta=cfg.beta, nstep=cfg.nstep, nev=cfg.nev, nw=cfg.nw, sectors=cfg.sectors, tij=cfg.tij)
        args.update(kw)
        s = cls(cfg.norb, cfg.nbath, cfg.U, cfg.J, cfg.ek, cfg.vk, **args)
        if tuned:
            from .autotune import apply
            apply(s, cfg.name)
        return s

    def close(self):
        if getattr(self, "h", None):
            lib().moed_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _check(self, rc):
        if rc:
            buf = ct.create_string_buffer(512)
            lib().moed_last_error(self.h, buf, 512)
            raise MoedError(rc, buf.value.decode())

    # ------------------------------------------------------------ basis / operators
    def dim(self, nu, nd):
        d = ct.c_int64()
        self._check(lib().moed_sector_dim(self.h, nu, nd, ct.byref(d)))
        return d.value

    def generate_basis(self, nu, nd):
        from math import comb
        up = np.zeros(comb(self.nsite, nu), np.int32)
        dn = np.zeros(comb(self.nsite, nd), np.int32)
        self._check(lib().moed_basis(self.h, nu, nd, up.ctypes.data_as(_ip), dn.ctypes.data_as(_ip)))
        return up, dn

    def set_bath(self, ek, vk):
        self._ek = _f64(ek).reshape(self.nsite).copy()
        self._vk = np.asarray(vk, np.float64).reshape(self.norb, self.nbath, order="F").reshape(-1, order="F").copy()
        self._check(lib().moed_set_bath(self.h, _d(self._ek), _d(self._vk)))

    def multiply_H(self, nu, nd, x):
        x = _f64(x)
        if x.size != self.dim(nu, nd):
            raise ValueError("vector length does not match the sector dimension")
        y = np.empty_like(x)
        self._check(lib().moed_hv(self.h, nu, nd, _d(x), _d(y)))
        return y

    def multiply_H_device(self, nu, nd, x_ptr, y_ptr, stream=None):
        self._check(lib().moed_hv_device(self.h, nu, nd, ct.c_void_p(x_ptr), ct.c_void_p(y_ptr),
                                         ct.c_void_p(stream) if stream else None))

    def apply_c(self, nu, nd, vec, pm, iorb, ispin):
        vec = _f64(vec)
        du, dd = (1, 0) if ispin == 1 else (0, 1)
        s = 1 if pm == 1 else -1
        nuo, ndo = nu + s * du, nd + s * dd
        out = np.empty(self.dim(nuo, ndo))
        self._check(lib().moed_apply_c(self.h, pm, iorb, ispin, nu, nd, _d(vec), _d(out)))
        return (nuo, ndo), out

    def lanczos_iteration(self, nu, nd, vec, nstep=None):
        nstep = self.nstep if nstep is None else nstep
        vec = _f64(vec)
        a, b = np.zeros(nstep + 1), np.zeros(nstep + 1)
        nsc = ct.c_int32()
        self._check(lib().moed_lanczos(self.h, nu, nd, _d(vec), nstep, ct.byref(nsc), _d(a), _d(b)))
        return nsc.value, a, b

    # ------------------------------------------------------------ solver / Green's function
    def ed_solve(self):
        n = self.nev
        val, prob = np.zeros(n), np.zeros(n)
        sec, lev = np.zeros(n, np.int32), np.zeros(n, np.int32)
        nc = ct.c_int32()
        self._check(lib().moed_solve(self.h, ct.byref(nc), _d(val), _d(prob), sec.ctypes.data_as(_ip),
                                     lev.ctypes.data_as(_ip)))
        k = nc.value
        return [(val[i], prob[i], int(sec[i]), int(lev[i])) for i in range(k)]

    def sector_eigenvalues(self, isector):
        ev = np.zeros(self.nev)
        self._check(lib().moed_sector_eigenvalues(self.h, isector, _d(ev)))
        return ev

    def eigvec(self, isector, ilevel):
        nu, nd = self.sectors[isector - 1]
        v = np.empty(self.dim(nu, nd))
        self._check(lib().moed_eigvec(self.h, isector, ilevel, _d(v)))
        return v

    def ed_calc_green_ftn(self, omega):
        omega = _f64(omega)
        G = np.zeros(2 * self.norb * len(omega))
        self._check(lib().moed_green(self.h, len(omega), _d(omega), _d(G)))
        return (G[0::2] + 1j * G[1::2]).reshape(self.norb, len(omega), order="F")

    def green_records(self):
        n = ct.c_int32()
        self._check(lib().moed_green_nrecords(self.h, ct.byref(n)))
        out = []
        for r in range(n.value):
            iev, iorb, ispin, even, nsp, nsn = (ct.c_int32() for _ in range(6))
            E, p = ct.c_double(), ct.c_double()
            ap, bp, an, bn = (np.zeros(self.nstep + 1) for _ in range(4))
            self._check(lib().moed_green_record(self.h, r, ct.byref(iev), ct.byref(iorb), ct.byref(ispin),
                                                ct.byref(E), ct.byref(p), ct.byref(even), ct.byref(nsp),
                                                ct.byref(nsn), _d(ap), _d(bp), _d(an), _d(bn)))
            out.append(dict(iev=iev.value, iorb=iorb.value, ispin=ispin.value, E=E.value, p=p.value,
                            even=bool(even.value), nsp=nsp.value, nsn=nsn.value, ap=ap, bp=bp, an=an, bn=bn))
        return out

    def green_cf(self, sign, E, nsc, a, b, omega):
        a, b, omega = _f64(a), _f64(b), _f64(omega)
        g = np.empty(2 * len(omega))
        self._check(lib().moed_green_cf(self.h, sign, E, nsc, _d(a), _d(b), len(omega), _d(omega), _d(g)))
        return g[0::2] + 1j * g[1::2]

    def green_cf_adjoint(self, sign, E, nsc, a, b, omega, gbar):
        a, b, omega = _f64(a), _f64(b), _f64(omega)
        gb = np.empty(2 * len(omega))
        gb[0::2], gb[1::2] = np.real(gbar), np.imag(gbar)
        ga, gbb = np.zeros(nsc + 1), np.zeros(nsc + 1)
        self._check(lib().moed_green_cf_adjoint(self.h, sign, E, nsc, _d(a), _d(b), len(omega), _d(omega), _d(gb),
                                                _d(ga), _d(gbb)))
        return ga, gbb

    # ------------------------------------------------------------ bath fit
    def calc_dev(self, ek, vk, ef, omega):
        omega = _f64(omega)
        vkf = np.asarray(vk, np.float64).reshape(self.norb, self.nbath, order="F").reshape(-1, order="F").copy()
        D = np.empty(2 * self.norb * len(omega))
        self._check(lib().moed_calc_dev(self.h, _d(_f64(ek)), _d(vkf), _d(_f64(ef)), len(omega), _d(omega), _d(D)))
        return (D[0::2] + 1j * D[1::2]).reshape(self.norb, len(omega), order="F")

    @property
    def nx(self):
        return self.nbath * (1 + self.norb) + self.norb

    def func_dfunc(self, x, omega, Dev):
        """batched FUNC and dfunc: x (nfit, nx) -> f (nfit,), g (nfit, nx)."""
        x = np.atleast_2d(_f64(x))
        nfit = x.shape[0]
        f, g = np.zeros(nfit), np.zeros((nfit, self.nx))
        D = _cplx_flat(Dev)
        omega = _f64(omega)
        self._check(lib().moed_chi2_grad(self.h, nfit, self.nx, _d(x), len(omega), _d(omega), _d(D), _d(f), _d(g)))
        return f, g

    def minimization(self, x, omega, Dev, ftol=1.0e-7):
        """batched minimization (func.f90:144 -> frprmn): returns (x, fret, iter, status)."""
        x = np.atleast_2d(_f64(x)).copy()
        nfit = x.shape[0]
        fret = np.zeros(nfit)
        it, st = np.zeros(nfit, np.int32), np.zeros(nfit, np.int32)
        D = _cplx_flat(Dev)
        omega = _f64(omega)
        self._check(lib().moed_minimize(self.h, nfit, self.nx, _d(x), len(omega), _d(omega), _d(D), ftol, _d(fret),
                                        it.ctypes.data_as(_ip), st.ctypes.data_as(_ip)))
        return x, fret, it, st

    # ------------------------------------------------------------ on-disk records

    def export_green(self, path):
        """apbpanbn.dat of the last ed_calc_green_ftn (ed_green.F90:112-129)"""
        self._check(lib().moed_export_green(self.h, os.fsencode(path)))

    def import_green(self, path):
        """read an apbpanbn.dat file into the solver's Green's-function records"""
        self._check(lib().moed_import_green(self.h, os.fsencode(path)))

    def export_eigvecs(self, directory):
        """the eigenvector files ed_solve writes (ed_solver.F90:58-61), every sector and level"""
        for isec in range(1, len(self.sectors) + 1):
            ev = self.sector_eigenvalues(isec)
            for lev in range(1, int(np.sum(ev < 1e300)) + 1):  # levels with an eigenvector
                export_eigvec(directory, isec, lev, self.eigvec(isec, lev))

    # ------------------------------------------------------------ real-axis spectral function

    def dos_diag(self, omega_real, small, ev, ap, bp, an, bn):
        """dos_diag (dos.f90:65-106) for one record: A(w) = -Im(G+ + G-)/pi."""
        om = _f64(omega_real)
        ap, bp, an, bn = _f64(ap), _f64(bp), _f64(an), _f64(bn)
        out = np.empty(len(om))
        self._check(lib().moed_dos_diag(self.h, len(om), _d(om), float(small), float(ev), len(ap) - 1, len(an) - 1,
                                        _d(ap), _d(bp), _d(an), _d(bn), _d(out)))
        return out

    def dos(self, omega_real, small=0.05, weighted=False):
        """dos (dos.f90:1-63) over the records of the last ed_calc_green_ftn: Aff (norb, nw)."""
        om = _f64(omega_real)
        out = np.empty(self.norb * len(om))
        self._check(lib().moed_dos(self.h, len(om), _d(om), float(small), int(bool(weighted)), _d(out)))
        return out.reshape(self.norb, len(om), order="F")

    # ------------------------------------------------------------ lattice step
    def delta_new(self, Gr, Dev, hk, rmu):
        """ed_delta_new + n_from_gksum on the device: returns (Gksum, new Dev, nocc)"""
        Gr = np.asarray(Gr, np.complex128)
        norb, nw = Gr.shape
        hk = _f64(np.asarray(hk).reshape(-1))
        nq = len(hk) // norb
        d = _cplx_flat(Dev)
        gks = np.empty(2 * norb * nw)
        nocc = np.empty(norb)
        self._check(lib().moed_delta_new(self.h, nw, nq, float(rmu), _d(hk), _d(_cplx_flat(Gr)), _d(d), _d(gks),
                                         _d(nocc)))
        tocx = lambda v: (v[0::2] + 1j * v[1::2]).reshape(norb, nw, order="F")
        return tocx(gks), tocx(d), nocc

    def set_hv_shape(self, threads=0, strip=0):
        self._check(lib().moed_set_hv_shape(self.h, threads, strip))

    def set_eig_chunk(self, vectors=0):
        self._check(lib().moed_set_eig_chunk(self.h, vectors))

    def set_persistent(self, max_dim=-1):
        """largest sector that the batched solve runs in its persistent work-queue launch
        (-1: default, 0: none)"""
        self._check(lib().moed_set_persistent(self.h, int(max_dim)))

    def set_fit_launch(self, iterations=0):
        """CG iterations per fit and kernel launch inside minimization (0: default)"""
        self._check(lib().moed_set_fit_launch(self.h, iterations))

    def fit_nfev(self, nfit):
        """FUNC evaluations of each fit of the last minimization"""
        n = np.zeros(nfit, np.int32)
        self._check(lib().moed_fit_nfev(self.h, nfit, n.ctypes.data_as(_ip)))
        return n

    @property
    def stream(self):
        return lib().moed_stream(self.h)
