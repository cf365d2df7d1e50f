"""The Fortran drop-in (module moed_ed of mo-dmft_amd/fortran/moed_iface.f90) driven by a program in
the shape of the reference's main.f90 (mo-dmft_amd/fortran/drop_in_check.f90): every result -- H x,
c+ x, the Lanczos coefficients, the kept states, Gr, the lattice step, the fitted bath and the
spectral function -- equals the same call made through the C ABI, bit for bit."""
import os
import struct
import subprocess

import numpy as np
import pytest

import moed
from moed import EDSolver, dmft
from moed.model import config, omega_grid, start_vector

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "mo-dmft_amd", "fortran", "drop_in_check")


class _Reader:
    def __init__(self, raw):
        self.raw, self.pos = raw, 0

    def f8(self, n):
        v = np.frombuffer(self.raw, "<f8", n, self.pos).copy()
        self.pos += 8 * n
        return v

    def i4(self, n=1):
        v = np.frombuffer(self.raw, "<i4", n, self.pos).copy()
        self.pos += 4 * n
        return v

    def c16(self, norb, nw):
        v = self.f8(2 * norb * nw)
        return (v[0::2] + 1j * v[1::2]).reshape(norb, nw, order="F")


def test_fortran_drop_in_matches_the_c_abi(tmp_path):
    if not os.path.exists(EXE):
        pytest.skip("the Fortran check program is not built (no amdflang when the library was built)")
    c = config("ns8_plain")
    nu, nd = c.sectors[0]
    norb, nbath, nq = c.norb, c.nbath, 100
    omr = (np.arange(1, 201) - 100) * 0.01
    ek, vk = np.asarray(c.ek, np.float64), np.asarray(c.vk, np.float64).reshape(norb, nbath)
    with EDSolver.from_config(c) as s:
        D = s.dim(nu, nd)
        x = start_vector(5, D)
        fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
        with open(fin, "wb") as f:
            f.write(struct.pack("<10i", norb, nbath, c.nstep, c.nev, c.nw, len(c.sectors), nu, nd, nq, len(omr)))
            f.write(struct.pack("<4d", c.U, c.J, c.beta, c.mu))
            f.write(ek.astype("<f8").tobytes() + vk.reshape(-1, order="F").astype("<f8").tobytes())
            f.write(np.array([a + b for a, b in c.sectors], "<i4").tobytes())
            f.write(np.array([a for a, b in c.sectors], "<i4").tobytes())
            f.write(x.astype("<f8").tobytes() + omr.astype("<f8").tobytes())
        print("running", EXE, flush=True)
        r = subprocess.run([EXE, str(fin), str(fout)], capture_output=True, text=True, timeout=240)
        assert r.returncode == 0, r.stdout + r.stderr

        # the same calls through the C ABI, in the program's order
        om = omega_grid(c.nw, c.beta)
        ef = ek[:norb].copy()
        y = s.multiply_H(nu, nd, x)
        (nuo, ndo), vout = s.apply_c(nu, nd, x, 1, 1, 1)
        Dout = s.dim(nuo, ndo)
        nsc, a, b = s.lanczos_iteration(nu, nd, x)
        kept = s.ed_solve()
        G = s.ed_calc_green_ftn(om)
        Dev = s.calc_dev(ek, vk, ef, om)
        gks, Dev, nocc = s.delta_new(G, Dev, moed.band_tb(norb, nq), c.mu)
        X, fret, _, st = s.minimization(dmft.ev_to_x(ek, vk, ef, norb, nbath)[None], om, Dev)
        assert st[0] == 0
        aff = s.dos(omr, 0.05)

    o = _Reader(fout.read_bytes())
    assert np.array_equal(o.f8(D), y)
    assert np.array_equal(o.f8(Dout), vout)
    assert o.i4()[0] == nsc
    assert np.array_equal(o.f8(c.nstep + 1), a) and np.array_equal(o.f8(c.nstep + 1), b)
    nev_calc = o.i4()[0]
    assert nev_calc == len(kept)
    for E, p, isec, ilev in kept:
        assert tuple(o.f8(2)) == (E, p) and tuple(o.i4(2)) == (isec, ilev)
    assert np.array_equal(o.c16(norb, c.nw), G)
    assert np.array_equal(o.c16(norb, c.nw), Dev)
    assert np.array_equal(o.c16(norb, c.nw), gks)
    assert np.array_equal(o.f8(norb), nocc)
    assert np.array_equal(o.f8(len(X[0])), X[0])
    assert o.f8(1)[0] == fret[0]
    assert np.array_equal(o.f8(norb * len(omr)).reshape(norb, len(omr), order="F"), aff)
    assert o.pos == len(o.raw)
