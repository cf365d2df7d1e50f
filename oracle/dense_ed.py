"""Brute-force second quantisation of the impurity Hamiltonian -- TEST INFRASTRUCTURE ONLY.

An implementation independent of the reference's operator code (ed_operators.F90) and
of oracle.c: the Hamiltonian is written as a list of fermion operator strings and each
string is applied to all basis states at once with vectorised Jordan-Wigner signs, for
any chosen mode ordering.  Two orderings give the same spectrum only if both are
consistent second quantisations, so it pins the signs of the reference's spin-flip and
pair-exchange terms (ed_operators.F90:167-271) through eigenvalues, and pins matrix
elements directly in the reference ordering (bit = (spin-1)*Nsite + site-1,
ed_operators.F90:63-66).

Terms (the Kanamori form the reference implements, ed_hamiltonian.F90:110-187):
  sum_{site,s} e_site n_{site s}
  + U sum_o n_{o up} n_{o dn} + sum_{o<o'} sum_{s,s'} (s==s' ? U-3J : U-2J) n_{o s} n_{o' s'}
  + sum_{s,o,b} V_{ob} (c+_{o s} c_{b s} + c+_{b s} c_{o s})
  + sum_{s, i!=j} t_ij c+_{i s} c_{j s}                  (extension, no reference term)
  - J sum_{i!=j} c+_{i up} c_{i dn} c+_{j dn} c_{j up}      (spin flip)
  - J sum_{i!=j} c+_{i up} c+_{i dn} c_{j up} c_{j dn}      (pair exchange, as applied)
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp

UP, DN = 0, 1


def strings(ns, n):
    """all ns-bit integers with popcount n, ascending (the reference's order)."""
    v = np.arange(1 << ns, dtype=np.int64)
    pc = np.zeros_like(v)
    for k in range(ns):
        pc += (v >> k) & 1
    return v[pc == n]


def popcount(a):
    a = a.astype(np.int64)
    c = np.zeros_like(a)
    while np.any(a):
        c += a & 1
        a = a >> 1
    return c


class Sector:
    """basis of sector (nu, nd): state = up + 2^ns * dn, index i = iu + nup * id."""

    def __init__(self, ns, nu, nd):
        self.ns, self.nu, self.nd = ns, nu, nd
        self.up = strings(ns, nu)
        self.dn = strings(ns, nd)
        self.nup, self.ndn = len(self.up), len(self.dn)
        self.dim = self.nup * self.ndn
        iu, idn = np.meshgrid(np.arange(self.nup), np.arange(self.ndn), indexing="xy")
        self.states = (self.up[iu.reshape(-1)] + (self.dn[idn.reshape(-1)] << ns)).astype(np.int64)
        self.rank_up = {int(s): k for k, s in enumerate(self.up)}
        self.rank_dn = {int(s): k for k, s in enumerate(self.dn)}

    def index(self, states):
        mask = (1 << self.ns) - 1
        lu = np.full(1 << self.ns, -1, np.int64)
        lu[self.up] = np.arange(self.nup)
        ld = np.full(1 << self.ns, -1, np.int64)
        ld[self.dn] = np.arange(self.ndn)
        return lu[states & mask] + self.nup * ld[states >> self.ns]


def ref_mode(ns, site, spin):
    """reference ordering: all up modes, then all down modes (site 0-based)."""
    return spin * ns + site


def interleaved_mode(ns, site, spin):
    """alternative ordering (site-major: 1up,1dn,2up,2dn,...), as in the notebook's SNEG model."""
    return 2 * site + spin


def _to_modes(states, ns, mode):
    out = np.zeros_like(states)
    for spin in (UP, DN):
        for site in range(ns):
            bit = (states >> (spin * ns + site)) & 1
            out |= bit << mode(ns, site, spin)
    return out


def _from_modes(mstates, ns, mode):
    out = np.zeros_like(mstates)
    for spin in (UP, DN):
        for site in range(ns):
            bit = (mstates >> mode(ns, site, spin)) & 1
            out |= bit << (spin * ns + site)
    return out


def apply_string(states, ops, ns, mode):
    """apply the operator product ops[0] ops[1] ... ops[-1] (rightmost acts first);
    each op = (dagger, site, spin).  Returns (target states in reference encoding, sign, valid)."""
    s = _to_modes(states, ns, mode)
    sign = np.ones(len(s), np.int64)
    valid = np.ones(len(s), bool)
    for dag, site, spin in reversed(ops):
        k = mode(ns, site, spin)
        occ = (s >> k) & 1
        valid &= (occ == 0) if dag else (occ == 1)
        par = popcount(s & ((1 << k) - 1)) & 1
        sign *= 1 - 2 * par
        s = s ^ (1 << k)
    return _from_modes(s, ns, mode), sign, valid


def terms(norb, nbath, U, J, ek, vk, tij=None):
    """list of (coeff, ops) for the Hamiltonian; vk is (norb, nbath); sites 0-based."""
    ns = norb + nbath
    T = []
    for spin in (UP, DN):
        for site in range(ns):
            T.append((ek[site], [(1, site, spin), (0, site, spin)]))
    for o in range(norb):
        T.append((U, [(1, o, UP), (0, o, UP), (1, o, DN), (0, o, DN)]))
    for o in range(norb):
        for p in range(o + 1, norb):
            for s1 in (UP, DN):
                for s2 in (UP, DN):
                    c = U - 3 * J if s1 == s2 else U - 2 * J
                    T.append((c, [(1, o, s1), (0, o, s1), (1, p, s2), (0, p, s2)]))
    for spin in (UP, DN):
        for o in range(norb):
            for b in range(nbath):
                v = vk[o, b]
                if v == 0.0:
                    continue
                T.append((v, [(1, o, spin), (0, norb + b, spin)]))
                T.append((v, [(1, norb + b, spin), (0, o, spin)]))
    if tij is not None:
        for spin in (UP, DN):
            for i in range(norb):
                for j in range(norb):
                    if i != j and tij[i, j] != 0.0:
                        T.append((tij[i, j], [(1, i, spin), (0, j, spin)]))
    for i in range(norb):
        for j in range(norb):
            if i == j:
                continue
            T.append((-J, [(1, i, UP), (0, i, DN), (1, j, DN), (0, j, UP)]))
            T.append((-J, [(1, i, UP), (1, i, DN), (0, j, UP), (0, j, DN)]))
    return T


def hamiltonian(sector, norb, nbath, U, J, ek, vk, tij=None, mode=ref_mode):
    """sparse H of the sector (CSR), H[i, j] = <i|H|j>."""
    ns = sector.ns
    rows, cols, vals = [], [], []
    src = np.arange(sector.dim)
    for c, ops in terms(norb, nbath, U, J, np.asarray(ek), np.asarray(vk).reshape(norb, nbath), tij):
        t, sg, ok = apply_string(sector.states, ops, ns, mode)
        if not ok.any():
            continue
        rows.append(sector.index(t[ok]))
        cols.append(src[ok])
        vals.append(c * sg[ok])
    if not rows:
        return sp.csr_matrix((sector.dim, sector.dim))
    r, cl, v = np.concatenate(rows), np.concatenate(cols), np.concatenate(vals).astype(np.float64)
    return sp.csr_matrix((v, (r, cl)), shape=(sector.dim, sector.dim))


def c_matrix(src, dst, site, spin, dagger, mode=ref_mode):
    """<dst|c(+)_{site spin}|src> as a sparse (dst.dim, src.dim) matrix."""
    t, sg, ok = apply_string(src.states, [(1 if dagger else 0, site, spin)], src.ns, mode)
    return sp.csr_matrix((sg[ok].astype(np.float64), (dst.index(t[ok]), np.arange(src.dim)[ok])),
                         shape=(dst.dim, src.dim))


def lehmann_green(norb, nbath, U, J, ek, vk, nu, nd, orb, spin, omega, tij=None, mode=ref_mode):
    """zero-temperature G_{orb spin}(i w) of the ground state of sector (nu, nd) by the Lehmann sum
    over the full spectra of the N+1 and N-1 sectors (dense eigh).  Returns (E0, G)."""
    ns = norb + nbath
    s0 = Sector(ns, nu, nd)
    H0 = hamiltonian(s0, norb, nbath, U, J, ek, vk, tij, mode).toarray()
    e0, v0 = np.linalg.eigh(H0)
    E0, psi = e0[0], v0[:, 0]
    du, dd = (1, 0) if spin == UP else (0, 1)
    G = np.zeros(len(omega), complex)
    z = 1j * np.asarray(omega)
    if nu + du <= ns and nd + dd <= ns:
        sp_ = Sector(ns, nu + du, nd + dd)
        Hp = hamiltonian(sp_, norb, nbath, U, J, ek, vk, tij, mode).toarray()
        ep, vp = np.linalg.eigh(Hp)
        amp = vp.T @ (c_matrix(s0, sp_, orb, spin, True, mode) @ psi)
        G += np.sum((amp ** 2)[:, None] / (z[None, :] - (ep - E0)[:, None]), axis=0)
    if nu - du >= 0 and nd - dd >= 0:
        sm = Sector(ns, nu - du, nd - dd)
        Hm = hamiltonian(sm, norb, nbath, U, J, ek, vk, tij, mode).toarray()
        em, vm = np.linalg.eigh(Hm)
        amp = vm.T @ (c_matrix(s0, sm, orb, spin, False, mode) @ psi)
        G += np.sum((amp ** 2)[:, None] / (z[None, :] + (em - E0)[:, None]), axis=0)
    return E0, G
