"""Draw the five workload parameter sets and write them as hex-float JSON (configs.json).

Generator: numpy.random.default_rng(seed) (PCG64), seeds 1001..1005.  Draw order per
config: bath energies e_k (one per bath site, ascending site index), then couplings V_k
(one per bath site), then impurity hoppings t_ij (upper triangle, row-major).  Bath site
k (0-based) belongs to orbital k // Nb; V is block-diagonal, so vk(o, k) = 0 for the
other orbitals.  Impurity levels ek(1:Norb) = e_imp - mu with e_imp = 0 and
mu = U/2 + (U' - J/2)(Norb - 1), U' = U - 2J.

Run:  python tests/golden/make_configs.py   (rewrites tests/golden/configs.json)
"""
import json
import os

import numpy as np

SPECS = [
    dict(name="ns8_plain", seed=1001, norb=2, nb=3, U=4.0, J=0.5, erange=(-2.0, 2.0), vrange=(0.2, 0.8),
         sectors=[(8, 4)], nstep=100, nw=512, beta=100.0, nev=1, tij=0.0),
    dict(name="ns12_plain", seed=1002, norb=2, nb=5, U=4.0, J=0.5, erange=(-3.0, 3.0), vrange=(0.1, 1.0),
         sectors=[(12, 6)], nstep=150, nw=1024, beta=100.0, nev=1, tij=0.0),
    dict(name="ns14_roofline", seed=1003, norb=2, nb=6, U=6.0, J=0.8, erange=(-3.0, 3.0), vrange=(0.1, 1.0),
         sectors=[(14, 7)], nstep=200, nw=2048, beta=200.0, nev=1, tij=0.0),
    dict(name="ns10_batched", seed=1004, norb=2, nb=4, U=4.0, J=0.5, erange=(-3.0, 3.0), vrange=(0.1, 1.0),
         sectors="all", nstep=100, nw=1024, beta=50.0, nev=8, tij=0.0),
    dict(name="ns15_fit", seed=1005, norb=3, nb=4, U=4.0, J=0.6, erange=(-3.0, 3.0), vrange=(0.1, 1.0),
         sectors=[(15, 7)], nstep=150, nw=4096, beta=200.0, nev=1, tij=0.3, nfit=256, ndmft=8),
]


def build(spec):
    rng = np.random.default_rng(spec["seed"])
    norb, nb = spec["norb"], spec["nb"]
    nbath = norb * nb
    ns = norb + nbath
    U, J = spec["U"], spec["J"]
    Up = U - 2 * J
    mu = U / 2 + (Up - J / 2) * (norb - 1)
    eb = rng.uniform(*spec["erange"], size=nbath)
    vb = rng.uniform(*spec["vrange"], size=nbath)
    tij = np.zeros((norb, norb))
    if spec["tij"]:
        for i in range(norb):
            for j in range(i + 1, norb):
                tij[i, j] = tij[j, i] = rng.uniform(-spec["tij"], spec["tij"])
    ek = np.concatenate([np.full(norb, 0.0 - mu), eb])
    vk = np.zeros((norb, nbath))
    for k in range(nbath):
        vk[k // nb, k] = vb[k]
    if spec["sectors"] == "all":
        sectors = [(nu + nd, nu) for nu in range(ns + 1) for nd in range(ns + 1)]
    else:
        sectors = spec["sectors"]
    out = dict(name=spec["name"], seed=spec["seed"], norb=norb, nbath=nbath, nb_per_orb=nb, nsite=ns,
               U=U.hex(), J=J.hex(), mu=float(mu).hex(), beta=float(spec["beta"]).hex(),
               ek=[float(v).hex() for v in ek], vk_colmajor=[float(v).hex() for v in vk.reshape(-1, order="F")],
               tij_colmajor=[float(v).hex() for v in tij.reshape(-1, order="F")] if spec["tij"] else None,
               sectors_nelec_nup=[list(s) for s in sectors], nstep=spec["nstep"], nw=spec["nw"],
               nev=spec["nev"])
    for k in ("nfit", "ndmft"):
        if k in spec:
            out[k] = spec[k]
    return out


def main():
    cfgs = [build(s) for s in SPECS]
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "configs.json")
    with open(path, "w") as f:
        json.dump(cfgs, f, indent=1)
    print("wrote", path)


if __name__ == "__main__":
    main()
