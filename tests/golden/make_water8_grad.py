"""Golden fixture: RHF gradient data of the (H2O)8 / 6-31G** cluster from the CPU oracle.

Geometry: hfx.geometry.water_cluster_grid((2, 2, 2), 5.6 Bohr spacing, +-0.4 Bohr jitter,
seed 1002), 200 Cartesian functions.  SCF: Schwarz and density thresholds 1e-12, DIIS 6,
energy tolerance 1e-10, density tolerance 1e-8.  Writes water8_grad.json next to this
script: the SCF energy, the oracle's analytic gradient (all 72 coordinates) and central
finite differences (step 1e-4 Bohr) of the oracle's SCF energies for the coordinates in
FD_COORDS, saved after every coordinate.  Run from the repository root:
    python tests/golden/make_water8_grad.py
"""
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "hartree-fock_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import numpy as np  # noqa: E402

import pyoracle  # noqa: E402
from hfx import Basis, geometry, lib  # noqa: E402

H = 1e-4
FD_COORDS = [(0, 0), (0, 1), (0, 2), (1, 0), (2, 1), (4, 2)]
SCF = dict(energy_tol=1e-10, density_tol=1e-8, schwarz_thresh=1e-12, density_thresh=1e-12, max_iter=100)
OUT = os.path.join(HERE, "water8_grad.json")


def main():
    O = pyoracle.load(fast=True)
    Z, xyz = geometry.water_cluster_grid(shape=(2, 2, 2), spacing=5.6, jitter=0.4, seed=1002)
    b = Basis("6-31g**", Z, xyz)

    def scf(x):
        r = lib.scf_run(O, Basis("6-31g**", Z, x), options=O.default_options(**SCF))
        assert r["converged"] == 1
        return r

    t = time.time()
    r = scf(xyz)
    out = dict(system="(H2O)8/6-31G**", n_aos=b.n_aos, step=H, scf=SCF, e_total=r["e_total"],
               iterations=r["iterations"], scf_seconds=time.time() - t)
    nocc = sum(b.Z) // 2
    P = 2 * r["D_alpha"]
    W = lib.energy_weighted_density(r["C_alpha"], r["eps_alpha"], nocc)
    t = time.time()
    g, parts, st = lib.rhf_gradient(lib.Engine(O, b), b, P, W, schwarz_thresh=1e-12, density_thresh=1e-12)
    out.update(grad=g.tolist(), parts=parts.tolist(), grad_seconds=time.time() - t, grad_quartets=st["n_quartets"])
    out["fd"] = []
    json.dump(out, open(OUT, "w"), indent=1)
    print("E", out["e_total"], "scf s", out["scf_seconds"], "grad s", out["grad_seconds"], flush=True)
    for (A, x) in FD_COORDS:
        xp, xm = xyz.copy(), xyz.copy()
        xp[A, x] += H
        xm[A, x] -= H
        fd = (scf(xp)["e_total"] - scf(xm)["e_total"]) / (2 * H)
        out["fd"].append([A, x, fd])
        json.dump(out, open(OUT, "w"), indent=1)
        print(A, x, fd, g[A, x], fd - g[A, x], flush=True)


if __name__ == "__main__":
    main()
