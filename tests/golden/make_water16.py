"""Golden fixture: FP64 SCF of the (H2O)16 / 6-31G** cluster on the CPU oracle.

Geometry: hfx.geometry.water_cluster_grid((4, 2, 2), 5.6 Bohr spacing, +-0.4 Bohr jitter,
seed 1002); Schwarz 1e-10, density threshold 1e-10, DIIS 6, energy tolerance 1e-8.
Writes water16_scf.json (energy, iteration count, per-iteration trace, norms of F and D)
next to this script.  Run from the repository root:  python tests/golden/make_water16.py
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


def main():
    O = pyoracle.load()
    Z, xyz = geometry.water_cluster_grid(shape=(4, 2, 2), spacing=5.6, jitter=0.4, seed=1002)
    b = Basis("6-31g**", Z, xyz)
    opts = O.default_options(energy_tol=1e-8, schwarz_thresh=1e-10, density_thresh=1e-10, diis_size=6)
    t = time.time()
    r = lib.scf_run(O, b, options=opts)
    out = dict(system="(H2O)16/6-31G**", n_aos=b.n_aos, e_total=r["e_total"], e_nuclear=r["e_nuclear"],
               iterations=r["iterations"], converged=r["converged"], trace=r["trace"].tolist(),
               F_fro=float(np.linalg.norm(r["F_alpha"])), D_fro=float(np.linalg.norm(r["D_alpha"])),
               F_diag=np.diag(r["F_alpha"]).tolist(), eps=r["eps_alpha"].tolist(), seconds=time.time() - t)
    with open(os.path.join(HERE, "water16_scf.json"), "w") as f:
        json.dump(out, f, indent=1)
    print(out["e_total"], out["iterations"], out["seconds"])


if __name__ == "__main__":
    main()
