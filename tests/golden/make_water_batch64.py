"""Golden fixture: the 64-molecule batch (seeded water clusters, n ~ U{1..8}, 6-31G, 13-104
functions each; hfx.geometry.water_batch(64, seed 1004)) as 64 single-molecule FP64 SCFs on the
CPU oracle, converged tightly (Schwarz and density thresholds 1e-12, energy tolerance 1e-10,
DIIS 6), so that the batched GPU SCF can be compared molecule by molecule at 1e-9 Eh without
the stopping iteration mattering.  Writes water_batch64.json next to this script.
Run from the repository root:  python tests/golden/make_water_batch64.py
"""
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "hartree-fock_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

OPTIONS = dict(energy_tol=1e-10, density_tol=1e-7, schwarz_thresh=1e-12, density_thresh=1e-12, diis_size=6,
               max_iter=100)


def molecules():
    from hfx import Basis, geometry
    return [Basis("6-31g", Z, x) for Z, x in geometry.water_batch(n_mol=64, seed=1004)]


def main():
    import pyoracle
    from hfx import lib
    O = pyoracle.load()
    mols = molecules()
    t = time.time()
    rows = []
    for m in mols:
        r = lib.scf_run(O, m, options=O.default_options(**OPTIONS), outputs=False)
        rows.append(dict(n_aos=m.n_aos, n_atoms=len(m.Z), e_total=r["e_total"], iterations=r["iterations"],
                         converged=r["converged"], last_dE=r["last_dE"]))
    out = dict(system="64 water clusters / 6-31G (seed 1004)", options=OPTIONS, molecules=rows,
               seconds=time.time() - t)
    with open(os.path.join(HERE, "water_batch64.json"), "w") as f:
        json.dump(out, f, indent=1)
    print(len(rows), "molecules,", sum(r["converged"] for r in rows), "converged, %.1fs" % out["seconds"])


if __name__ == "__main__":
    main()
