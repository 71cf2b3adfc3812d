"""Golden fixture: the C20H42 / cc-pVDZ (expanded, 510 Cartesian functions, 246 shells) workload.

Geometry: hfx.geometry.alkane_chain(20 carbons, seed 1003).  Two stages:

  python tests/golden/make_c20h42.py --orbitals gpurun_out/c20h42_orbitals.npz
      (on a GPU) runs the engine's FP64 SCF (Schwarz 1e-10, density 1e-10, DIIS 6, energy
      tolerance 1e-8) and stores its occupied orbitals C_occ (510 x 81) and energy.  These are
      the fixture's *input*: a converged, realistic density.  (The oracle's own SCF of this
      molecule takes hours on 8 cores, one full build about ten minutes.)
  python tests/golden/make_c20h42.py --from gpurun_out/c20h42_orbitals.npz
      (on the CPU, where the oracle is built) computes the *expected outputs* with the
      deterministic oracle: its Schwarz table Q, and its full direct build at the fixed density
      D = sum_k C_k C_k^T (dD_tot = 2D, dD_alpha = D, Schwarz and density thresholds 1e-10, that Q):
      survivor count, list hashes, class counts, E2 = tr(D (J - K)), the electronic energy
      2 tr(D h) + E2 with the oracle's h = T + V, diag(J), diag(K), Frobenius norms and 600
      sampled elements of J and K.

Writes c20h42_fixture.npz (C_occ, lower triangle of Q) and c20h42_fixture.json next to this
script.  D is formed by density_from_orbitals() (a fixed-order sum of outer products with IEEE
elementwise operations: no BLAS, no fused operations), so any host rebuilds it bit for bit.
"""
import argparse
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "hartree-fock_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import numpy as np  # noqa: E402

N_SAMPLES = 600
SCF_OPTIONS = dict(energy_tol=1e-8, schwarz_thresh=1e-10, density_thresh=1e-10, diis_size=6, max_iter=60)


def density_from_orbitals(C):
    """D = sum_k C[:, k] C[:, k]^T accumulated in column order with separate IEEE multiply and add,
    so every host computes the same bits."""
    n = C.shape[0]
    D = np.zeros((n, n))
    for k in range(C.shape[1]):
        D += np.multiply.outer(C[:, k], C[:, k])
    return D


def workload():
    from hfx import Basis, geometry
    Z, xyz = geometry.alkane_chain(n_carbon=20, seed=1003)
    return Basis("cc-pvdz-expanded", Z, xyz)


def tril_to_full(t, n):
    M = np.zeros((n, n))
    M[np.tril_indices(n)] = t
    return M + np.tril(M, -1).T


def load_fixture():
    """(basis, C_occ, D, Q, reference dict) of the committed fixture."""
    with open(os.path.join(HERE, "c20h42_fixture.json")) as f:
        ref = json.load(f)
    z = np.load(os.path.join(HERE, "c20h42_fixture.npz"))
    b = workload()
    C = z["C_occ"]
    return b, C, density_from_orbitals(C), tril_to_full(z["Q_tril"], b.n_shells), ref


def gpu_orbitals(path):
    from hfx import lib
    L = lib.load()
    b = workload()
    t = time.time()
    r = lib.scf_run(L, b, options=L.default_options(**SCF_OPTIONS))
    nocc = b.n_electrons() // 2
    np.savez(path, C_occ=np.ascontiguousarray(r["C_alpha"][:, :nocc]), e_total=r["e_total"],
             iterations=r["iterations"], converged=r["converged"])
    print("GPU SCF E=%.10f it=%d conv=%d %.1fs -> %s" % (r["e_total"], r["iterations"], r["converged"],
                                                        time.time() - t, path), flush=True)


def oracle_outputs(path):
    import pyoracle
    from hfx import lib
    src = np.load(path)
    assert int(src["converged"]) == 1
    C_occ = np.ascontiguousarray(src["C_occ"])
    b = workload()
    n, nS = b.n_aos, b.n_shells
    assert C_occ.shape == (n, b.n_electrons() // 2)
    O = pyoracle.load()
    e = lib.Engine(O, b)
    t = time.time()
    Q = e.schwarz()
    h = e.kinetic() + e.nuclear(b.Z, b.xyz)
    print("oracle Q, h: %.1fs" % (time.time() - t), flush=True)
    D = density_from_orbitals(C_occ)
    t = time.time()
    J, K, _, st = e.jk(2 * D, D, Q=Q, schwarz_thresh=1e-10, density_thresh=1e-10)
    dt = time.time() - t
    print("oracle build %.0fs, %d quartets" % (dt, st["n_quartets"]), flush=True)
    Zs = np.asarray(b.Z, float)
    R = np.linalg.norm(b.xyz[:, None, :] - b.xyz[None, :, :], axis=-1)
    iu = np.triu_indices(len(Zs), 1)
    e_nuc = float(np.sum(Zs[iu[0]] * Zs[iu[1]] / R[iu]))
    E2 = float(np.sum(D * (J - K)))
    rng = np.random.default_rng(510)
    ii = rng.integers(0, n, N_SAMPLES)
    jj = rng.integers(0, n, N_SAMPLES)
    out = dict(system="C20H42/cc-pVDZ expanded", n_aos=n, n_shells=nS, n_occ=int(C_occ.shape[1]),
               density_source="orbitals of the HIP engine's FP64 SCF (%s), E = %.10f, %d iterations" % (
                   "settings " + json.dumps(SCF_OPTIONS), float(src["e_total"]), int(src["iterations"])),
               gpu_scf_e_total=float(src["e_total"]), oracle_build_seconds=dt,
               n_quartets=st["n_quartets"], hash_sum=str(st["hash_sum"]), hash_xor=str(st["hash_xor"]),
               class_count=st["class_count"], model_gflop=st["model_gflop"], E2=E2,
               e_electronic=float(2 * np.sum(D * h)) + E2, e_nuclear=e_nuc,
               J_diag=np.diag(J).tolist(), K_diag=np.diag(K).tolist(),
               J_fro=float(np.linalg.norm(J)), K_fro=float(np.linalg.norm(K)),
               sample_i=ii.tolist(), sample_j=jj.tolist(), sample_J=J[ii, jj].tolist(), sample_K=K[ii, jj].tolist())
    np.savez_compressed(os.path.join(HERE, "c20h42_fixture.npz"), C_occ=C_occ, Q_tril=Q[np.tril_indices(nS)])
    with open(os.path.join(HERE, "c20h42_fixture.json"), "w") as f:
        json.dump(out, f, indent=1)
    print("E2 = %.12f, E_total(D) = %.10f" % (E2, out["e_electronic"] + e_nuc))


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    g = ap.add_mutually_exclusive_group(required=True)
    g.add_argument("--orbitals", help="GPU stage: write the SCF orbitals here")
    g.add_argument("--from", dest="src", help="CPU stage: oracle outputs at the density of these orbitals")
    a = ap.parse_args()
    if a.orbitals:
        gpu_orbitals(a.orbitals)
    else:
        oracle_outputs(a.src)
