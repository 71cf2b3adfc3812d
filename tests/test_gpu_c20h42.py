"""The headline workload, C20H42 / cc-pVDZ expanded (510 Cartesian functions), on the HIP engine
against the deterministic oracle's full build at a fixed converged density
(tests/golden/c20h42_fixture.*, written by tests/golden/make_c20h42.py).  At N = 510 the FP64
classes with d shells on both sides run without LDS exchange rows (the rows and the Boys slice
would exceed the 64 KB per block), so this also covers the direct-atomic exchange path."""
import os
import sys

import numpy as np
import pytest

from hfx import autotune, device, lib

pytestmark = pytest.mark.gpu

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden"))
import make_c20h42  # noqa: E402

SCF = make_c20h42.SCF_OPTIONS


@pytest.fixture(scope="module")
def c20(engine_lib):
    b, C, D, Q, ref = make_c20h42.load_fixture()
    assert b.n_aos == ref["n_aos"] == 510 and b.n_shells == ref["n_shells"] == 246
    return b, D, Q, ref, lib.Engine(engine_lib, b)


def _check_against_oracle(J, K, D, ref, tol_e2, tol_el):
    E2 = float(np.sum(D * (J - K)))
    assert abs(E2 - ref["E2"]) <= tol_e2, (E2, ref["E2"])
    i, j = np.asarray(ref["sample_i"]), np.asarray(ref["sample_j"])
    assert np.abs(J[i, j] - np.asarray(ref["sample_J"])).max() <= tol_el
    assert np.abs(K[i, j] - np.asarray(ref["sample_K"])).max() <= tol_el
    assert np.abs(np.diag(J) - np.asarray(ref["J_diag"])).max() <= tol_el
    assert np.abs(np.diag(K) - np.asarray(ref["K_diag"])).max() <= tol_el
    return E2


def test_c20h42_fp64_graph_build_matches_oracle(c20):
    """FP64 build, graph-replayed from device-resident buffers, with the oracle's Q: the same
    work list (count and both hashes), |dE2| <= 1e-9 Eh, sampled |dJ|, |dK| <= 1e-8."""
    b, D, Q, ref, g = c20
    n = b.n_aos
    print("c20h42 fp64 graph", flush=True)
    J, K, _, st = g.jk(2 * D, D, Q=Q)  # eager host-buffer build, with statistics
    assert st["n_quartets"] == ref["n_quartets"]
    assert (str(st["hash_sum"]), str(st["hash_xor"])) == (ref["hash_sum"], ref["hash_xor"])
    assert st["class_count"] == ref["class_count"]
    _check_against_oracle(J, K, D, ref, 1e-9, 1e-8)
    # the whole FP64 energy at the fixed density: 2 tr(D h) + E2 (+ E_nn) with the GPU's one-electron
    # matrices and build, against the oracle's
    h = g.kinetic() + g.nuclear(b.Z, b.xyz)
    e_el = float(2 * np.sum(D * h) + np.sum(D * (J - K)))
    assert abs(e_el - ref["e_electronic"]) <= 1e-9, (e_el, ref["e_electronic"])
    bufs = [device.DeviceArray.from_host(x) for x in (2 * D, D)] + [device.DeviceArray(n * n) for _ in range(2)]
    try:
        for rep in range(3):
            lib.jk_device(g, *bufs, Q=np.ascontiguousarray(Q.reshape(-1, order="F")),
                          flags=lib.HFX_JK_USE_GRAPH | lib.HFX_JK_NO_STATS)
            device.synchronize()
            Jg, Kg = [x.download().reshape(n, n, order="F") for x in bufs[2:]]
            _check_against_oracle(Jg, Kg, D, ref, 1e-9, 1e-8)
            assert abs(np.linalg.norm(Jg) - ref["J_fro"]) <= 1e-8 and abs(np.linalg.norm(Kg) - ref["K_fro"]) <= 1e-8
    finally:
        for x in bufs:
            x.free()


def test_c20h42_mixed_graph_build_at_the_autotuned_cut(c20):
    """The headline build: mixed precision at the cut the autotuner picks (as the benchmark does),
    graph-replayed: |dE2| <= 1e-6 Eh against the oracle's FP64 build."""
    b, D, Q, ref, g = c20
    print("c20h42 autotune + mixed graph", flush=True)
    cut, table = autotune.tune_cut(g, D, target=1e-7, Q=Q)
    assert cut > 0.0, table
    for rep in range(2):
        J, K, _, st = g.jk(2 * D, D, Q=Q, precision=lib.HFX_MIXED, fp32_cut=cut, flags=lib.HFX_JK_USE_GRAPH)
        assert st["n_quartets"] == ref["n_quartets"] and st["n_fp32"] > 0.9 * st["n_quartets"]
        E2 = _check_against_oracle(J, K, D, ref, 1e-6, 1e-5)
        print("cut %.0e: |dE2| vs oracle %.2e, fp32 %d / %d" % (cut, abs(E2 - ref["E2"]), st["n_fp32"],
                                                                 st["n_quartets"]), flush=True)


def test_c20h42_scf_fp64_and_mixed(engine_lib, c20):
    """FP64 SCF with the benchmark's settings, and the mixed-precision SCF that finishes with FP64
    builds: both reach the oracle's energy at the fixed converged density (E = 2 tr(D h) + E2 + E_nn;
    the fixed density is itself a converged SCF density, so the energies agree to the SCF's own
    convergence, 1e-7 Eh) and each other to 1e-6 Eh."""
    b, D, Q, ref, g = c20
    print("c20h42 scf", flush=True)
    e_ref = ref["e_electronic"] + ref["e_nuclear"]
    r64 = lib.scf_run(engine_lib, b, options=engine_lib.default_options(**SCF), outputs=False)
    assert r64["converged"] == 1
    assert abs(r64["e_nuclear"] - ref["e_nuclear"]) <= 1e-9
    # E_total of the SCF is formed with its last (DIIS-extrapolated) Fock matrix and stops at |dE| < 1e-8:
    # it agrees with the energy functional at the converged density to a few 1e-8
    assert abs(r64["e_total"] - e_ref) <= 1e-7, (r64["e_total"], e_ref)
    rmx = lib.scf_run(engine_lib, b, options=engine_lib.default_options(
        precision=lib.HFX_MIXED, fp32_cut=1e-2, final_fp64=1, jk_flags=lib.HFX_JK_USE_GRAPH, **SCF), outputs=False)
    assert rmx["converged"] == 1 and 0 < rmx["iterations_mixed"] < rmx["iterations"]
    assert abs(rmx["e_total"] - r64["e_total"]) <= 1e-6
    assert abs(rmx["e_total"] - e_ref) <= 1e-6
    print("c20h42 scf: fp64 %.10f (%d it), mixed %.10f (%d+%d it), oracle E(D) %.10f" % (
        r64["e_total"], r64["iterations"], rmx["e_total"], rmx["iterations_mixed"],
        rmx["iterations"] - rmx["iterations_mixed"], e_ref), flush=True)


def test_c20h42_scf_device_linalg(engine_lib, c20):
    """The mixed-precision SCF with the iteration's linear algebra on the device (rocSOLVER's
    eigensolver, DIIS and densities in HBM) over the benchmark's engine: the oracle's energy at
    the converged density to 1e-6 Eh, as the host-driven run."""
    b, D, Q, ref, g = c20
    print("c20h42 device scf", flush=True)
    e_ref = ref["e_electronic"] + ref["e_nuclear"]
    kw = dict(precision=lib.HFX_MIXED, fp32_cut=1e-2, final_fp64=1, jk_flags=lib.HFX_JK_USE_GRAPH,
              device_linalg=1, **SCF)
    r = lib.scf_run(engine_lib, b, options=engine_lib.default_options(**kw), outputs=False, engine=g)
    assert r["converged"] == 1 and 0 < r["iterations_mixed"] < r["iterations"]
    assert abs(r["e_total"] - e_ref) <= 1e-6
    print("c20h42 device scf: %.10f (%d+%d it) in %.3f s (J/K %.3f s, rest %.3f s, library load %.3f s)" % (
        r["e_total"], r["iterations_mixed"], r["iterations"] - r["iterations_mixed"], r["time_total_s"],
        r["time_jk_s"], r["time_diag_s"], r["time_linalg_load_s"]), flush=True)
