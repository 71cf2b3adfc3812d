"""The SCF iteration with its linear algebra on the device (device_linalg = 1: DIIS, X^T F X,
rocSOLVER's eigensolver, densities and energy in HBM) against the host driver's iteration over
the same engine, and against the oracle's SCF (reference SCF::run, src/SCF.cpp:23-136)."""
import numpy as np
import pytest

import molecules
from hfx import lib

pytestmark = pytest.mark.gpu

DET = lib.HFX_JK_DETERMINISTIC
# Plain Roothaan steps at tight thresholds: the stopping iteration is well defined and the
# energies agree far below the convergence tolerance (the DIIS energy is not variational).
TIGHT = dict(energy_tol=1e-11, density_tol=1e-9, use_diis=0, max_iter=150, schwarz_thresh=1e-14,
             density_thresh=1e-14)


def _run(L, b, dev, charge=0, mult=1, **kw):
    return lib.scf_run(L, b, charge=charge, multiplicity=mult, options=L.default_options(device_linalg=dev, **kw))


def _check_orbitals(r, S, nocc, beta=False):
    """C^T S C = 1, F C = S C eps, D = C_occ C_occ^T (the eigenproblem the iteration solved)."""
    s = "beta" if beta else "alpha"
    C, F, D, eps = r["C_" + s], r["F_" + s], r["D_" + s], r["eps_" + s]
    assert np.abs(C.T @ S @ C - np.eye(len(eps))).max() <= 1e-9
    assert np.all(np.diff(eps) >= -1e-12)
    assert np.abs(D - C[:, :nocc] @ C[:, :nocc].T).max() <= 1e-10


def test_rhf_tight_matches_host_and_oracle(engine_lib, oracle):
    b = molecules.water_dimer("6-31g**")
    print("device scf rhf tight", b.n_aos, flush=True)
    rd = _run(engine_lib, b, 1, jk_flags=DET, **TIGHT)
    rh = _run(engine_lib, b, 0, jk_flags=DET, **TIGHT)
    ro = _run(oracle, b, 0, **TIGHT)
    assert rd["converged"] == rh["converged"] == ro["converged"] == 1
    assert abs(rd["iterations"] - rh["iterations"]) <= 1
    assert abs(rd["e_total"] - rh["e_total"]) <= 1e-10
    assert abs(rd["e_total"] - ro["e_total"]) <= 1e-10
    assert np.abs(rd["D_alpha"] - rh["D_alpha"]).max() <= 1e-8
    assert np.abs(rd["F_alpha"] - ro["F_alpha"]).max() <= 1e-8
    assert np.abs(rd["eps_alpha"] - ro["eps_alpha"]).max() <= 1e-8
    assert np.abs(rd["h_core"] - rh["h_core"]).max() == 0.0 and np.abs(rd["S"] - rh["S"]).max() == 0.0
    _check_orbitals(rd, rd["S"], 10)
    assert rd["n_jk_builds"] == rd["iterations"]
    tr = rd["trace"]
    assert tr.shape == (rd["iterations"], 4) and tr[-1, 1] == rd["last_dE"] and tr[-1, 2] == rd["last_dD"]
    assert rd["time_jk_s"] > 0 and rd["time_diag_s"] > 0 and rd["time_linalg_load_s"] >= 0


def test_diis_defaults_match_host_and_oracle(engine_lib, oracle):
    """DIIS 6 (the reference's default): the device's DIIS (Gram matrix by DGEMM, QR with column
    pivoting in one thread) follows the host's to the same energy and iteration count.  The
    DIIS-extrapolated energy of this system oscillates at 1e-8 for tens of iterations before it
    settles, so the run stops at 1e-7."""
    b = molecules.water_dimer("6-31g**")
    print("device scf diis", flush=True)
    kw = dict(energy_tol=1e-7, schwarz_thresh=1e-12, density_thresh=1e-12, jk_flags=DET)
    rd = _run(engine_lib, b, 1, **kw)
    rh = _run(engine_lib, b, 0, **kw)
    kw.pop("jk_flags")
    ro = _run(oracle, b, 0, **kw)
    assert rd["converged"] == rh["converged"] == ro["converged"] == 1
    assert rd["iterations"] == rh["iterations"] == ro["iterations"]
    assert abs(rd["e_total"] - rh["e_total"]) <= 1e-9 and abs(rd["e_total"] - ro["e_total"]) <= 1e-9
    assert rd["last_diis_err"] < 1e-6
    assert np.abs(rd["trace"] - rh["trace"]).max() <= 1e-8
    _check_orbitals(rd, rd["S"], 10)


@pytest.mark.parametrize("diis", [0, 1])
def test_uhf_damping_and_level_shift(engine_lib, oracle, diis):
    """H2O+ doublet UHF (no degenerate orbitals, so the state is unique) with damping for the first
    iterations and a level shift on both spins: the device run takes the host's decisions (damping
    and shift by iteration, dE and the HOMO-LUMO gaps read back from the device), follows the
    host's trajectory and removes the shift from the reported virtuals."""
    b = molecules.h2o("6-31g**")
    print("device scf uhf damp/shift diis", diis, flush=True)
    tol = dict(energy_tol=1e-7, density_tol=1e-6) if diis else dict(energy_tol=1e-10, density_tol=1e-8)
    kw = dict(use_diis=diis, max_iter=200, damp=40, max_damp_iter=4, **tol,
              level_shift=0.3, lshift_gap_tol=5.0, max_lshift_iter=6, schwarz_thresh=1e-13, density_thresh=1e-13)
    rd = _run(engine_lib, b, 1, charge=1, mult=2, jk_flags=DET, **kw)
    rh = _run(engine_lib, b, 0, charge=1, mult=2, jk_flags=DET, **kw)
    ro = _run(oracle, b, 0, charge=1, mult=2, **kw)
    assert rd["converged"] == rh["converged"] == ro["converged"] == 1
    assert rd["iterations"] == rh["iterations"] == ro["iterations"]
    assert abs(rd["e_total"] - rh["e_total"]) <= 1e-9 and abs(rd["e_total"] - ro["e_total"]) <= 1e-9
    assert np.abs(rd["trace"][:, 0] - rh["trace"][:, 0]).max() <= 1e-8
    for s in ("alpha", "beta"):
        assert np.abs(rd["eps_" + s] - rh["eps_" + s]).max() <= 1e-7
        assert np.abs(rd["D_" + s] - rh["D_" + s]).max() <= 1e-6
    _check_orbitals(rd, rd["S"], 5)
    _check_orbitals(rd, rd["S"], 4, beta=True)


def test_rhf_level_shift_only(engine_lib):
    """Level shift throughout an RHF run (the shift stays on when the SCF stops): the reported
    virtual energies have it removed, as the host driver's."""
    b = molecules.water_dimer("6-31g")
    kw = dict(energy_tol=1e-10, density_tol=1e-8, max_iter=200, level_shift=0.5, lshift_gap_tol=10.0,
              max_lshift_iter=500, use_diis=0, jk_flags=DET, schwarz_thresh=1e-13, density_thresh=1e-13)
    rd = _run(engine_lib, b, 1, **kw)
    rh = _run(engine_lib, b, 0, **kw)
    assert rd["converged"] == rh["converged"] == 1
    assert abs(rd["e_total"] - rh["e_total"]) <= 1e-9
    assert np.abs(rd["eps_alpha"] - rh["eps_alpha"]).max() <= 1e-7


@pytest.mark.parametrize("switch_tol", [1e-4, 0.0])
def test_mixed_stages(engine_lib, switch_tol):
    """Mixed builds then FP64 from a full build, both stage switches: the device iteration ends
    within 1e-8 Eh of the host's and 1e-6 Eh of the FP64 SCF."""
    b = molecules.ethane()
    print("device scf mixed", switch_tol, flush=True)
    kw = dict(energy_tol=1e-9, precision=lib.HFX_MIXED, fp32_cut=1e-4, final_fp64=1, switch_tol=switch_tol)
    rd = _run(engine_lib, b, 1, **kw)
    rh = _run(engine_lib, b, 0, **kw)
    r64 = _run(engine_lib, b, 1, energy_tol=1e-9)
    assert rd["converged"] == rh["converged"] == r64["converged"] == 1
    assert 0 < rd["iterations_mixed"] < rd["iterations"]
    assert abs(rd["e_total"] - rh["e_total"]) <= 1e-8
    assert abs(rd["e_total"] - r64["e_total"]) <= 1e-6
    assert abs(rd["e_total_mixed"] - rh["e_total_mixed"]) <= 1e-6


def test_conventional_and_graph_builds(engine_lib):
    """Stored integrals (direct = 0) and graph-replayed direct builds under the device iteration."""
    b = molecules.water_dimer("6-31g**")
    rc = _run(engine_lib, b, 1, direct=0, jk_flags=DET, **TIGHT)
    rh = _run(engine_lib, b, 0, direct=0, jk_flags=DET, **TIGHT)
    assert rc["converged"] == rh["converged"] == 1
    assert abs(rc["e_total"] - rh["e_total"]) <= 1e-10
    rg = _run(engine_lib, b, 1, jk_flags=lib.HFX_JK_USE_GRAPH, **TIGHT)
    assert rg["converged"] == 1 and abs(rg["e_total"] - rh["e_total"]) <= 1e-9


def test_refusals(engine_lib):
    b = molecules.h2o("sto-3g")
    with pytest.raises(lib.HfxError, match="diis_size"):
        _run(engine_lib, b, 1, diis_size=16)
    with pytest.raises(lib.HfxError, match="one molecule"):
        lib.scf_run_batched(engine_lib, [b, b], options=engine_lib.default_options(device_linalg=1))
    r = _run(engine_lib, b, 1)  # the engine is still usable
    assert r["converged"] == 1
