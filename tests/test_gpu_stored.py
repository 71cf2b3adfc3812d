"""The conventional (stored-integral) path on the GPU: hfx_store_eris + hfx_jk(HFX_JK_STORED)
against the oracle's packed-tensor restatement of IntegralEngine::electronRepulsionTensor /
SCF::buildFockMatrix and against the GPU's own direct build with the density test off."""
import numpy as np
import pytest

import molecules
from hfx import Basis, geometry, lib

pytestmark = pytest.mark.gpu

MOLS = {
    "h2o_631gss": lambda: molecules.h2o("6-31g**"),
    "dimer_631gss": lambda: molecules.water_dimer("6-31g**"),
    "ethane_ccpvdz": lambda: molecules.ethane(),
}


def _density(n, nocc, seed, scale=0.1):
    rng = np.random.default_rng(seed)
    C = rng.standard_normal((n, nocc))
    return C @ C.T * scale


@pytest.mark.parametrize("mol", list(MOLS))
def test_stored_jk_matches_oracle_and_direct(engine_lib, oracle, mol):
    b = MOLS[mol]()
    print("stored jk", mol, flush=True)
    g, o = lib.Engine(engine_lib, b), lib.Engine(oracle, b)
    Q = o.schwarz()
    D = _density(b.n_aos, max(1, b.n_electrons() // 2), 21)
    ng, nbytes = g.store_eris(1e-10, Q=Q)
    no, _ = o.store_eris(1e-10, Q=Q)
    assert ng == no and nbytes > 0
    Jg, Kg, _, sg = g.jk(2 * D, D, schwarz_thresh=1e-10, flags=lib.HFX_JK_STORED)
    Jo, Ko, _, so = o.jk(2 * D, D, schwarz_thresh=1e-10, flags=lib.HFX_JK_STORED)
    assert sg["n_quartets"] == ng
    scale = max(1.0, np.abs(Jo).max())
    assert np.abs(Jg - Jo).max() <= 1e-11 * scale
    assert np.abs(Kg - Ko).max() <= 1e-11 * scale
    assert np.abs(Jg - Jg.T).max() == 0.0 and np.abs(Kg - Kg.T).max() == 0.0
    # the direct build with every density test passing visits the same quartets
    Jd, Kd, _, sd = g.jk(2 * D, D, Q=Q, schwarz_thresh=1e-10, density_thresh=0.0)
    assert sd["n_quartets"] == ng
    assert np.abs(Jg - Jd).max() <= 1e-12 * scale
    assert np.abs(Kg - Kd).max() <= 1e-12 * scale


def test_stored_uhf_matches_oracle(engine_lib, oracle):
    b = molecules.oh_radical()
    g, o = lib.Engine(engine_lib, b), lib.Engine(oracle, b)
    Da, Db = _density(b.n_aos, 5, 1), _density(b.n_aos, 4, 2)
    g.store_eris(1e-10)
    o.store_eris(1e-10)
    Jg, Kag, Kbg, _ = g.jk(Da + Db, Da, Db, schwarz_thresh=1e-10, flags=lib.HFX_JK_STORED)
    Jo, Kao, Kbo, _ = o.jk(Da + Db, Da, Db, schwarz_thresh=1e-10, flags=lib.HFX_JK_STORED)
    for x, y in ((Jg, Jo), (Kag, Kao), (Kbg, Kbo)):
        assert np.abs(x - y).max() <= 1e-11 * max(1.0, np.abs(y).max())


def test_stored_reuse_linearity_graph_and_release(engine_lib):
    """One store serves many densities (J/K linear in D), graph replays equal the eager build,
    a new store replaces the old one (and the captured graph), and release ends stored builds."""
    b = molecules.ethane()
    g = lib.Engine(engine_lib, b)
    n = b.n_aos
    D1, D2 = _density(n, 9, 3), _density(n, 9, 4)
    n1, _ = g.store_eris(1e-12)
    J1, K1, _, _ = g.jk(D1, D1, schwarz_thresh=1e-14, flags=lib.HFX_JK_STORED)
    J2, K2, _, _ = g.jk(D2, D2, schwarz_thresh=1e-14, flags=lib.HFX_JK_STORED)
    J12, K12, _, _ = g.jk(D1 + D2, D1 + D2, schwarz_thresh=1e-14, flags=lib.HFX_JK_STORED)
    assert np.abs(J12 - J1 - J2).max() < 1e-11 and np.abs(K12 - K1 - K2).max() < 1e-11
    fl = lib.HFX_JK_STORED | lib.HFX_JK_USE_GRAPH
    for _ in range(3):
        Jr, Kr, _, _ = g.jk(D1, D1, schwarz_thresh=1e-14, flags=fl)
        assert np.abs(Jr - J1).max() < 1e-12 and np.abs(Kr - K1).max() < 1e-12
    n2, _ = g.store_eris(1e-6)
    assert n2 < n1
    Jc, Kc, _, _ = g.jk(D1, D1, schwarz_thresh=1e-6, flags=fl)
    Jd, Kd, _, _ = g.jk(D1, D1, schwarz_thresh=1e-6, density_thresh=0.0)
    assert np.abs(Jc - Jd).max() < 1e-12 and np.abs(Kc - Kd).max() < 1e-12
    with pytest.raises(lib.HfxError):
        g.jk(D1, D1, flags=lib.HFX_JK_STORED, precision=lib.HFX_MIXED, fp32_cut=1e-2)
    g.release_eris()
    with pytest.raises(lib.HfxError):
        g.jk(D1, D1, flags=lib.HFX_JK_STORED)


def test_stored_batched_matches_single(engine_lib):
    mols = [molecules.h2o("6-31g**"), molecules.methane(), molecules.water_dimer("6-31g**")]
    eb = lib.Engine(engine_lib, mols)
    eb.store_eris(1e-10)
    Ds = [_density(m.n_aos, max(1, m.n_electrons() // 2), 30 + i) for i, m in enumerate(mols)]
    Jb, Kb, _, _ = eb.jk([2 * d for d in Ds], Ds, schwarz_thresh=1e-10, flags=lib.HFX_JK_STORED)
    for m, d, jb, kb in zip(mols, Ds, Jb, Kb):
        e = lib.Engine(engine_lib, m)
        e.store_eris(1e-10)
        J, K, _, _ = e.jk(2 * d, d, schwarz_thresh=1e-10, flags=lib.HFX_JK_STORED)
        assert np.abs(jb - J).max() < 1e-12 * max(1.0, np.abs(J).max())
        assert np.abs(kb - K).max() < 1e-12 * max(1.0, np.abs(K).max())


# Tight settings remove the incremental-build noise floor and make the stopping iteration
# well defined: plain Roothaan steps (the DIIS-extrapolated energy is not variational and
# hovers near 1e-10), thresholds 1e-14, energy tolerance 1e-11.  The GPU runs use the
# deterministic (fixed-point) accumulation, so every GPU run is bitwise reproducible.
TIGHT = dict(energy_tol=1e-11, density_tol=1e-9, use_diis=0, max_iter=150, schwarz_thresh=1e-14,
             density_thresh=1e-14)


def test_conventional_scf_matches_direct_and_oracle(engine_lib, oracle):
    """DIRECT false: the GPU driver on stored integrals reaches the direct full-build energy and
    the oracle's conventional energy to 1e-9 Eh, in the same number of iterations
    (reference conventional build src/SCF.cpp:233-521)."""
    b = molecules.water_dimer("6-31g**")
    det = lib.HFX_JK_DETERMINISTIC
    rc = lib.scf_run(engine_lib, b, options=engine_lib.default_options(direct=0, jk_flags=det, **TIGHT))
    rd = lib.scf_run(engine_lib, b, options=engine_lib.default_options(direct=1, full_builds=1, jk_flags=det, **TIGHT))
    ro = lib.scf_run(oracle, b, options=oracle.default_options(direct=0, **TIGHT))
    assert rc["converged"] == 1 and rd["converged"] == 1 and ro["converged"] == 1
    assert rc["iterations"] == rd["iterations"] == ro["iterations"]
    assert abs(rc["e_total"] - rd["e_total"]) <= 1e-9
    assert abs(rc["e_total"] - ro["e_total"]) <= 1e-9
    assert np.abs(rc["F_alpha"] - ro["F_alpha"]).max() <= 1e-8
    # and the deterministic conventional run repeats bit for bit
    rc2 = lib.scf_run(engine_lib, b, options=engine_lib.default_options(direct=0, jk_flags=det, **TIGHT))
    assert rc2["e_total"] == rc["e_total"] and np.array_equal(rc2["F_alpha"], rc["F_alpha"])


def test_stored_water8_against_direct(engine_lib):
    """(H2O)8 / 6-31G** (200 functions): stored and direct (density test off) builds agree."""
    Z, xyz = geometry.water_cluster_grid(shape=(2, 2, 2), spacing=5.6, jitter=0.4, seed=1002)
    b = Basis("6-31g**", Z, xyz)
    g = lib.Engine(engine_lib, b)
    D = _density(b.n_aos, 40, 8, scale=0.02)
    n, nbytes = g.store_eris(1e-10)
    print("water8 stored", n, nbytes, flush=True)
    Js, Ks, _, _ = g.jk(2 * D, D, schwarz_thresh=1e-10, flags=lib.HFX_JK_STORED)
    Jd, Kd, _, sd = g.jk(2 * D, D, schwarz_thresh=1e-10, density_thresh=0.0)
    assert sd["n_quartets"] == n
    scale = max(1.0, np.abs(Jd).max())
    assert np.abs(Js - Jd).max() <= 1e-12 * scale
    assert np.abs(Ks - Kd).max() <= 1e-12 * scale


def test_conventional_batched_scf_matches_direct(engine_lib, oracle):
    """DIRECT false in the batched driver: one store for all molecules, full-density builds through
    it; per-molecule energies equal the direct full-build batched run and the oracle's to 1e-9 Eh,
    in the same number of iterations."""
    mols = [molecules.h2o("6-31g**"), molecules.water_dimer("6-31g"), molecules.methane()]
    det = lib.HFX_JK_DETERMINISTIC
    rc = lib.scf_run_batched(engine_lib, mols, options=engine_lib.default_options(direct=0, jk_flags=det, **TIGHT))
    rd = lib.scf_run_batched(engine_lib, mols, options=engine_lib.default_options(direct=1, full_builds=1,
                                                                                  jk_flags=det, **TIGHT))
    ro = lib.scf_run_batched(oracle, mols, options=oracle.default_options(direct=0, **TIGHT))
    for a, b, c in zip(rc, rd, ro):
        assert a["converged"] == 1 and b["converged"] == 1 and c["converged"] == 1
        assert a["iterations"] == b["iterations"] == c["iterations"]
        assert abs(a["e_total"] - b["e_total"]) <= 1e-9
        assert abs(a["e_total"] - c["e_total"]) <= 1e-9
