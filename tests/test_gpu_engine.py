"""Parity of the HIP engine (through its C ABI) with the CPU oracle on the same inputs."""
import numpy as np
import pytest

import molecules
import os_reference
from hfx import lib

pytestmark = pytest.mark.gpu

MOLS = {
    "h2o_sto3g": lambda: molecules.h2o("sto-3g"),
    "h2o_631gss": lambda: molecules.h2o("6-31g**"),
    "dimer_631gss": lambda: molecules.water_dimer("6-31g**"),
    "methane_ccpvdz": lambda: molecules.methane(),
    "ethane_ccpvdz": lambda: molecules.ethane(),
}


@pytest.fixture(scope="module")
def pairs(engine_lib, oracle):
    """(basis, GPU engine, oracle engine) per test molecule, created once."""
    out = {}
    for k, f in MOLS.items():
        b = f()
        out[k] = (b, lib.Engine(engine_lib, b), lib.Engine(oracle, b))
    return out


def _density(n, nocc, seed, scale=0.1):
    rng = np.random.default_rng(seed)
    C = rng.standard_normal((n, nocc))
    return C @ C.T * scale


@pytest.mark.parametrize("mol", list(MOLS))
def test_one_electron(pairs, mol):
    b, g, o = pairs[mol]
    print("one-electron", mol, flush=True)
    for name in ("overlap", "kinetic"):
        x, y = getattr(g, name)(), getattr(o, name)()
        assert np.abs(x - y).max() <= 1e-12 * max(1.0, np.abs(y).max()), name
    x, y = g.nuclear(b.Z, b.xyz), o.nuclear(b.Z, b.xyz)
    assert np.abs(x - y).max() <= 1e-11 * max(1.0, np.abs(y).max())


def test_one_electron_against_obara_saika(pairs):
    """GPU S, T, V of every shell pair of methane / cc-pVDZ ((la, lb) through (d, d)) against the
    independent Obara-Saika recurrences of tests/os_reference.py, not only against the oracle."""
    b, g, _ = pairs["methane_ccpvdz"]
    S, T, V = g.overlap(), g.kinetic(), g.nuclear(b.Z, b.xyz)
    off = b.flat()["ao_offset"]
    for sa, A in enumerate(b.shells):
        for sb, B in enumerate(b.shells[: sa + 1]):
            ref = os_reference.one_electron_block(b, sa, sb, b.Z, b.xyz)
            for M, want in zip((S, T, V), ref):
                got = M[off[sa]:off[sa] + A.nao, off[sb]:off[sb] + B.nao]
                assert np.abs(got - want).max() <= 1e-11 * max(1.0, np.abs(want).max()), (sa, sb)


@pytest.mark.parametrize("mol", list(MOLS))
def test_schwarz(pairs, mol):
    b, g, o = pairs[mol]
    print("schwarz", mol, flush=True)
    qg, qo = g.schwarz(), o.schwarz()
    assert np.all(qg == qg.T)
    assert np.abs(qg - qo).max() <= 1e-13 * qo.max()


@pytest.mark.parametrize("mol", ["h2o_631gss", "ethane_ccpvdz"])
def test_eri_quartets(pairs, mol):
    b, g, o = pairs[mol]
    print("eri quartets", mol, flush=True)
    nS = b.n_shells
    rng = np.random.default_rng(1)
    quartets = [(a, bb, c, d) for a in range(nS) for bb in range(a + 1) for c in range(nS) for d in range(c + 1)
                if a * (a + 1) // 2 + bb >= c * (c + 1) // 2 + d]
    if len(quartets) > 4000:
        idx = rng.choice(len(quartets), 4000, replace=False)
        quartets = [quartets[i] for i in sorted(idx)]
    # also non-canonical orders, which the ABI accepts with the reference's zero rules
    quartets += [(1, 4, 2, 0), (0, 0, 3, 3), (2, 5, 5, 2)]
    bg, bo = g.eri_quartets(quartets), o.eri_quartets(quartets)
    worst = 0.0
    for q, x, y in zip(quartets, bg, bo):
        assert x.shape == y.shape
        assert np.array_equal(x == 0, y == 0) or np.abs(x - y).max() < 1e-14, q
        worst = max(worst, float(np.abs(x - y).max()))
    assert worst < 1e-12


@pytest.mark.parametrize("mol", list(MOLS))
def test_jk_matches_oracle(pairs, mol):
    b, g, o = pairs[mol]
    print("jk", mol, flush=True)
    n = b.n_aos
    Da = _density(n, max(1, b.n_electrons() // 2), 5)
    Q = o.schwarz()  # the same Schwarz table: screening decisions must coincide exactly
    Jg, Kg, _, sg = g.jk(2 * Da, Da, Q=Q)
    Jo, Ko, _, so = o.jk(2 * Da, Da, Q=Q)
    assert sg["n_quartets"] == so["n_quartets"]
    assert sg["class_count"] == so["class_count"]
    assert sg["hash_sum"] == so["hash_sum"] and sg["hash_xor"] == so["hash_xor"]
    assert sg["model_gflop"] == pytest.approx(so["model_gflop"], rel=1e-12)
    scale = max(1.0, np.abs(Jo).max())
    assert np.abs(Jg - Jo).max() <= 1e-11 * scale
    assert np.abs(Kg - Ko).max() <= 1e-11 * scale
    assert np.abs(Jg - Jg.T).max() == 0.0 and np.abs(Kg - Kg.T).max() == 0.0


@pytest.mark.parametrize("mol", ["dimer_631gss", "ethane_ccpvdz"])
def test_jk_screened_density(pairs, mol):
    """Small, sparse density increments: density screening removes most quartets."""
    b, g, o = pairs[mol]
    print("jk screened", mol, flush=True)
    n = b.n_aos
    D = _density(n, 3, 9, scale=1e-7)
    D[np.abs(D) < 2e-9] = 0.0
    Q = o.schwarz()
    Jg, Kg, _, sg = g.jk(2 * D, D, Q=Q, schwarz_thresh=1e-10, density_thresh=1e-8)
    Jo, Ko, _, so = o.jk(2 * D, D, Q=Q, schwarz_thresh=1e-10, density_thresh=1e-8)
    assert (sg["n_quartets"], sg["hash_sum"], sg["hash_xor"]) == (so["n_quartets"], so["hash_sum"], so["hash_xor"])
    assert 0 < sg["n_quartets"]
    assert np.abs(Jg - Jo).max() <= 1e-16 + 1e-11 * np.abs(Jo).max()
    assert np.abs(Kg - Ko).max() <= 1e-16 + 1e-11 * np.abs(Ko).max()


def test_jk_uhf(engine_lib, oracle):
    b = molecules.oh_radical()
    g, o = lib.Engine(engine_lib, b), lib.Engine(oracle, b)
    n = b.n_aos
    Da, Db = _density(n, 5, 1), _density(n, 4, 2)
    Q = o.schwarz()
    Jg, Kag, Kbg, sg = g.jk(Da + Db, Da, Db, Q=Q)
    Jo, Kao, Kbo, so = o.jk(Da + Db, Da, Db, Q=Q)
    assert (sg["n_quartets"], sg["hash_sum"]) == (so["n_quartets"], so["hash_sum"])
    for x, y in ((Jg, Jo), (Kag, Kao), (Kbg, Kbo)):
        assert np.abs(x - y).max() <= 1e-11 * max(1.0, np.abs(y).max())


def test_jk_against_bruteforce_tensor(pairs):
    """GPU J/K against einsum over the dense tensor of the oracle's ERIs (no screening)."""
    b, g, o = pairs["h2o_631gss"]
    G = os_reference.dense_eri(o, b)
    D = _density(b.n_aos, 5, 3)
    J, K, _, _ = g.jk(2 * D, D, schwarz_thresh=0.0, density_thresh=0.0)
    assert np.abs(J - np.einsum("ijkl,kl->ij", G, 2 * D)).max() < 1e-11
    assert np.abs(K - np.einsum("ijkl,jk->il", G, D)).max() < 1e-11


def test_jk_mixed_precision(pairs):
    b, g, o = pairs["ethane_ccpvdz"]
    D = _density(b.n_aos, 9, 4)
    J64, K64, _, s64 = g.jk(2 * D, D)
    Jm, Km, _, sm = g.jk(2 * D, D, precision=lib.HFX_MIXED, fp32_cut=1e-3)
    assert sm["n_quartets"] == s64["n_quartets"] and sm["hash_sum"] == s64["hash_sum"]
    assert 0 < sm["n_fp32"] < sm["n_quartets"]
    # FP32 quartets have bound < 1e-3: their absolute error stays far below 1e-6
    assert np.abs(Jm - J64).max() < 1e-6
    assert np.abs(Km - K64).max() < 1e-6


def test_batched_engine_matches_single(engine_lib):
    from hfx import geometry, Basis
    mols = geometry.water_batch(n_mol=5, seed=21, nmax=3)
    bases = [Basis("6-31g", Z, x) for Z, x in mols]
    gb = lib.Engine(engine_lib, bases)
    Ds = [_density(bb.n_aos, 5 * (len(bb.Z) // 3), 30 + i) for i, bb in enumerate(bases)]
    Jb, Kb, _, sb = gb.jk([2 * D for D in Ds], Ds)
    tot = 0
    for i, bb in enumerate(bases):
        e1 = lib.Engine(engine_lib, bb)
        J1, K1, _, s1 = e1.jk(2 * Ds[i], Ds[i])
        tot += s1["n_quartets"]
        assert np.abs(Jb[i] - J1).max() < 1e-12 * max(1, np.abs(J1).max())
        assert np.abs(Kb[i] - K1).max() < 1e-12 * max(1, np.abs(K1).max())
    assert sb["n_quartets"] == tot
    Sb = gb.overlap()
    for i, bb in enumerate(bases):
        assert np.abs(Sb[i] - lib.Engine(engine_lib, bb).overlap()).max() < 1e-14


# ---------------------------------------------------------------------------------------
# Whole SCF through the same host driver over each engine
# ---------------------------------------------------------------------------------------
def test_scf_h2o_sto3g(engine_lib, oracle):
    """H2O/STO-3G with the reference's settings (Schwarz 1e-10, density 1e-10, DIIS 6,
    energy tolerance 1e-8)."""
    b = molecules.h2o("sto-3g")
    kw = dict(energy_tol=1e-8, schwarz_thresh=1e-10, density_thresh=1e-10, diis_size=6)
    rg = lib.scf_run(engine_lib, b, options=engine_lib.default_options(**kw))
    ro = lib.scf_run(oracle, b, options=oracle.default_options(**kw))
    assert rg["converged"] == 1 and ro["converged"] == 1
    assert rg["iterations"] == ro["iterations"]
    assert abs(rg["e_total"] - ro["e_total"]) <= 1e-9
    assert np.abs(rg["F_alpha"] - ro["F_alpha"]).max() <= 1e-8


def test_scf_tight_parity(engine_lib, oracle):
    """Tight thresholds remove the incremental-build noise floor: energies agree to 1e-10."""
    b = molecules.water_dimer("6-31g**")
    # plain Roothaan steps: the DIIS-extrapolated energy is not variational and hovers near 1e-10
    kw = dict(energy_tol=1e-11, density_tol=1e-9, use_diis=0, max_iter=150, schwarz_thresh=1e-14,
              density_thresh=1e-14)
    rg = lib.scf_run(engine_lib, b, options=engine_lib.default_options(**kw))
    ro = lib.scf_run(oracle, b, options=oracle.default_options(**kw))
    assert rg["converged"] == 1 and ro["converged"] == 1
    assert abs(rg["e_total"] - ro["e_total"]) <= 1e-10
    assert np.abs(rg["F_alpha"] - ro["F_alpha"]).max() <= 1e-8


@pytest.mark.parametrize("switch_tol", [1e-4, 0.0])
def test_scf_mixed_recovers_fp64(engine_lib, switch_tol):
    """Mixed builds, then FP64 from a full build: switching early (DIIS history kept) or after
    converging in mixed precision (fresh DIIS) both end within 1e-6 Eh of the FP64 SCF."""
    b = molecules.ethane()
    kw = dict(energy_tol=1e-9)
    r64 = lib.scf_run(engine_lib, b, options=engine_lib.default_options(**kw))
    rmx = lib.scf_run(engine_lib, b, options=engine_lib.default_options(precision=lib.HFX_MIXED, fp32_cut=1e-4,
                                                                         final_fp64=1, switch_tol=switch_tol, **kw))
    assert r64["converged"] == 1 and rmx["converged"] == 1
    assert abs(rmx["e_total"] - r64["e_total"]) <= 1e-6
    assert 0 < rmx["iterations_mixed"] < rmx["iterations"]
