"""Graph-captured builds and the batched multi-molecule SCF on the HIP engine."""
import numpy as np
import pytest

import molecules
from hfx import Basis, device, geometry, lib

pytestmark = pytest.mark.gpu


def _density(n, nocc, seed, scale=0.1):
    rng = np.random.default_rng(seed)
    C = rng.standard_normal((n, nocc))
    return C @ C.T * scale


def _close(a, b, rel=1e-12):
    return np.abs(a - b).max() <= rel * max(1.0, np.abs(b).max())


@pytest.fixture(scope="module")
def ethane(engine_lib):
    b = molecules.ethane()
    return b, lib.Engine(engine_lib, b)


def test_graph_replay_matches_eager(ethane):
    """Replays of the captured build pipeline agree with eager builds, with eager builds and
    host synchronizations between the replays, and a re-capture after a parameter change."""
    b, g = ethane
    n = b.n_aos
    print("graph replay ethane", n, flush=True)
    Ds = [_density(n, 9, 100 + i) for i in range(4)]
    for rep, D in enumerate(Ds + Ds[:2]):
        Je, Ke, _, se = g.jk(2 * D, D)
        Jg, Kg, _, sg = g.jk(2 * D, D, flags=lib.HFX_JK_USE_GRAPH)
        assert (sg["n_quartets"], sg["hash_sum"], sg["hash_xor"]) == (se["n_quartets"], se["hash_sum"], se["hash_xor"])
        assert _close(Jg, Je) and _close(Kg, Ke), rep
    # a different threshold set re-captures; then mixed precision through the graph
    D = Ds[0]
    Je, Ke, _, se = g.jk(2 * D, D, schwarz_thresh=1e-12, density_thresh=1e-12)
    Jg, Kg, _, sg = g.jk(2 * D, D, schwarz_thresh=1e-12, density_thresh=1e-12, flags=lib.HFX_JK_USE_GRAPH)
    assert sg["n_quartets"] == se["n_quartets"] and _close(Jg, Je) and _close(Kg, Ke)
    for rep in range(3):
        Jm, Km, _, sm = g.jk(2 * D, D, precision=lib.HFX_MIXED, fp32_cut=1e-4)
        Jmg, Kmg, _, smg = g.jk(2 * D, D, precision=lib.HFX_MIXED, fp32_cut=1e-4, flags=lib.HFX_JK_USE_GRAPH)
        assert smg["n_fp32"] == sm["n_fp32"] and smg["hash_sum"] == sm["hash_sum"]
        assert _close(Jmg, Jm) and _close(Kmg, Km), rep


def test_graph_replay_uhf_device_buffers(ethane):
    """UHF builds on device-resident buffers through the graph match eager host-buffer builds."""
    b, g = ethane
    n = b.n_aos
    print("graph replay uhf device", n, flush=True)
    for rep in range(3):
        Da, Db = _density(n, 9, 200 + rep), _density(n, 8, 300 + rep)
        Je, Kae, Kbe, _ = g.jk(Da + Db, Da, Db)
        bufs = [device.DeviceArray.from_host(x) for x in (Da + Db, Da, Db)]
        outs = [device.DeviceArray(n * n) for _ in range(3)]
        lib.jk_device(g, bufs[0], bufs[1], outs[0], outs[1], dD_beta=bufs[2], Kb=outs[2],
                      flags=lib.HFX_JK_USE_GRAPH)
        device.synchronize()
        Jg, Kag, Kbg = [o.download().reshape(n, n, order="F") for o in outs]
        assert _close(Jg, Je) and _close(Kag, Kae) and _close(Kbg, Kbe), rep
        for x in bufs + outs:
            x.free()


def test_scf_graph_flag_matches_eager(engine_lib):
    """The SCF driver with graph-replayed builds converges to the eager energy."""
    b = molecules.water_dimer("6-31g**")
    kw = dict(energy_tol=1e-8, max_iter=60)
    r0 = lib.scf_run(engine_lib, b, options=engine_lib.default_options(**kw))
    r1 = lib.scf_run(engine_lib, b, options=engine_lib.default_options(jk_flags=lib.HFX_JK_USE_GRAPH, **kw))
    assert r0["converged"] == 1 and r1["converged"] == 1
    assert abs(r0["e_total"] - r1["e_total"]) <= 1e-9
    assert np.abs(r0["F_alpha"] - r1["F_alpha"]).max() <= 1e-8


def _batch(n_mol, seed, nmax):
    return [Basis("6-31g", Z, x) for Z, x in geometry.water_batch(n_mol=n_mol, seed=seed, nmax=nmax)]


def test_batched_scf_matches_oracle(engine_lib, oracle):
    """Batched FP64 SCF over molecules of uneven size: every molecule's energy and
    iteration count match the oracle's single-molecule SCF."""
    bases = _batch(6, 1004, 3)
    print("batched scf", [b.n_aos for b in bases], flush=True)
    kw = dict(energy_tol=1e-8, max_iter=60)
    rb = lib.scf_run_batched(engine_lib, bases, options=engine_lib.default_options(**kw))
    for b, r in zip(bases, rb):
        ro = lib.scf_run(oracle, b, options=oracle.default_options(**kw), outputs=False)
        assert r["converged"] == 1 and ro["converged"] == 1
        assert abs(r["e_total"] - ro["e_total"]) <= 1e-9, (b.n_aos, r["e_total"], ro["e_total"])


@pytest.mark.parametrize("switch_tol", [1e-4, 0.0])
def test_batched_scf_mixed_graph(engine_lib, switch_tol):
    """Batched mixed-precision SCF with graph-replayed builds: the FP64 stage (after an early
    switch, or after converging in mixed precision: then that stop is within 1e-6 of FP64)
    agrees with the plain FP64 batched run."""
    bases = _batch(12, 1004, 8)
    print("batched mixed", switch_tol, [b.n_aos for b in bases], flush=True)
    kw = dict(energy_tol=1e-8, max_iter=60)
    r64 = lib.scf_run_batched(engine_lib, bases, options=engine_lib.default_options(**kw))
    rmx = lib.scf_run_batched(engine_lib, bases, options=engine_lib.default_options(
        precision=lib.HFX_MIXED, fp32_cut=1e-4, final_fp64=1, switch_tol=switch_tol, jk_flags=lib.HFX_JK_USE_GRAPH,
        **kw))
    for a, m in zip(r64, rmx):
        assert a["converged"] == 1 and m["converged"] == 1
        assert 0 < m["iterations_mixed"] < m["iterations"]
        if switch_tol == 0.0:
            assert abs(m["e_total_mixed"] - a["e_total"]) <= 1e-6
        # both runs stop at |dE| < 1e-8 between iterations: their energies agree to a few 1e-8
        assert abs(m["e_total"] - a["e_total"]) <= 1e-7
