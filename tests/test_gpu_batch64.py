"""The 64-molecule batched workload (seeded water clusters, n ~ U{1..8}, 6-31G, 13-104 functions
per molecule; hfx.geometry.water_batch(64, seed 1004)) on the HIP engine, molecule by molecule
against the oracle's single-molecule SCF energies in tests/golden/water_batch64.json
(written by tests/golden/make_water_batch64.py)."""
import json
import os

import pytest

from hfx import Basis, geometry, lib

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def batch():
    with open(os.path.join(HERE, "golden", "water_batch64.json")) as f:
        ref = json.load(f)
    mols = [Basis("6-31g", Z, x) for Z, x in geometry.water_batch(n_mol=64, seed=1004)]
    assert [m.n_aos for m in mols] == [r["n_aos"] for r in ref["molecules"]]
    return mols, ref


def test_batch64_fp64_matches_oracle(engine_lib, batch):
    """FP64 batched SCF (every iteration's Fock builds of all molecules still iterating in one
    hfx_jk call) with the fixture's tight settings: |dE| <= 1e-9 Eh per molecule."""
    mols, ref = batch
    print("batch64 fp64", sum(m.n_aos for m in mols), flush=True)
    r = lib.scf_run_batched(engine_lib, mols, options=engine_lib.default_options(**ref["options"]))
    for i, (x, y) in enumerate(zip(r, ref["molecules"])):
        assert x["converged"] == 1, i
        assert abs(x["e_total"] - y["e_total"]) <= 1e-9, (i, x["e_total"], y["e_total"])


@pytest.mark.parametrize("cut", [1e-2, 1e-4])
def test_batch64_mixed_graph_matches_oracle(engine_lib, batch, cut):
    """Mixed-precision, graph-replayed batched SCF (FP32 below the cut, then FP64 builds from a full
    FP64 build to convergence): |dE| <= 1e-6 Eh per molecule against the oracle."""
    mols, ref = batch
    print("batch64 mixed graph", cut, flush=True)
    opts = engine_lib.default_options(precision=lib.HFX_MIXED, fp32_cut=cut, final_fp64=1,
                                      jk_flags=lib.HFX_JK_USE_GRAPH, **ref["options"])
    r = lib.scf_run_batched(engine_lib, mols, options=opts)
    worst = 0.0
    for i, (x, y) in enumerate(zip(r, ref["molecules"])):
        assert x["converged"] == 1, i
        assert 0 < x["iterations_mixed"] <= x["iterations"], i
        worst = max(worst, abs(x["e_total"] - y["e_total"]))
    print("batch64 mixed max |dE| %.2e" % worst, flush=True)
    assert worst <= 1e-6
