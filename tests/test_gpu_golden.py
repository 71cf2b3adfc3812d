"""The (H2O)16 / 6-31G** FP64 SCF (400 Cartesian functions, every class through dddd) on the
HIP engine against the oracle's run stored in tests/golden/water16_scf.json."""
import json
import os

import numpy as np
import pytest

from hfx import Basis, geometry, lib

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


def test_water16_fp64_scf_matches_golden(engine_lib):
    with open(os.path.join(HERE, "golden", "water16_scf.json")) as f:
        ref = json.load(f)
    Z, xyz = geometry.water_cluster_grid(shape=(4, 2, 2), spacing=5.6, jitter=0.4, seed=1002)
    b = Basis("6-31g**", Z, xyz)
    assert b.n_aos == ref["n_aos"] == 400
    print("water16 scf", b.n_aos, flush=True)
    opts = engine_lib.default_options(energy_tol=1e-8, schwarz_thresh=1e-10, density_thresh=1e-10, diis_size=6)
    r = lib.scf_run(engine_lib, b, options=opts)
    assert r["converged"] == 1 and r["iterations"] == ref["iterations"]
    assert abs(r["e_nuclear"] - ref["e_nuclear"]) <= 1e-9
    assert abs(r["e_total"] - ref["e_total"]) <= 1e-9
    assert np.abs(np.diag(r["F_alpha"]) - np.asarray(ref["F_diag"])).max() <= 1e-8
    assert np.abs(r["eps_alpha"] - np.asarray(ref["eps"])).max() <= 1e-8
    assert abs(np.linalg.norm(r["F_alpha"]) - ref["F_fro"]) <= 1e-8
