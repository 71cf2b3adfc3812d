"""RHF nuclear gradient of the HIP engine: against the oracle's analytic gradient on the same
densities, against finite differences of SCF energies, in mixed precision, and through a
torch.autograd.Function."""
import json
import os

import numpy as np
import pytest

import gradient_tools as gt
import molecules
from hfx import Basis, geometry, lib

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


def _water8():
    Z, xyz = geometry.water_cluster_grid(shape=(2, 2, 2), spacing=5.6, jitter=0.4, seed=1002)
    return Basis("6-31g**", Z, xyz)


@pytest.mark.parametrize("mol", ["h2o_sto3g", "h2o_631gss", "dimer_631gss", "ethane"])
def test_gradient_matches_oracle(engine_lib, oracle, mol):
    b = {"h2o_sto3g": lambda: molecules.h2o("sto-3g"), "h2o_631gss": lambda: molecules.h2o("6-31g**"),
         "dimer_631gss": lambda: molecules.water_dimer("6-31g**"), "ethane": molecules.ethane}[mol]()
    print("gradient vs oracle", mol, b.n_aos, flush=True)
    r = gt.scf(oracle, b, full_builds=0, energy_tol=1e-9, density_tol=1e-7)
    P, W = gt.densities(r, b)
    kw = dict(schwarz_thresh=1e-12, density_thresh=1e-12)
    gg, pg, sg = lib.rhf_gradient(lib.Engine(engine_lib, b), b, P, W, **kw)
    go, po, so = lib.rhf_gradient(lib.Engine(oracle, b), b, P, W, **kw)
    assert sg["n_quartets"] == so["n_quartets"]
    for k in range(4):
        assert np.abs(pg[k] - po[k]).max() <= 1e-9, k
    assert np.abs(gg - go).max() <= 1e-9


def test_water8_gradient_against_finite_differences(engine_lib):
    """All 72 coordinates of (H2O)8 / 6-31G**: central differences (h = 1e-4 Bohr) of the
    engine's SCF energies, the 144 displaced molecules run as one batched SCF.

    The energies come from full Fock builds: with incremental builds the contributions that the
    density test skips accumulate along each molecule's SCF path, so two paths to the same
    density can differ by ~1e-9 Eh, which the 1/(2h) of the differences turns into ~1e-5."""
    b = _water8()
    print("water8 gradient vs FD", b.n_aos, flush=True)
    kw = dict(full_builds=1, energy_tol=1e-10, density_tol=1e-8, schwarz_thresh=1e-12, density_thresh=1e-12)
    r, g, parts, st = gt.analytic(engine_lib, b, **kw)
    fd = gt.finite_differences_batched(engine_lib, b, **kw)
    err = np.abs(fd - g).max()
    A, x = np.unravel_index(np.argmax(np.abs(fd - g)), g.shape)
    print("max |analytic - FD| = %.2e Eh/Bohr at atom %d, axis %d; max |g| = %.2e" % (err, A, x, np.abs(g).max()),
          flush=True)
    assert err <= 1e-6
    assert np.abs(g.sum(0)).max() <= 1e-9


def test_water8_gradient_against_oracle_fixture(engine_lib):
    """The oracle's FP64 analytic gradient and its finite differences, stored in
    tests/golden/water8_grad.json."""
    path = os.path.join(HERE, "golden", "water8_grad.json")
    if not os.path.exists(path):
        pytest.skip("fixture not generated")
    try:
        ref = json.load(open(path))
    except ValueError:
        pytest.skip("fixture incomplete")
    b = _water8()
    kw = dict(**ref["scf"])
    r, g, parts, st = gt.analytic(engine_lib, b, **kw)
    assert abs(r["e_total"] - ref["e_total"]) <= 1e-9
    assert np.abs(g - np.asarray(ref["grad"])).max() <= 1e-7
    for A, x, fd in ref["fd"]:
        assert abs(g[A, x] - fd) <= 1e-6, (A, x)


def test_gradient_mixed_precision(engine_lib):
    b = _water8()
    print("mixed gradient", flush=True)
    r = gt.scf(engine_lib, b, full_builds=0, energy_tol=1e-9, density_tol=1e-7)
    P, W = gt.densities(r, b)
    e = lib.Engine(engine_lib, b)
    g64, _, s64 = lib.rhf_gradient(e, b, P, W)
    gmx, _, smx = lib.rhf_gradient(e, b, P, W, precision=lib.HFX_MIXED, fp32_cut=1e-5)
    assert 0 < smx["n_fp32"] < smx["n_quartets"] == s64["n_quartets"]
    assert np.abs(gmx - g64).max() <= 1e-6
    # the same screened quartets, every one with at least one primitive quartet
    assert (smx["class_prims"] == s64["class_prims"]).all() and s64["class_prims"].sum() >= s64["n_quartets"]
    # per-class timing runs the class kernels one after another on one stream instead of spread
    # over four: the same gradient up to the order of the atomic additions
    gser, _, sser = lib.rhf_gradient(e, b, P, W, flags=lib.HFX_JK_CLASS_TIMING)
    assert np.abs(gser - g64).max() <= 1e-11 and (sser["class_ms"] > 0).sum() > 10


def test_autograd_function():
    """The SCF energy as a torch function of the coordinates: backward gives the analytic
    gradient, which matches finite differences.  Runs in a fresh process: PyTorch's ROCm
    libraries cannot be loaded into a process that already holds the system rocBLAS/rocSOLVER
    (the device-linear-algebra SCF tests load them), so torch is imported there first."""
    import importlib.util
    if importlib.util.find_spec("torch") is None:  # found without importing it here
        pytest.skip("torch is not installed")
    import subprocess
    import sys
    env = dict(os.environ)
    env["PYTHONPATH"] = os.pathsep.join([HERE, os.path.join(os.path.dirname(HERE), "hartree-fock_amd")] +
                                        ([env["PYTHONPATH"]] if env.get("PYTHONPATH") else []))
    print("autograd function (child process)", flush=True)
    p = subprocess.run([sys.executable, "-c", "import torch, gradient_tools as gt; gt.autograd_check()"], env=env,
                       capture_output=True, text=True, timeout=240)
    assert p.returncode == 0, p.stdout[-2000:] + p.stderr[-2000:]
    res = json.loads(p.stdout.strip().splitlines()[-1])
    assert res["n"] == 4 and res["max_err"] <= 1e-6, res
