"""The compiled C++ caller of include/hfx.h (tests/c_caller/hfx_c_caller.cpp, built by
hartree-fock_amd/build.py with g++): it links and starts without a GPU, and on the GPU its
results equal those of the same calls made through ctypes and of the oracle."""
import os
import subprocess

import numpy as np
import pytest

import molecules
from hfx import lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "c_caller", "hfx_c_caller")


def test_c_caller_is_built_and_links():
    assert os.path.exists(BIN), "build the engine first (python hartree-fock_amd/build.py)"
    r = subprocess.run([BIN], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2 and "usage" in r.stderr  # loaded libhfx_mi355x.so; no device call made


def _write_input(path, b, D):
    f = b.flat()
    with open(path, "wb") as fh:
        np.asarray([f["n_shells"], f["n_prims"], f["n_aos"], f["n_norm"]], np.int32).tofile(fh)
        for k in ("l", "n_prim", "prim_offset", "ao_offset", "norm_offset"):
            np.asarray(f[k], np.int32).tofile(fh)
        for k in ("centers", "alpha", "coeff", "norm"):
            np.asarray(f[k], np.float64).tofile(fh)
        np.asarray(f["cart"], np.int32).tofile(fh)
        np.asarray(D, np.float64).reshape(-1, order="F").tofile(fh)


@pytest.mark.gpu
def test_c_caller_on_gpu(tmp_path, engine_lib, oracle):
    b = molecules.water_dimer("6-31g**")
    n, nS = b.n_aos, b.n_shells
    rng = np.random.default_rng(3)
    C = rng.standard_normal((n, 10))
    D = C @ C.T * 0.05
    _write_input(tmp_path / "in.bin", b, D)
    print("c caller", n, flush=True)
    r = subprocess.run([BIN, str(tmp_path / "in.bin"), str(tmp_path / "out.bin")], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    raw = np.fromfile(tmp_path / "out.bin", np.float64, count=5 * n * n + nS * nS)
    nq = int(np.fromfile(tmp_path / "out.bin", np.int64, offset=raw.nbytes)[0])
    S, off = raw[:n * n].reshape(n, n, order="F"), n * n
    Q = raw[off:off + nS * nS].reshape(nS, nS, order="F")
    off += nS * nS
    J, K, Jg, Kg = [raw[off + k * n * n: off + (k + 1) * n * n].reshape(n, n, order="F") for k in range(4)]
    g = lib.Engine(engine_lib, b)
    assert np.abs(S - g.overlap()).max() <= 1e-14 and np.array_equal(Q, g.schwarz())
    o = lib.Engine(oracle, b)
    Jo, Ko, _, so = o.jk(2 * D, D, Q=Q)
    assert nq == so["n_quartets"]
    for x, y in ((J, Jo), (K, Ko), (Jg, Jo), (Kg, Ko)):
        assert np.abs(x - y).max() <= 1e-11 * max(1.0, np.abs(y).max())
