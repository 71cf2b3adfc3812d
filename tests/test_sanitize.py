"""The CPU oracle and the package's host SCF driver (hfx_scf.cpp, linalg.hpp) under
AddressSanitizer + UndefinedBehaviorSanitizer: oracle/sanitize_check (make -C oracle sanitize)
runs every host path once (integrals, J/K, direct / damped / level-shifted / conventional /
UHF / batched SCF, gradient) and must finish with no sanitizer report."""
import os
import subprocess

import numpy as np
import pytest

import molecules

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ODIR = os.path.join(ROOT, "oracle")
BIN = os.path.join(ODIR, "sanitize_check")


@pytest.fixture(scope="module")
def sanitized():
    subprocess.check_call(["make", "-s", "-C", ODIR, "sanitize"])
    return BIN


def _write(path, b, charge=0, mult=1):
    f = b.flat()
    with open(path, "wb") as fh:
        np.asarray([f["n_shells"], f["n_prims"], f["n_aos"], f["n_norm"]], np.int32).tofile(fh)
        for k in ("l", "n_prim", "prim_offset", "ao_offset", "norm_offset"):
            np.asarray(f[k], np.int32).tofile(fh)
        for k in ("centers", "alpha", "coeff", "norm"):
            np.asarray(f[k], np.float64).tofile(fh)
        np.asarray(f["cart"], np.int32).tofile(fh)
        np.asarray([len(b.Z)], np.int32).tofile(fh)
        np.asarray(b.Z, np.float64).tofile(fh)
        np.asarray(b.xyz, np.float64).reshape(-1).tofile(fh)
        np.asarray([charge, mult], np.int32).tofile(fh)


@pytest.mark.parametrize("mol", ["h2o_631gss", "oh_cation_triplet"])
def test_host_code_under_asan_ubsan(sanitized, tmp_path, mol):
    if mol == "h2o_631gss":
        b, ch, mu = molecules.h2o("6-31g**"), 0, 1
    else:
        b, ch, mu = molecules.oh_radical("6-31g"), 1, 3
    _write(tmp_path / "in.bin", b, ch, mu)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1",
               OMP_NUM_THREADS="2")
    r = subprocess.run([sanitized, str(tmp_path / "in.bin")], capture_output=True, text=True, timeout=600, env=env)
    report = r.stderr
    assert "ERROR: AddressSanitizer" not in report and "runtime error:" not in report, report[-4000:]
    assert "LeakSanitizer" not in report, report[-4000:]
    assert r.returncode == 0, (r.stdout, report[-4000:])
    assert "direct" in r.stdout
