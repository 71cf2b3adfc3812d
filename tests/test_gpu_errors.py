"""Error behaviour of the HIP engine's C ABI: every failure comes back as a status code with a
message (lib.HfxError here), never as a crash, and leaves the engine usable."""
import copy
import ctypes as C

import numpy as np
import pytest

import molecules
from hfx import lib
from hfx.basis import Shell, _norm_factors

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def water(engine_lib):
    b = molecules.h2o("6-31g**")
    return b, lib.Engine(engine_lib, b)


def test_nonfinite_inputs_are_rejected(water):
    b, g = water
    n = b.n_aos
    D = np.eye(n) * 0.1
    bad = D.copy()
    bad[3, 2] = np.nan
    with pytest.raises(lib.HfxError, match="non-finite"):
        g.jk(bad, D)
    bad[3, 2] = np.inf
    with pytest.raises(lib.HfxError, match="non-finite"):
        g.jk(D, D, bad)
    Q = g.schwarz()
    Qb = Q.copy()
    Qb[0, 0] = np.nan
    with pytest.raises(lib.HfxError, match="non-finite"):
        g.jk(D, D, Q=Qb)
    # the engine is still usable afterwards
    J, K, _, st = g.jk(2 * D, D)
    assert st["n_quartets"] > 0 and np.isfinite(J).all() and np.isfinite(K).all()


def test_bad_arguments_are_rejected(water):
    b, g = water
    D = np.eye(b.n_aos) * 0.1
    with pytest.raises(lib.HfxError, match="precision"):
        g.jk(D, D, precision=7)
    nS = b.n_shells
    offs = np.zeros(1, np.int64)
    out = np.zeros(16)
    for q in ([0, 0, 0, nS], [-1, 0, 0, 0]):
        abcd = np.asarray(q, np.int32)
        rc = g.lib.dll.hfx_eri_quartets(g.h, 1, abcd.ctypes.data_as(lib._i32p),
                                       offs.ctypes.data_as(C.POINTER(C.c_int64)), out.ctypes.data_as(lib._f64p))
        assert rc == lib.HFX_ERR_INVALID
        assert b"out of range" in g.lib.dll.hfx_last_error()
    with pytest.raises(lib.HfxError):
        g.store_eris(-1.0)
    assert g.eri_quartets([(0, 0, 0, 0)])[0].shape == (1, 1, 1, 1)


def test_f_shells_are_unsupported(engine_lib):
    b = copy.deepcopy(molecules.h2o("6-31g**"))
    sh = b.shells[-1]
    exps = [0.8]
    b.shells[-1] = Shell(l=3, exponents=exps, coefficients=[1.0], atom=sh.atom, center=sh.center,
                         norm=_norm_factors(3, exps))
    b._flat = None
    with pytest.raises(lib.HfxError, match="angular momentum"):
        lib.Engine(engine_lib, b)
