"""hfx_one_electron_noalloc: the allocation-free variant of hfx_overlap / hfx_kinetic / hfx_nuclear
(staging in a static device array), called through the C ABI and compared with the allocating
entry points bit for bit, on two calls in a row, the second with a smaller input."""
import ctypes as C

import numpy as np
import pytest

import molecules
from hfx import lib

pytestmark = pytest.mark.gpu

_f64p = C.POINTER(C.c_double)
_i32p = C.POINTER(C.c_int32)


def _noalloc(g, kind, Z=None, xyz=None, atom_mol=None):
    out = np.zeros(g.n_elems)
    Zc = None if Z is None else np.ascontiguousarray(np.asarray(Z, np.float64))
    Xc = None if xyz is None else np.ascontiguousarray(np.asarray(xyz, np.float64).reshape(-1))
    Ac = None if atom_mol is None else np.ascontiguousarray(np.asarray(atom_mol, np.int32))
    n = 0 if Zc is None else len(Zc)
    rc = g.lib.dll.hfx_one_electron_noalloc(g.h, kind, n, None if Zc is None else Zc.ctypes.data_as(_f64p),
                                            None if Xc is None else Xc.ctypes.data_as(_f64p),
                                            None if Ac is None else Ac.ctypes.data_as(_i32p), out.ctypes.data_as(_f64p))
    return rc, g.unpack(out)


def test_noalloc_matches_the_allocating_calls_twice(engine_lib):
    big, small = molecules.ethane(), molecules.h2o("6-31g**")  # the second call has the smaller input
    for b in (big, small):
        print("noalloc", b.n_aos, flush=True)
        g = lib.Engine(engine_lib, b)
        for kind, ref in ((0, g.overlap()), (1, g.kinetic()), (2, g.nuclear(b.Z, b.xyz))):
            rc, M = _noalloc(g, kind, b.Z, b.xyz)
            assert rc == lib.HFX_OK, engine_lib.dll.hfx_last_error()
            assert np.array_equal(M, ref), (b.n_aos, kind)


def test_noalloc_batched_and_limits(engine_lib):
    mols = [molecules.h2o("6-31g"), molecules.water_dimer("6-31g")]
    g = lib.Engine(engine_lib, mols)
    Z = np.concatenate([m.Z for m in mols])
    xyz = np.concatenate([m.xyz for m in mols])
    am = np.concatenate([np.full(len(m.Z), i) for i, m in enumerate(mols)])
    ref = g.nuclear(Z, xyz, atom_mol=am)
    rc, V = _noalloc(g, 2, Z, xyz, am)
    assert rc == lib.HFX_OK and all(np.array_equal(x, y) for x, y in zip(V, ref))
    # atoms not grouped by molecule, too many point charges, a bad kind: clear refusals
    rc, _ = _noalloc(g, 2, Z[::-1], xyz[::-1], am[::-1])
    assert rc == lib.HFX_ERR_INVALID and b"grouped" in engine_lib.dll.hfx_last_error()
    g1 = lib.Engine(engine_lib, mols[0])
    rc, _ = _noalloc(g1, 2, np.ones(5000), np.zeros((5000, 3)))
    assert rc == lib.HFX_ERR_INVALID and b"limit" in engine_lib.dll.hfx_last_error()
    rc, _ = _noalloc(g1, 3)
    assert rc == lib.HFX_ERR_INVALID
