"""ctypes access to the CPU oracle library (test infrastructure only).

Only tests/, the smoke check and the benchmark's CPU-baseline leg use this module.  The
oracle exports the same C ABI as the HIP engine (include/hfx.h), so the package's ``Library``
and ``Engine`` classes drive it unchanged; the oracle-only entry points are bound here.
"""
import ctypes as C
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(os.path.dirname(HERE), "hartree-fock_amd"))

from hfx import lib as _lib  # noqa: E402

DET_PATH = os.path.join(HERE, "liboracle_hfx.so")
FAST_PATH = os.path.join(HERE, "liboracle_hfx_fast.so")

_f64p = C.POINTER(C.c_double)


def load(fast=False):
    L = _lib.Library(FAST_PATH if fast else DET_PATH)
    d = L.dll
    d.oracle_boys.argtypes = [C.c_int32, C.c_double, _f64p]
    d.oracle_boys.restype = None
    d.oracle_boys_series.argtypes = [C.c_int32, C.c_double]
    d.oracle_boys_series.restype = C.c_double
    d.oracle_hermite.argtypes = [C.c_int32, C.c_int32, C.c_double, C.c_double, C.c_double, _f64p]
    d.oracle_hermite.restype = None
    d.oracle_hermite_index.argtypes = [C.c_int32] * 5
    d.oracle_hermite_index.restype = C.c_int32
    d.oracle_hermite_r.argtypes = [C.c_int32, C.c_double, _f64p, _f64p]
    d.oracle_hermite_r.restype = None
    d.oracle_model_flop.argtypes = [C.c_int32] * 4
    d.oracle_model_flop.restype = C.c_double
    d.oracle_jk_range.argtypes = [C.c_void_p, C.POINTER(_lib.HfxJkArgs), C.c_int32, C.c_int32, _f64p, _f64p, _f64p,
                                  C.POINTER(_lib.HfxJkStats)]
    return L


def build():
    """Compile both oracle libraries (make in this directory)."""
    import subprocess
    subprocess.check_call(["make", "-s", "-C", HERE, "-j4"])
