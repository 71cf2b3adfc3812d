"""The HIP engine library loads and exports every symbol the C headers declare (no GPU
calls), and the package refuses to run without it."""
import ctypes as C
import os
import re

import pytest

from hfx import lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for h in ("hfx.h", "hfx_scf.h"):
        src = open(os.path.join(ROOT, "include", h)).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        names |= set(re.findall(r"\b(hfx_[a-z_0-9]+)\s*\(", src))
    return names


def test_header_symbols_are_bound():
    assert declared_symbols() <= set(lib.ABI_SYMBOLS)


def test_library_exports_all_symbols():
    assert os.path.exists(lib.LIB_PATH), "build the engine first (python hartree-fock_amd/build.py)"
    d = C.CDLL(lib.LIB_PATH)
    for name in declared_symbols():
        assert hasattr(d, name), name


def test_missing_library_fails_loudly(tmp_path):
    with pytest.raises(lib.HfxError):
        lib.load(str(tmp_path / "nope.so"))


def test_oracle_exports_the_same_abi(oracle):
    for name in declared_symbols():
        assert hasattr(oracle.dll, name), name


def test_status_codes_match_the_header():
    """The Python mirror of the status codes and flags equals the enums of include/hfx.h."""
    text = open(os.path.join(ROOT, "include", "hfx.h")).read()
    for name, value in re.findall(r"\b(HFX_(?:OK|ERR_\w+|JK_\w+|FP64|MIXED))\s*=\s*(\d+)", text):
        assert getattr(lib, name) == int(value), name
