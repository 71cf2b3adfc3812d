"""Host side of the MI355X direct-SCF Fock-build engine: basis layout, workload geometries and
the ctypes binding of the C ABI (include/hfx.h, include/hfx_scf.h)."""
from . import basis, geometry, lib  # noqa: F401
from .basis import Basis  # noqa: F401
