"""Minimal HIP runtime access through ctypes: device buffers for device-resident J/K inputs.

Only allocation, copies and synchronization; all compute goes through the engine library.
"""
import ctypes as C

import numpy as np

_hip = None


def _lib():
    global _hip
    if _hip is None:
        # The runtime functions are looked up through the engine library's handle (dlsym
        # searches the library's own dependencies), so these buffers live in the same HIP
        # runtime instance as the engine, whichever other copies (torch's) the process maps.
        from . import lib
        _hip = lib.load().dll
        _hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        _hip.hipFree.argtypes = [C.c_void_p]
        _hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        _hip.hipDeviceSynchronize.argtypes = []
        _hip.hipSetDevice.argtypes = [C.c_int]
        _hip.hipGetDeviceCount.argtypes = [C.POINTER(C.c_int)]
        _hip.hipStreamCreateWithFlags.argtypes = [C.POINTER(C.c_void_p), C.c_uint]
        _hip.hipStreamDestroy.argtypes = [C.c_void_p]
        _hip.hipStreamSynchronize.argtypes = [C.c_void_p]
        _hip.hipStreamBeginCapture.argtypes = [C.c_void_p, C.c_int]
        _hip.hipStreamEndCapture.argtypes = [C.c_void_p, C.POINTER(C.c_void_p)]
        _hip.hipGraphInstantiate.argtypes = [C.POINTER(C.c_void_p), C.c_void_p, C.c_void_p, C.c_char_p, C.c_size_t]
        _hip.hipGraphLaunch.argtypes = [C.c_void_p, C.c_void_p]
        _hip.hipGraphExecDestroy.argtypes = [C.c_void_p]
        _hip.hipGraphDestroy.argtypes = [C.c_void_p]
    return _hip


H2D, D2H, D2D = 1, 2, 3


def device_count():
    n = C.c_int(0)
    if _lib().hipGetDeviceCount(C.byref(n)) != 0:
        return 0
    return n.value


def synchronize():
    rc = _lib().hipDeviceSynchronize()
    if rc != 0:
        raise RuntimeError("hipDeviceSynchronize failed: %d" % rc)


class DeviceArray:
    """A float64 buffer in device memory."""

    def __init__(self, n):
        self.n = int(n)
        p = C.c_void_p()
        rc = _lib().hipMalloc(C.byref(p), max(1, self.n) * 8)
        if rc != 0:
            raise MemoryError("hipMalloc(%d) failed: %d" % (self.n * 8, rc))
        self.ptr = p

    @classmethod
    def from_host(cls, a):
        a = np.ascontiguousarray(np.asarray(a, np.float64).reshape(-1, order="F"))
        d = cls(a.size)
        d.upload(a)
        return d

    def upload(self, a):
        a = np.ascontiguousarray(a, np.float64)
        assert a.size == self.n
        rc = _lib().hipMemcpy(self.ptr, a.ctypes.data_as(C.c_void_p), self.n * 8, H2D)
        if rc != 0:
            raise RuntimeError("hipMemcpy H2D failed: %d" % rc)

    def download(self):
        out = np.empty(self.n)
        rc = _lib().hipMemcpy(out.ctypes.data_as(C.c_void_p), self.ptr, self.n * 8, D2H)
        if rc != 0:
            raise RuntimeError("hipMemcpy D2H failed: %d" % rc)
        return out

    def copy_from(self, other):
        """Device-to-device copy of another buffer of the same size (the runtime's copy kernel)."""
        assert other.n == self.n
        rc = _lib().hipMemcpy(self.ptr, other.ptr, self.n * 8, D2D)
        if rc != 0:
            raise RuntimeError("hipMemcpy D2D failed: %d" % rc)

    def free(self):
        if self.ptr:
            _lib().hipFree(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def _ok(rc, what):
    if rc != 0:
        raise RuntimeError("%s failed: %d" % (what, rc))


class Stream:
    """A non-blocking HIP stream at the default priority (for hfx_jk_async callers)."""

    def __init__(self):
        self.h = C.c_void_p()
        _ok(_lib().hipStreamCreateWithFlags(C.byref(self.h), 1), "hipStreamCreateWithFlags")

    def synchronize(self):
        _ok(_lib().hipStreamSynchronize(self.h), "hipStreamSynchronize")

    def begin_capture(self):
        _ok(_lib().hipStreamBeginCapture(self.h, 1), "hipStreamBeginCapture")  # thread-local mode

    def end_capture(self):
        """Ends the capture and returns the instantiated graph."""
        g = C.c_void_p()
        _ok(_lib().hipStreamEndCapture(self.h, C.byref(g)), "hipStreamEndCapture")
        return Graph(g)

    def destroy(self):
        if self.h:
            _lib().hipStreamDestroy(self.h)
            self.h = None


class Graph:
    """An instantiated graph captured from a Stream."""

    def __init__(self, g):
        self.g = g
        self.x = C.c_void_p()
        _ok(_lib().hipGraphInstantiate(C.byref(self.x), g, None, None, 0), "hipGraphInstantiate")

    def launch(self, stream):
        _ok(_lib().hipGraphLaunch(self.x, stream.h), "hipGraphLaunch")

    def destroy(self):
        if self.x:
            _lib().hipGraphExecDestroy(self.x)
            _lib().hipGraphDestroy(self.g)
            self.x = None
