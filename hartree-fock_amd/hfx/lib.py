"""ctypes binding of the engine ABI (include/hfx.h) and the SCF driver (include/hfx_scf.h).

``Library(path)`` binds any shared library that exports the hfx.h symbols.  ``load()``
returns the HIP engine built in this package (``lib/libhfx_mi355x.so``) and raises if it is
missing: there is no CPU fallback behind the product path.
"""
import ctypes as C
import os

import numpy as np

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_DIR, "lib", "libhfx_mi355x.so")

HFX_OK = 0
HFX_ERR_INVALID = 1
HFX_ERR_UNSUPPORTED = 2
HFX_ERR_NOMEM = 3
HFX_ERR_DEVICE = 4
HFX_ERR_NONFINITE = 5
HFX_ERR_NOT_CONVERGED = 6
HFX_FP64 = 0
HFX_MIXED = 1
HFX_JK_DEVICE_PTRS = 1
HFX_JK_USE_GRAPH = 2
HFX_JK_NO_STATS = 4
HFX_JK_CLASS_TIMING = 8
HFX_JK_STORED = 16
HFX_JK_DETERMINISTIC = 32
N_CLASSES = 21

ERRORS = {1: "invalid argument", 2: "unsupported", 3: "out of memory", 4: "device error",
          5: "non-finite input", 6: "not converged"}

_i32p = C.POINTER(C.c_int32)
_f64p = C.POINTER(C.c_double)


class HfxBasis(C.Structure):
    _fields_ = [("n_shells", C.c_int32), ("n_prims", C.c_int32), ("n_aos", C.c_int32), ("n_norm", C.c_int32),
                ("l", _i32p), ("n_prim", _i32p), ("prim_offset", _i32p), ("ao_offset", _i32p),
                ("norm_offset", _i32p), ("centers", _f64p), ("alpha", _f64p), ("coeff", _f64p),
                ("norm", _f64p), ("cart", _i32p)]


class HfxJkArgs(C.Structure):
    _fields_ = [("dD_tot", _f64p), ("dD_alpha", _f64p), ("dD_beta", _f64p), ("Q", _f64p),
                ("schwarz_thresh", C.c_double), ("density_thresh", C.c_double), ("precision", C.c_int32),
                ("fp32_cut", C.c_double), ("flags", C.c_int32)]


class HfxJkStats(C.Structure):
    _fields_ = [("n_quartets", C.c_int64), ("n_fp32", C.c_int64), ("class_count", C.c_int64 * N_CLASSES),
                ("hash_sum", C.c_uint64), ("hash_xor", C.c_uint64), ("model_gflop", C.c_double),
                ("device_ms", C.c_double), ("class_gflop64", C.c_double * N_CLASSES),
                ("class_gflop32", C.c_double * N_CLASSES), ("class_ms64", C.c_double * N_CLASSES),
                ("class_ms32", C.c_double * N_CLASSES), ("exec_gflop", C.c_double),
                ("class_exec_gflop64", C.c_double * N_CLASSES), ("class_exec_gflop32", C.c_double * N_CLASSES)]

    def as_dict(self):
        return dict(n_quartets=self.n_quartets, n_fp32=self.n_fp32, class_count=list(self.class_count),
                    hash_sum=self.hash_sum, hash_xor=self.hash_xor, model_gflop=self.model_gflop,
                    device_ms=self.device_ms, class_gflop64=list(self.class_gflop64),
                    class_gflop32=list(self.class_gflop32), class_ms64=list(self.class_ms64),
                    class_ms32=list(self.class_ms32), exec_gflop=self.exec_gflop,
                    class_exec_gflop64=list(self.class_exec_gflop64),
                    class_exec_gflop32=list(self.class_exec_gflop32))


class HfxGradArgs(C.Structure):
    _fields_ = [("P", _f64p), ("W", _f64p), ("n_atoms", C.c_int32), ("Z", _f64p), ("xyz", _f64p),
                ("shell_atom", _i32p), ("schwarz_thresh", C.c_double), ("density_thresh", C.c_double),
                ("precision", C.c_int32), ("fp32_cut", C.c_double), ("flags", C.c_int32)]


class HfxGradStats(C.Structure):
    _fields_ = [("n_quartets", C.c_int64), ("n_fp32", C.c_int64), ("device_ms", C.c_double),
                ("class_ms", C.c_double * 21), ("class_prims", C.c_int64 * 21)]


class HfxScfOptions(C.Structure):
    _fields_ = [("max_iter", C.c_int32), ("energy_tol", C.c_double), ("density_tol", C.c_double),
                ("use_diis", C.c_int32), ("diis_size", C.c_int32), ("diis_error_tol", C.c_double),
                ("schwarz_thresh", C.c_double), ("density_thresh", C.c_double), ("unrestricted", C.c_int32),
                ("guess_mix", C.c_int32), ("damp", C.c_int32), ("max_damp_iter", C.c_int32),
                ("stop_damp_thresh", C.c_double), ("level_shift", C.c_double), ("lshift_gap_tol", C.c_double),
                ("max_lshift_iter", C.c_int32), ("stop_lshift_thresh", C.c_double), ("precision", C.c_int32),
                ("fp32_cut", C.c_double), ("final_fp64", C.c_int32), ("jk_flags", C.c_int32),
                ("full_builds", C.c_int32), ("direct", C.c_int32),
                ("switch_tol", C.c_double), ("device_linalg", C.c_int32), ("orbital_space", _f64p),
                ("n_orbital_space", C.c_int32)]


class HfxScfResult(C.Structure):
    _fields_ = [("e_total", C.c_double), ("e_electronic", C.c_double), ("e_nuclear", C.c_double),
                ("iterations", C.c_int32), ("converged", C.c_int32), ("last_dE", C.c_double),
                ("last_dD", C.c_double), ("last_diis_err", C.c_double), ("e_total_mixed", C.c_double),
                ("iterations_mixed", C.c_int32), ("time_total_s", C.c_double), ("time_init_s", C.c_double),
                ("time_jk_s", C.c_double), ("time_diag_s", C.c_double), ("n_jk_builds", C.c_int64),
                ("n_quartets_last", C.c_int64),
                ("F_alpha", _f64p), ("D_alpha", _f64p), ("C_alpha", _f64p), ("eps_alpha", _f64p),
                ("F_beta", _f64p), ("D_beta", _f64p), ("C_beta", _f64p), ("eps_beta", _f64p),
                ("h_core", _f64p), ("S", _f64p), ("trace", _f64p), ("trace_cap", C.c_int32),
                ("time_linalg_load_s", C.c_double), ("n_mo", C.c_int32)]


ABI_SYMBOLS = ["hfx_create", "hfx_create_batched", "hfx_destroy", "hfx_sizes", "hfx_overlap", "hfx_kinetic",
               "hfx_nuclear", "hfx_one_electron_noalloc", "hfx_schwarz", "hfx_eri_quartets", "hfx_jk",
               "hfx_jk_async", "hfx_set_launch", "hfx_last_error",
               "hfx_scf_default_options", "hfx_scf_run", "hfx_scf_run_engine", "hfx_scf_run_batched",
               "hfx_scf_run_batched_engine", "hfx_rhf_gradient", "hfx_store_eris", "hfx_release_eris"]


def _p(a, t=_f64p):
    return None if a is None else a.ctypes.data_as(t)


class HfxError(RuntimeError):
    pass


class Library:
    """A loaded library exporting the hfx.h ABI."""

    def __init__(self, path):
        if not os.path.exists(path):
            raise HfxError("engine library not found: %s (build it first)" % path)
        self.path = path
        self.dll = C.CDLL(path)
        d = self.dll
        d.hfx_last_error.restype = C.c_char_p
        d.hfx_create.argtypes = [C.POINTER(HfxBasis), C.c_int, C.POINTER(C.c_void_p)]
        d.hfx_create_batched.argtypes = [C.c_int32, C.POINTER(HfxBasis), C.c_int, C.POINTER(C.c_void_p)]
        d.hfx_destroy.argtypes = [C.c_void_p]
        d.hfx_sizes.argtypes = [C.c_void_p] + [C.POINTER(C.c_int64)] * 4
        for f in ("hfx_overlap", "hfx_kinetic", "hfx_schwarz"):
            getattr(d, f).argtypes = [C.c_void_p, _f64p]
        d.hfx_nuclear.argtypes = [C.c_void_p, C.c_int32, _f64p, _f64p, _i32p, _f64p]
        if hasattr(d, "hfx_one_electron_noalloc"):
            d.hfx_one_electron_noalloc.argtypes = [C.c_void_p, C.c_int32, C.c_int32, _f64p, _f64p, _i32p, _f64p]
        d.hfx_eri_quartets.argtypes = [C.c_void_p, C.c_int64, _i32p, C.POINTER(C.c_int64), _f64p]
        d.hfx_jk.argtypes = [C.c_void_p, C.POINTER(HfxJkArgs), _f64p, _f64p, _f64p, C.POINTER(HfxJkStats)]
        if hasattr(d, "hfx_set_launch"):
            d.hfx_set_launch.argtypes = [C.c_void_p, C.c_int32, C.c_int32, C.c_int32]
        if hasattr(d, "hfx_jk_async"):  # (a library built before this entry point lacks it)
            d.hfx_jk_async.argtypes = [C.c_void_p, C.POINTER(HfxJkArgs), _f64p, _f64p, _f64p, C.c_void_p]
        d.hfx_store_eris.argtypes = [C.c_void_p, _f64p, C.c_double, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]
        d.hfx_release_eris.argtypes = [C.c_void_p]
        d.hfx_rhf_gradient.argtypes = [C.c_void_p, C.POINTER(HfxGradArgs), _f64p, _f64p, C.POINTER(HfxGradStats)]
        d.hfx_scf_default_options.argtypes = [C.POINTER(HfxScfOptions)]
        d.hfx_scf_default_options.restype = None
        d.hfx_scf_run.argtypes = [C.POINTER(HfxBasis), C.c_int32, _f64p, _f64p, C.c_int32, C.c_int32,
                                  C.POINTER(HfxScfOptions), C.c_int, C.POINTER(HfxScfResult)]
        d.hfx_scf_run_engine.argtypes = [C.c_void_p, C.c_int32, _f64p, _f64p, C.c_int32, C.c_int32,
                                         C.POINTER(HfxScfOptions), C.POINTER(HfxScfResult)]
        d.hfx_scf_run_batched.argtypes = [C.c_int32, C.POINTER(HfxBasis), C.c_int32, _f64p, _f64p, _i32p, _i32p,
                                          _i32p, C.POINTER(HfxScfOptions), C.c_int, C.POINTER(HfxScfResult)]
        d.hfx_scf_run_batched_engine.argtypes = [C.c_void_p, C.c_int32, _i32p, C.c_int32, _f64p, _f64p, _i32p,
                                                 _i32p, _i32p, C.POINTER(HfxScfOptions), C.POINTER(HfxScfResult)]

    def check(self, rc, what):
        if rc != HFX_OK:
            msg = self.dll.hfx_last_error()
            raise HfxError("%s failed (%s): %s" % (what, ERRORS.get(rc, rc), msg.decode() if msg else ""))

    def default_options(self, **kw):
        o = HfxScfOptions()
        self.dll.hfx_scf_default_options(C.byref(o))
        for k, v in kw.items():
            if not hasattr(o, k):
                raise KeyError(k)
            setattr(o, k, v)
        return o


class _BasisHolder:
    """Keeps the numpy arrays behind an HfxBasis struct alive."""

    def __init__(self, basis):
        f = basis.flat()
        self.arrays = f
        b = HfxBasis()
        b.n_shells, b.n_prims, b.n_aos, b.n_norm = f["n_shells"], f["n_prims"], f["n_aos"], f["n_norm"]
        for k in ("l", "n_prim", "prim_offset", "ao_offset", "norm_offset", "cart"):
            setattr(b, k, _p(f[k], _i32p))
        for k in ("centers", "alpha", "coeff", "norm"):
            setattr(b, k, _p(f[k]))
        self.struct = b


class Engine:
    """One engine handle (single molecule, or a batch of molecules)."""

    def __init__(self, library, basis, device=0):
        self.lib = library
        bases = basis if isinstance(basis, (list, tuple)) else [basis]
        self.bases = list(bases)
        self._holders = [_BasisHolder(b) for b in self.bases]
        h = C.c_void_p()
        if len(self.bases) == 1 and not isinstance(basis, (list, tuple)):
            rc = library.dll.hfx_create(C.byref(self._holders[0].struct), device, C.byref(h))
        else:
            arr = (HfxBasis * len(self.bases))(*[x.struct for x in self._holders])
            rc = library.dll.hfx_create_batched(len(self.bases), arr, device, C.byref(h))
        library.check(rc, "hfx_create")
        self.h = h
        n = [C.c_int64() for _ in range(4)]
        library.check(library.dll.hfx_sizes(self.h, *[C.byref(x) for x in n]), "hfx_sizes")
        self.n_aos, self.n_shells, self.n_mol, self.n_elems = [x.value for x in n]
        self.sizes = [b.n_aos for b in self.bases]
        self.shell_counts = [b.n_shells for b in self.bases]

    def close(self):
        if self.h:
            self.lib.dll.hfx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # packed N_m x N_m blocks <-> list of matrices
    def unpack(self, flat):
        out, off = [], 0
        for n in self.sizes:
            out.append(flat[off:off + n * n].reshape(n, n, order="F"))
            off += n * n
        return out if len(out) > 1 else out[0]

    def pack(self, mats):
        if isinstance(mats, np.ndarray) and mats.ndim == 2:
            mats = [mats]
        return np.ascontiguousarray(np.concatenate([np.asarray(m, np.float64).reshape(-1, order="F") for m in mats]))

    def _mat(self, fn):
        out = np.zeros(self.n_elems)
        self.lib.check(getattr(self.lib.dll, fn)(self.h, _p(out)), fn)
        return self.unpack(out)

    def overlap(self):
        return self._mat("hfx_overlap")

    def kinetic(self):
        return self._mat("hfx_kinetic")

    def nuclear(self, Z, xyz, atom_mol=None):
        Z = np.ascontiguousarray(np.asarray(Z, np.float64))
        xyz = np.ascontiguousarray(np.asarray(xyz, np.float64).reshape(-1))
        am = None if atom_mol is None else np.ascontiguousarray(np.asarray(atom_mol, np.int32))
        out = np.zeros(self.n_elems)
        self.lib.check(self.lib.dll.hfx_nuclear(self.h, len(Z), _p(Z), _p(xyz), _p(am, _i32p), _p(out)),
                       "hfx_nuclear")
        return self.unpack(out)

    def schwarz(self):
        n = sum(s * s for s in self.shell_counts)
        out = np.zeros(n)
        self.lib.check(self.lib.dll.hfx_schwarz(self.h, _p(out)), "hfx_schwarz")
        res, off = [], 0
        for s in self.shell_counts:
            res.append(out[off:off + s * s].reshape(s, s, order="F"))
            off += s * s
        return res if len(res) > 1 else res[0]

    def eri_quartets(self, quartets):
        """Reference-layout blocks of contracted ERIs for a list of (a,b,c,d) shell quartets."""
        q = np.ascontiguousarray(np.asarray(quartets, np.int32).reshape(-1, 4))
        sh = self.bases[0].shells
        sizes = [sh[a].nao * sh[b].nao * sh[c].nao * sh[d].nao for a, b, c, d in q]
        offs = np.zeros(len(q), np.int64)
        if len(q):
            offs[1:] = np.cumsum(sizes)[:-1]
        out = np.zeros(int(sum(sizes)))
        self.lib.check(self.lib.dll.hfx_eri_quartets(self.h, len(q), _p(q, _i32p), _p(offs, C.POINTER(C.c_int64)),
                                                     _p(out)), "hfx_eri_quartets")
        res = []
        for (a, b, c, d), o, s in zip(q, offs, sizes):
            res.append(out[o:o + s].reshape(sh[a].nao, sh[b].nao, sh[c].nao, sh[d].nao))
        return res

    def store_eris(self, schwarz_thresh=1e-10, Q=None):
        """Compute and keep in HBM the integral blocks of every shell quartet passing the Schwarz
        test (IntegralEngine::electronRepulsionTensor, src/IntegralEngine.cpp:635-692); later
        jk(..., flags=HFX_JK_STORED) builds digest them.  Returns (stored quartets, bytes)."""
        q = None
        if Q is not None:
            qs = Q if isinstance(Q, (list, tuple)) else [Q]
            q = np.ascontiguousarray(np.concatenate([np.asarray(x, np.float64).reshape(-1, order="F") for x in qs]))
        n, b = C.c_int64(0), C.c_int64(0)
        self.lib.check(self.lib.dll.hfx_store_eris(self.h, _p(q), schwarz_thresh, C.byref(n), C.byref(b)),
                       "hfx_store_eris")
        return n.value, b.value

    def set_launch(self, blocks_per_cu=0, chunk_len=0, streams=0):
        """hfx_set_launch: class-kernel launch parameters of this engine (0 keeps a value)."""
        self.lib.check(self.lib.dll.hfx_set_launch(self.h, blocks_per_cu, chunk_len, streams), "hfx_set_launch")

    def release_eris(self):
        self.lib.check(self.lib.dll.hfx_release_eris(self.h), "hfx_release_eris")

    def jk(self, dD_tot, dD_alpha, dD_beta=None, Q=None, schwarz_thresh=1e-10, density_thresh=1e-10,
           precision=HFX_FP64, fp32_cut=0.0, flags=0):
        dt = self.pack(dD_tot)
        da = self.pack(dD_alpha)
        db = None if dD_beta is None else self.pack(dD_beta)
        q = None
        if Q is not None:
            qs = Q if isinstance(Q, (list, tuple)) else [Q]
            q = np.ascontiguousarray(np.concatenate([np.asarray(x, np.float64).reshape(-1, order="F") for x in qs]))
        a = HfxJkArgs()
        a.dD_tot, a.dD_alpha, a.dD_beta, a.Q = _p(dt), _p(da), _p(db), _p(q)
        a.schwarz_thresh, a.density_thresh = schwarz_thresh, density_thresh
        a.precision, a.fp32_cut, a.flags = precision, fp32_cut, flags
        J = np.zeros(self.n_elems)
        Ka = np.zeros(self.n_elems)
        Kb = np.zeros(self.n_elems) if db is not None else None
        st = HfxJkStats()
        self.lib.check(self.lib.dll.hfx_jk(self.h, C.byref(a), _p(J), _p(Ka), _p(Kb), C.byref(st)), "hfx_jk")
        return self.unpack(J), self.unpack(Ka), (None if Kb is None else self.unpack(Kb)), st.as_dict()


def rhf_gradient(engine, basis, P, W, schwarz_thresh=1e-10, density_thresh=1e-10, precision=HFX_FP64,
                 fp32_cut=0.0, flags=0):
    """RHF nuclear gradient (n_atoms, 3) from the total density P and the energy-weighted
    density W.  Returns (grad, parts, stats) with parts[k] (n_atoms, 3) for k = nuclear
    repulsion, -W.dS, P.d(T+V), two-electron."""
    nat = len(basis.Z)
    Z = np.ascontiguousarray(np.asarray(basis.Z, np.float64))
    xyz = np.ascontiguousarray(basis.xyz.reshape(-1))
    sat = np.ascontiguousarray([sh.atom for sh in basis.shells], np.int32)
    Pf = np.ascontiguousarray(np.asarray(P, np.float64).reshape(-1, order="F"))
    Wf = np.ascontiguousarray(np.asarray(W, np.float64).reshape(-1, order="F"))
    a = HfxGradArgs()
    a.P, a.W, a.n_atoms, a.Z, a.xyz, a.shell_atom = _p(Pf), _p(Wf), nat, _p(Z), _p(xyz), _p(sat, _i32p)
    a.schwarz_thresh, a.density_thresh = schwarz_thresh, density_thresh
    a.precision, a.fp32_cut, a.flags = precision, fp32_cut, flags
    g = np.zeros(3 * nat)
    parts = np.zeros(4 * 3 * nat)
    st = HfxGradStats()
    engine.lib.check(engine.lib.dll.hfx_rhf_gradient(engine.h, C.byref(a), _p(g), _p(parts), C.byref(st)),
                     "hfx_rhf_gradient")
    return g.reshape(nat, 3), parts.reshape(4, nat, 3), dict(n_quartets=st.n_quartets, n_fp32=st.n_fp32,
                                                            device_ms=st.device_ms, class_ms=np.array(st.class_ms),
                                                            class_prims=np.array(st.class_prims))


def energy_weighted_density(C, eps, n_occ):
    """W = 2 sum_{i < n_occ} eps_i C_i C_i^T (RHF)."""
    Co = C[:, :n_occ]
    return 2.0 * (Co * eps[:n_occ]) @ Co.T


def jk_device(engine, dD_tot, dD_alpha, J, Ka, dD_beta=None, Kb=None, Q=None, schwarz_thresh=1e-10,
              density_thresh=1e-10, precision=HFX_FP64, fp32_cut=0.0, flags=0):
    """One build on device-resident buffers (hfx.device.DeviceArray); returns the stats dict."""
    a = HfxJkArgs()
    vp = lambda d: None if d is None else C.cast(d.ptr, _f64p)
    a.dD_tot, a.dD_alpha, a.dD_beta = vp(dD_tot), vp(dD_alpha), vp(dD_beta)
    a.Q = None if Q is None else _p(Q)
    a.schwarz_thresh, a.density_thresh = schwarz_thresh, density_thresh
    a.precision, a.fp32_cut, a.flags = precision, fp32_cut, flags | HFX_JK_DEVICE_PTRS
    st = HfxJkStats()
    engine.lib.check(engine.lib.dll.hfx_jk(engine.h, C.byref(a), vp(J), vp(Ka), vp(Kb), C.byref(st)), "hfx_jk")
    return st.as_dict()


def jk_async(engine, dD_tot, dD_alpha, J, Ka, stream, dD_beta=None, Kb=None, Q=None, schwarz_thresh=1e-10,
             density_thresh=1e-10, precision=HFX_FP64, fp32_cut=0.0, flags=0):
    """hfx_jk_async on device buffers (hfx.device.DeviceArray) and a HIP stream (hfx.device.Stream,
    a c_void_p handle, or None for the null stream): enqueued, not waited for."""
    a = HfxJkArgs()
    vp = lambda d: None if d is None else C.cast(d.ptr, _f64p)
    a.dD_tot, a.dD_alpha, a.dD_beta = vp(dD_tot), vp(dD_alpha), vp(dD_beta)
    a.Q = None if Q is None else _p(Q)
    a.schwarz_thresh, a.density_thresh = schwarz_thresh, density_thresh
    a.precision, a.fp32_cut, a.flags = precision, fp32_cut, flags | HFX_JK_DEVICE_PTRS
    h = getattr(stream, "h", stream)  # hfx.device.Stream, a c_void_p or None
    engine.lib.check(engine.lib.dll.hfx_jk_async(engine.h, C.byref(a), vp(J), vp(Ka), vp(Kb), h), "hfx_jk_async")


def scf_run(library, basis, charge=0, multiplicity=1, options=None, device=0, outputs=True, trace_cap=200,
            engine=None, spherical=False):
    """Run the library's SCF driver (hfx_scf_run, or hfx_scf_run_engine over `engine`); returns a
    dict with energies and (optionally) matrices.  spherical=True runs in the orbital space of
    the spherical-harmonic functions of the shells (Basis.spherical_transform); the matrices
    stay in the Cartesian AO basis."""
    o = options if options is not None else library.default_options()
    r = HfxScfResult()
    n = basis.n_aos
    keep = {}
    if spherical:
        T = np.asfortranarray(basis.spherical_transform())
        keep["_T"] = T
        o2 = HfxScfOptions()
        C.memmove(C.byref(o2), C.byref(o), C.sizeof(o))
        o2.orbital_space = T.ctypes.data_as(_f64p)
        o2.n_orbital_space = T.shape[1]
        o = o2
    if outputs:
        for k in ("F_alpha", "D_alpha", "C_alpha", "F_beta", "D_beta", "C_beta", "h_core", "S"):
            keep[k] = np.zeros(n * n)
            setattr(r, k, _p(keep[k]))
        for k in ("eps_alpha", "eps_beta"):
            keep[k] = np.zeros(n)
            setattr(r, k, _p(keep[k]))
    keep["trace"] = np.zeros(4 * trace_cap)
    r.trace = _p(keep["trace"])
    r.trace_cap = trace_cap
    hb = _BasisHolder(basis)
    Z = np.ascontiguousarray(np.asarray(basis.Z, np.float64))
    xyz = np.ascontiguousarray(basis.xyz.reshape(-1))
    if engine is None:
        rc = library.dll.hfx_scf_run(C.byref(hb.struct), len(Z), _p(Z), _p(xyz), charge, multiplicity, C.byref(o),
                                     device, C.byref(r))
        library.check(rc, "hfx_scf_run")
    else:
        rc = library.dll.hfx_scf_run_engine(engine.h, len(Z), _p(Z), _p(xyz), charge, multiplicity, C.byref(o),
                                            C.byref(r))
        library.check(rc, "hfx_scf_run_engine")
    out = {k: getattr(r, k) for k, _ in HfxScfResult._fields_ if not k.startswith(("F_", "D_", "C_", "eps_", "h_", "S", "trace"))}
    it = r.iterations
    out["trace"] = keep["trace"][:4 * min(it, trace_cap)].reshape(-1, 4)
    for k, v in keep.items():
        if k in ("trace", "_T"):
            continue
        out[k] = v.reshape(n, n, order="F") if v.size == n * n else v
    return out


def scf_run_batched(library, bases, charges=None, multiplicities=None, options=None, device=0, engine=None):
    """Run the batched SCF driver over a list of Basis objects (one per molecule).

    Returns a list of per-molecule dicts (energies, iterations, convergence, timings)."""
    o = options if options is not None else library.default_options()
    n_mol = len(bases)
    res = (HfxScfResult * n_mol)()
    Z = np.ascontiguousarray(np.concatenate([np.asarray(b.Z, np.float64) for b in bases]))
    xyz = np.ascontiguousarray(np.concatenate([b.xyz.reshape(-1) for b in bases]))
    atom_mol = np.ascontiguousarray(np.concatenate([np.full(len(b.Z), m, np.int32) for m, b in enumerate(bases)]))
    ch = None if charges is None else np.ascontiguousarray(charges, np.int32)
    mu = None if multiplicities is None else np.ascontiguousarray(multiplicities, np.int32)
    if engine is None:
        holders = [_BasisHolder(b) for b in bases]
        arr = (HfxBasis * n_mol)(*[x.struct for x in holders])
        rc = library.dll.hfx_scf_run_batched(n_mol, arr, len(Z), _p(Z), _p(xyz), _p(atom_mol, _i32p),
                                             _p(ch, _i32p), _p(mu, _i32p), C.byref(o), device, res)
        library.check(rc, "hfx_scf_run_batched")
    else:
        n_aos = np.ascontiguousarray([b.n_aos for b in bases], np.int32)
        rc = library.dll.hfx_scf_run_batched_engine(engine.h, n_mol, _p(n_aos, _i32p), len(Z), _p(Z), _p(xyz),
                                                    _p(atom_mol, _i32p), _p(ch, _i32p), _p(mu, _i32p),
                                                    C.byref(o), res)
        library.check(rc, "hfx_scf_run_batched_engine")
    keys = [k for k, _ in HfxScfResult._fields_ if not k.startswith(("F_", "D_", "C_", "eps_", "h_", "S", "trace"))]
    return [{k: getattr(r, k) for k in keys} for r in res]


_LIB = None


def load(path=None):
    """The HIP engine library (lib/libhfx_mi355x.so, or the file HFX_LIBRARY names).
    Raises HfxError when it has not been built."""
    global _LIB
    if path is not None:
        return Library(path)
    if _LIB is None:
        _LIB = Library(os.environ.get("HFX_LIBRARY") or LIB_PATH)
    return _LIB
