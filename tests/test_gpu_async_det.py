"""Deterministic (fixed-point) J/K and the stream-taking hfx_jk_async, including a build
captured into a hipGraph by the caller, on the HIP engine."""
import numpy as np
import pytest

import molecules
from hfx import device, lib

pytestmark = pytest.mark.gpu

DET = lib.HFX_JK_DETERMINISTIC


def _density(n, nocc, seed, scale=0.1):
    rng = np.random.default_rng(seed)
    C = rng.standard_normal((n, nocc))
    return C @ C.T * scale


@pytest.fixture(scope="module")
def ethane(engine_lib):
    b = molecules.ethane()
    return b, lib.Engine(engine_lib, b)


def test_deterministic_builds_are_bitwise_reproducible(ethane, oracle):
    """Repeated deterministic builds (eager, graph-replayed, mixed, UHF) give identical bits, and
    agree with the ordinary FP64-atomic build and the oracle to rounding."""
    b, g = ethane
    n = b.n_aos
    print("deterministic ethane", n, flush=True)
    D = _density(n, 9, 5)
    J0, K0, _, s0 = g.jk(2 * D, D, flags=DET)
    for flags in (DET, DET, DET | lib.HFX_JK_USE_GRAPH, DET | lib.HFX_JK_USE_GRAPH, DET | lib.HFX_JK_USE_GRAPH):
        J, K, _, s = g.jk(2 * D, D, flags=flags)
        assert np.array_equal(J, J0) and np.array_equal(K, K0), flags
        assert (s["n_quartets"], s["hash_sum"]) == (s0["n_quartets"], s0["hash_sum"])
    Jf, Kf, _, _ = g.jk(2 * D, D)  # FP64 atomics (order-dependent last bits)
    assert np.abs(J0 - Jf).max() <= 1e-12 * max(1.0, np.abs(Jf).max())
    assert np.abs(K0 - Kf).max() <= 1e-12 * max(1.0, np.abs(Kf).max())
    o = lib.Engine(oracle, b)
    Q = o.schwarz()
    Jd, Kd, _, sd = g.jk(2 * D, D, Q=Q, flags=DET)
    Jo, Ko, _, so = o.jk(2 * D, D, Q=Q)
    assert (sd["n_quartets"], sd["hash_sum"], sd["hash_xor"]) == (so["n_quartets"], so["hash_sum"], so["hash_xor"])
    assert np.abs(Jd - Jo).max() <= 1e-11 * max(1.0, np.abs(Jo).max())
    assert np.abs(Kd - Ko).max() <= 1e-11 * max(1.0, np.abs(Ko).max())
    # mixed precision and UHF
    Jm0, Km0, _, _ = g.jk(2 * D, D, precision=lib.HFX_MIXED, fp32_cut=1e-3, flags=DET)
    Jm1, Km1, _, _ = g.jk(2 * D, D, precision=lib.HFX_MIXED, fp32_cut=1e-3, flags=DET | lib.HFX_JK_USE_GRAPH)
    assert np.array_equal(Jm0, Jm1) and np.array_equal(Km0, Km1)
    Da, Db = _density(n, 9, 6), _density(n, 8, 7)
    r0 = g.jk(Da + Db, Da, Db, flags=DET)
    r1 = g.jk(Da + Db, Da, Db, flags=DET)
    for x, y in zip(r0[:3], r1[:3]):
        assert np.array_equal(x, y)


@pytest.mark.parametrize("name", ["C8H18/6-31G", "C8H18/cc-pVDZ", "(H2O)8/6-31G**"])
def test_atomic_and_fixed_point_builds_agree_on_every_class(engine_lib, name):
    """The ordinary build (FP64 atomics, LDS exchange rows with flushes, per-XCD copies) and the
    deterministic one (fixed point, no LDS rows) take different digest paths through every class
    kernel; on molecules with thousands of bra-pair switches they must agree to rounding, FP64 and
    mixed.  The two accumulation modes are separately compiled kernels, so in mixed mode their
    FP32 integrals may differ in the last FP32 bit (about 1e-7 relative of quartets bounded by
    the cut, 1e-3 here): the mixed comparison allows 1e-9 absolute."""
    from hfx import Basis, geometry
    if name.startswith("C8H18"):
        Z, x = geometry.alkane_chain(n_carbon=8, seed=1003)
        b = Basis("6-31g" if "6-31G" in name else "cc-pvdz-expanded", Z, x)
    else:
        Z, x = geometry.water_cluster_grid(shape=(2, 2, 2), spacing=5.6, jitter=0.4, seed=1002)
        b = Basis("6-31g**", Z, x)
    g = lib.Engine(engine_lib, b)
    D = _density(b.n_aos, b.n_electrons() // 2, 11, scale=0.05)
    print("atomic vs fixed", name, b.n_aos, flush=True)
    for kw, tol in ((dict(), 1e-12), (dict(precision=lib.HFX_MIXED, fp32_cut=1e-3), 1e-9)):
        Jd, Kd, _, _ = g.jk(2 * D, D, flags=DET, **kw)
        for flags in (0, lib.HFX_JK_USE_GRAPH, lib.HFX_JK_CLASS_TIMING):
            J, K, _, _ = g.jk(2 * D, D, flags=flags, **kw)
            assert np.abs(J - Jd).max() <= tol * max(1.0, np.abs(Jd).max()), (kw, flags)
            assert np.abs(K - Kd).max() <= tol * max(1.0, np.abs(Kd).max()), (kw, flags)


def test_deterministic_stored_builds(engine_lib):
    b = molecules.water_dimer("6-31g**")
    g = lib.Engine(engine_lib, b)
    D = _density(b.n_aos, 10, 9)
    g.store_eris(1e-12)
    fl = lib.HFX_JK_STORED | DET
    J0, K0, _, _ = g.jk(2 * D, D, schwarz_thresh=1e-12, flags=fl)
    for flags in (fl, fl | lib.HFX_JK_USE_GRAPH, fl | lib.HFX_JK_USE_GRAPH):
        J, K, _, _ = g.jk(2 * D, D, schwarz_thresh=1e-12, flags=flags)
        assert np.array_equal(J, J0) and np.array_equal(K, K0)
    g.release_eris()


def _bufs(n, D):
    return (device.DeviceArray.from_host(2 * D), device.DeviceArray.from_host(D), device.DeviceArray(n * n),
            device.DeviceArray(n * n))


def test_async_on_caller_stream_matches_blocking_build(ethane):
    b, g = ethane
    n = b.n_aos
    print("async eager", n, flush=True)
    st = device.Stream()
    try:
        for rep, flags in enumerate((DET, DET | lib.HFX_JK_USE_GRAPH, DET)):
            D = _density(n, 9, 40 + rep)
            Je, Ke, _, _ = g.jk(2 * D, D, flags=DET)
            dt, da, J, K = _bufs(n, D)
            lib.jk_async(g, dt, da, J, K, st, flags=flags)
            st.synchronize()
            Ja, Ka = J.download().reshape(n, n, order="F"), K.download().reshape(n, n, order="F")
            assert np.array_equal(Ja, Je) and np.array_equal(Ka, Ke), rep
            # a blocking build right after the async one is ordered behind it
            lib.jk_async(g, dt, da, J, K, st, flags=DET)
            Je2, Ke2, _, _ = g.jk(2 * D, D, flags=DET)
            assert np.array_equal(Je2, Je)
            st.synchronize()
            for x in (dt, da, J, K):
                x.free()
    finally:
        st.destroy()


def test_async_build_captured_by_the_caller(ethane):
    """The caller captures hfx_jk_async into its own hipGraph (as it would a whole SCF iteration)
    and replays it for several densities, with eager builds and host synchronizations between
    the replays; every replay equals the blocking build of its density bit for bit."""
    b, g = ethane
    n = b.n_aos
    print("async captured", n, flush=True)
    g.jk(*(lambda D: (2 * D, D))(_density(n, 9, 1)), flags=DET)  # sorts the lists for the engine's Q
    st = device.Stream()
    D0 = _density(n, 9, 60)
    dt, da, J, K = _bufs(n, D0)
    try:
        st.begin_capture()
        try:
            lib.jk_async(g, dt, da, J, K, st, flags=DET)
        finally:
            graph = st.end_capture()
        for rep in range(4):
            D = _density(n, 9, 60 + rep)
            dt.upload(np.ascontiguousarray((2 * D).reshape(-1, order="F")))
            da.upload(np.ascontiguousarray(D.reshape(-1, order="F")))
            graph.launch(st)
            st.synchronize()
            Jr, Kr = J.download().reshape(n, n, order="F"), K.download().reshape(n, n, order="F")
            Je, Ke, _, _ = g.jk(2 * D, D, flags=DET)  # eager work between the replays
            assert np.array_equal(Jr, Je) and np.array_equal(Kr, Ke), rep
        graph.destroy()
    finally:
        for x in (dt, da, J, K):
            x.free()
        st.destroy()


def test_async_refusals(ethane):
    b, g = ethane
    n = b.n_aos
    D = _density(n, 9, 70)
    dt, da, J, K = _bufs(n, D)
    st = device.Stream()
    try:
        a = lib.HfxJkArgs()
        import ctypes as C
        a.dD_tot, a.dD_alpha = C.cast(dt.ptr, C.POINTER(C.c_double)), C.cast(da.ptr, C.POINTER(C.c_double))
        a.schwarz_thresh, a.density_thresh, a.flags = 1e-10, 1e-10, 0  # host-pointer flag missing
        rc = g.lib.dll.hfx_jk_async(g.h, C.byref(a), C.cast(J.ptr, C.POINTER(C.c_double)),
                                    C.cast(K.ptr, C.POINTER(C.c_double)), None, st.h)
        assert rc == lib.HFX_ERR_INVALID
        with pytest.raises(lib.HfxError):
            lib.jk_async(g, dt, da, J, K, st, flags=lib.HFX_JK_CLASS_TIMING)
        # a graph launch and a re-sort are refused while the caller's stream is capturing
        Q2 = g.schwarz() * 1.0000001
        st.begin_capture()
        try:
            with pytest.raises(lib.HfxError):
                lib.jk_async(g, dt, da, J, K, st, flags=lib.HFX_JK_USE_GRAPH)
            with pytest.raises(lib.HfxError):
                lib.jk_async(g, dt, da, J, K, st, Q=np.ascontiguousarray(Q2.reshape(-1, order="F")))
        finally:
            st.end_capture().destroy()
        st.synchronize()
    finally:
        for x in (dt, da, J, K):
            x.free()
        st.destroy()
