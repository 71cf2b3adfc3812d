"""hfx_set_launch: the class-kernel launch shape (blocks per CU, work-list chunk length, build
streams) changes only the order of the FP64 accumulations, never which quartets are computed."""
import numpy as np
import pytest

import molecules
from hfx import lib

pytestmark = pytest.mark.gpu


def _density(n, nocc, seed, scale=0.1):
    rng = np.random.default_rng(seed)
    C = rng.standard_normal((n, nocc))
    return C @ C.T * scale


def test_launch_shapes_agree(engine_lib):
    b = molecules.ethane()
    g = lib.Engine(engine_lib, b)
    D = _density(b.n_aos, 9, 21)
    print("launch shapes ethane", b.n_aos, flush=True)
    ref = {}
    for kw in (dict(), dict(precision=lib.HFX_MIXED, fp32_cut=1e-3)):
        ref[kw.get("precision", 0)] = g.jk(2 * D, D, **kw)
    for blocks, chunk, streams in ((1, 256, 1), (4, 65536, 2), (3, 1024, 4), (2, 2048, 3)):
        g.set_launch(blocks, chunk, streams)
        for kw in (dict(), dict(precision=lib.HFX_MIXED, fp32_cut=1e-3)):
            J0, K0, _, s0 = ref[kw.get("precision", 0)]
            for flags in (0, lib.HFX_JK_USE_GRAPH, lib.HFX_JK_USE_GRAPH):
                J, K, _, s = g.jk(2 * D, D, flags=flags, **kw)
                assert (s["n_quartets"], s["hash_sum"]) == (s0["n_quartets"], s0["hash_sum"]), (blocks, chunk, kw)
                assert np.abs(J - J0).max() <= 1e-12 * max(1.0, np.abs(J0).max()), (blocks, chunk, streams, kw)
                assert np.abs(K - K0).max() <= 1e-12 * max(1.0, np.abs(K0).max()), (blocks, chunk, streams, kw)
        Jd, Kd, _, _ = g.jk(2 * D, D, flags=lib.HFX_JK_DETERMINISTIC)
        Jd0, Kd0, _, _ = g.jk(2 * D, D, flags=lib.HFX_JK_DETERMINISTIC)
        assert np.array_equal(Jd, Jd0) and np.array_equal(Kd, Kd0)


def test_launch_shape_refusals(engine_lib):
    g = lib.Engine(engine_lib, molecules.water_dimer())
    for bad in ((17, 0, 0), (-1, 0, 0), (0, 255, 0), (0, 65537, 0), (0, 0, 5), (0, 0, -2)):
        with pytest.raises(RuntimeError):
            g.set_launch(*bad)
    g.set_launch(0, 0, 0)  # keeps everything
