"""Autotuner for the mixed-precision cut and the class kernels' launch parameters.

tune_cut: quartets whose bound Q(ab) Q(cd) max|dD| falls below the cut run in FP32.  For a given
density the tuner measures, for a ladder of cuts, the change of the two-electron energy
tr(D_a (J - K_a)) against the FP64 build and the build time, and returns the fastest cut whose
energy error stays within the target.

tune_launch: the persistent blocks per CU of each class launch, the run of list entries a block
claims at a time, and the number of streams a build is spread over (hfx_set_launch), timed on
graph-replayed device-resident builds; the fastest median wins.  The reference's analogue is the
fixed OpenMP schedule(dynamic,32) of its quartet loop (src/SCF.cpp:592).
"""
import itertools

import numpy as np

from . import device, lib


def two_electron_energy(Da, J, Ka, Db=None, Kb=None):
    if Db is None:
        return float(np.sum(Da * (J - Ka)))
    return 0.5 * float(np.sum(Da * (J - Ka)) + np.sum(Db * (J - Kb)))


def tune_cut(engine, Da, target=5e-7, cuts=(1.0, 0.3, 0.1, 3e-2, 1e-2, 3e-3, 1e-3, 3e-4, 1e-4, 1e-5, 1e-6),
             Q=None, schwarz_thresh=1e-10, density_thresh=1e-10, repeats=3):
    """Returns (cut, table) with table = [(cut, |dE|, seconds, n_fp32, n_quartets)].

    Every cut's build is run `repeats` times; its time is the fastest build's device time (HIP
    events around the build).  Among the cuts whose |dE| meets the target, the fastest wins
    (ties go to the larger cut); 0.0 (all FP64) when none does."""
    Dt = 2 * Da
    kw = dict(Q=Q, schwarz_thresh=schwarz_thresh, density_thresh=density_thresh)
    J, K, _, st = engine.jk(Dt, Da, **kw)
    t64 = min([st["device_ms"]] + [engine.jk(Dt, Da, **kw)[3]["device_ms"] for _ in range(repeats - 1)]) * 1e-3
    E64 = two_electron_energy(Da, J, K)
    table = [(0.0, 0.0, t64, 0, st["n_quartets"])]
    best, best_t = 0.0, None
    for cut in sorted(cuts, reverse=True):
        mk = dict(kw, precision=lib.HFX_MIXED, fp32_cut=cut)
        Jm, Km, _, sm = engine.jk(Dt, Da, **mk)
        dt = min([sm["device_ms"]] + [engine.jk(Dt, Da, **mk)[3]["device_ms"] for _ in range(repeats - 1)]) * 1e-3
        dE = abs(two_electron_energy(Da, Jm, Km) - E64)
        table.append((cut, dE, dt, sm["n_fp32"], sm["n_quartets"]))
        if dE <= target and (best_t is None or dt < best_t * 0.99):
            best, best_t = cut, dt
    return best, table


def tune_launch(engine, Da, Q=None, precision=lib.HFX_FP64, cut=0.0, blocks=(1, 2, 3, 4), chunks=(1024, 2048, 4096),
                streams=(2, 4), repeats=5, schwarz_thresh=1e-10, density_thresh=1e-10):
    """Times graph-replayed builds on device buffers for every (blocks per CU, chunk length,
    streams) combination, leaves the engine on the fastest, and returns (best, table) with
    table = [((blocks, chunk, streams), median device ms)]."""
    n = Da.shape[0]
    bufs = (device.DeviceArray.from_host(2 * Da), device.DeviceArray.from_host(Da), device.DeviceArray(n * n),
            device.DeviceArray(n * n))
    q = None if Q is None else np.ascontiguousarray(np.asarray(Q, np.float64).reshape(-1, order="F"))
    kw = dict(Q=q, schwarz_thresh=schwarz_thresh, density_thresh=density_thresh, precision=precision, fp32_cut=cut,
              flags=lib.HFX_JK_USE_GRAPH | lib.HFX_JK_NO_STATS)
    table = []
    try:
        for b, c, s in itertools.product(blocks, chunks, streams):
            engine.set_launch(b, c, s)
            lib.jk_device(engine, *bufs, **kw)  # capture
            ms = [lib.jk_device(engine, *bufs, **kw)["device_ms"] for _ in range(repeats)]
            table.append(((b, c, s), float(np.median(ms))))
    finally:
        for x in bufs:
            x.free()
    best = min(table, key=lambda r: r[1])[0]
    engine.set_launch(*best)
    return best, table
