"""Autotuner for the mixed-precision cut.

Quartets whose bound Q(ab) Q(cd) max|dD| falls below the cut run in FP32.  For a given
density the tuner measures, for a ladder of cuts, the change of the two-electron energy
tr(D_a (J - K_a)) against the FP64 build and the build time, and returns the fastest cut whose
energy error stays within the target.
"""
import numpy as np

from . import lib


def two_electron_energy(Da, J, Ka, Db=None, Kb=None):
    if Db is None:
        return float(np.sum(Da * (J - Ka)))
    return 0.5 * float(np.sum(Da * (J - Ka)) + np.sum(Db * (J - Kb)))


def tune_cut(engine, Da, target=1e-7, cuts=(1.0, 0.3, 0.1, 3e-2, 1e-2, 3e-3, 1e-3, 3e-4, 1e-4, 1e-5, 1e-6),
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
