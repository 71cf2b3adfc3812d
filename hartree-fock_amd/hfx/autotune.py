"""Autotuner for the mixed-precision cut.

Quartets whose bound Q(ab) Q(cd) max|dD| falls below the cut run in FP32.  For a given
density the tuner measures, for a ladder of cuts, the change of the two-electron energy
tr(D_a (J - K_a)) against the FP64 build and the build time, and returns the largest cut
whose energy error stays within the target.
"""
import time

import numpy as np

from . import lib


def two_electron_energy(Da, J, Ka, Db=None, Kb=None):
    if Db is None:
        return float(np.sum(Da * (J - Ka)))
    return 0.5 * float(np.sum(Da * (J - Ka)) + np.sum(Db * (J - Kb)))


def tune_cut(engine, Da, target=1e-7, cuts=(1e-2, 3e-3, 1e-3, 3e-4, 1e-4, 3e-5, 1e-5, 1e-6), Q=None,
             schwarz_thresh=1e-10, density_thresh=1e-10):
    """Returns (cut, table) with table = [(cut, |dE|, seconds, n_fp32, n_quartets)]."""
    Dt = 2 * Da
    t = time.perf_counter()
    J, K, _, st = engine.jk(Dt, Da, Q=Q, schwarz_thresh=schwarz_thresh, density_thresh=density_thresh)
    t64 = time.perf_counter() - t
    E64 = two_electron_energy(Da, J, K)
    table = [(0.0, 0.0, t64, 0, st["n_quartets"])]
    best = 0.0
    for cut in sorted(cuts, reverse=True):
        t = time.perf_counter()
        Jm, Km, _, sm = engine.jk(Dt, Da, Q=Q, schwarz_thresh=schwarz_thresh, density_thresh=density_thresh,
                                  precision=lib.HFX_MIXED, fp32_cut=cut)
        dt = time.perf_counter() - t
        dE = abs(two_electron_energy(Da, Jm, Km) - E64)
        table.append((cut, dE, dt, sm["n_fp32"], sm["n_quartets"]))
        if dE <= target and best == 0.0:
            best = cut
    return best, table
