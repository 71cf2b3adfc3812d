"""Independent integral evaluation for pinning the oracle (test infrastructure).

Obara-Saika recurrences (Obara & Saika, J. Chem. Phys. 84, 3963 (1986)) for primitive
Cartesian Gaussian ERIs, with the Boys function from mpmath at 30 digits.  Nothing here
shares code with the McMurchie-Davidson oracle or engine, so agreement pins both.
"""
from functools import lru_cache

import mpmath
import numpy as np

mpmath.mp.dps = 30


@lru_cache(maxsize=None)
def boys(m, T):
    """F_m(T) = int_0^1 t^(2m) exp(-T t^2) dt."""
    T = mpmath.mpf(T)
    if T == 0:
        return 1.0 / (2 * m + 1)
    return float(mpmath.gammainc(m + mpmath.mpf(0.5), 0, T) / (2 * T ** (m + mpmath.mpf(0.5))))


def _prim_eri(a, b, c, d, A, B, C, D, al, be, ga, de):
    p, q = al + be, ga + de
    P = (al * A + be * B) / p
    Q = (ga * C + de * D) / q
    rho = p * q / (p + q)
    PA, PB, QC, QD, PQ = P - A, P - B, Q - C, Q - D, P - Q
    T = rho * float(PQ @ PQ)
    Kab = np.exp(-al * be / p * float((A - B) @ (A - B)))
    Kcd = np.exp(-ga * de / q * float((C - D) @ (C - D)))
    pre = 2 * np.pi ** 2.5 / (p * q * np.sqrt(p + q)) * Kab * Kcd

    @lru_cache(maxsize=None)
    def I(a, b, c, d, m):
        if min(a + b + c + d) < 0:
            return 0.0
        for idx, v in enumerate((a, b, c, d)):
            for i in range(3):
                if v[i] > 0:
                    break
            else:
                continue
            break
        else:
            return pre * boys(m, T)
        e = [0, 0, 0]
        e[i] = 1
        e = tuple(e)
        dec = lambda v: tuple(x - y for x, y in zip(v, e))
        if idx == 0:
            a1 = dec(a)
            r = PA[i] * I(a1, b, c, d, m) - rho / p * PQ[i] * I(a1, b, c, d, m + 1)
            if a1[i] > 0:
                r += a1[i] / (2 * p) * (I(dec(a1), b, c, d, m) - rho / p * I(dec(a1), b, c, d, m + 1))
            if b[i] > 0:
                r += b[i] / (2 * p) * (I(a1, dec(b), c, d, m) - rho / p * I(a1, dec(b), c, d, m + 1))
            if c[i] > 0:
                r += c[i] / (2 * (p + q)) * I(a1, b, dec(c), d, m + 1)
            if d[i] > 0:
                r += d[i] / (2 * (p + q)) * I(a1, b, c, dec(d), m + 1)
            return r
        if idx == 1:
            b1 = dec(b)
            r = PB[i] * I(a, b1, c, d, m) - rho / p * PQ[i] * I(a, b1, c, d, m + 1)
            if b1[i] > 0:
                r += b1[i] / (2 * p) * (I(a, dec(b1), c, d, m) - rho / p * I(a, dec(b1), c, d, m + 1))
            if c[i] > 0:
                r += c[i] / (2 * (p + q)) * I(a, b1, dec(c), d, m + 1)
            if d[i] > 0:
                r += d[i] / (2 * (p + q)) * I(a, b1, c, dec(d), m + 1)
            return r
        if idx == 2:
            c1 = dec(c)
            r = QC[i] * I(a, b, c1, d, m) + rho / q * PQ[i] * I(a, b, c1, d, m + 1)
            if c1[i] > 0:
                r += c1[i] / (2 * q) * (I(a, b, dec(c1), d, m) - rho / q * I(a, b, dec(c1), d, m + 1))
            if d[i] > 0:
                r += d[i] / (2 * q) * (I(a, b, c1, dec(d), m) - rho / q * I(a, b, c1, dec(d), m + 1))
            return r
        d1 = dec(d)
        r = QD[i] * I(a, b, c, d1, m) + rho / q * PQ[i] * I(a, b, c, d1, m + 1)
        if d1[i] > 0:
            r += d1[i] / (2 * q) * (I(a, b, c, dec(d1), m) - rho / q * I(a, b, c, dec(d1), m + 1))
        return r

    return I(tuple(a), tuple(b), tuple(c), tuple(d), 0)


def contracted_eri(basis, sa, sb, sc, sd):
    """Full (naoA, naoB, naoC, naoD) block of normalized contracted ERIs."""
    from hfx.basis import cart_components
    sh = [basis.shells[s] for s in (sa, sb, sc, sd)]
    comps = [cart_components(s.l) for s in sh]
    cen = [np.array(s.center) for s in sh]
    out = np.zeros([len(c) for c in comps])
    for ia, ca in enumerate(comps[0]):
        for ib, cb in enumerate(comps[1]):
            for ic, cc in enumerate(comps[2]):
                for id_, cd in enumerate(comps[3]):
                    v = 0.0
                    for pa in range(sh[0].nprim):
                        for pb in range(sh[1].nprim):
                            for pc in range(sh[2].nprim):
                                for pd in range(sh[3].nprim):
                                    n = (sh[0].norm[pa * sh[0].nao + ia] * sh[1].norm[pb * sh[1].nao + ib] *
                                         sh[2].norm[pc * sh[2].nao + ic] * sh[3].norm[pd * sh[3].nao + id_])
                                    c = (sh[0].coefficients[pa] * sh[1].coefficients[pb] *
                                         sh[2].coefficients[pc] * sh[3].coefficients[pd])
                                    v += n * c * _prim_eri(ca, cb, cc, cd, cen[0], cen[1], cen[2], cen[3],
                                                           sh[0].exponents[pa], sh[1].exponents[pb],
                                                           sh[2].exponents[pc], sh[3].exponents[pd])
                    out[ia, ib, ic, id_] = v
    return out


def dense_eri(engine, basis):
    """Dense N^4 ERI tensor assembled from an engine's hfx_eri_quartets blocks, using the
    canonical entries only (the reference's zero rules) and the 8-fold symmetry."""
    nS, N = basis.n_shells, basis.n_aos
    quartets = [(a, b, c, d) for a in range(nS) for b in range(a + 1) for c in range(nS) for d in range(c + 1)
                if a * (a + 1) // 2 + b >= c * (c + 1) // 2 + d]
    blocks = engine.eri_quartets(quartets)
    off = basis.flat()["ao_offset"]
    G = np.zeros((N, N, N, N))
    for (a, b, c, d), blk in zip(quartets, blocks):
        for A in range(blk.shape[0]):
            for B in range(blk.shape[1]):
                for Cc in range(blk.shape[2]):
                    for D in range(blk.shape[3]):
                        i, j, k, l = off[a] + A, off[b] + B, off[c] + Cc, off[d] + D
                        if j > i or l > k:
                            continue
                        if a == c and b == d and i * (i + 1) // 2 + j < k * (k + 1) // 2 + l:
                            continue
                        v = blk[A, B, Cc, D]
                        for (w, x, y, z) in ((i, j, k, l), (j, i, k, l), (i, j, l, k), (j, i, l, k),
                                             (k, l, i, j), (l, k, i, j), (k, l, j, i), (l, k, j, i)):
                            G[w, x, y, z] = v
    return G
