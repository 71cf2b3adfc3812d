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


# ---------------------------------------------------------------------------------------
# One-electron integrals by Obara-Saika recurrences (Obara & Saika 1986; the forms of
# Helgaker, Jorgensen & Olsen, Molecular Electronic-Structure Theory, sec. 9.3 and 9.10):
#   overlap   S_{i+1,j} = X_PA S_ij + (i S_{i-1,j} + j S_{i,j-1}) / 2p,  S_00 = sqrt(pi/p) exp(-mu X_AB^2)
#   kinetic   T_{i+1,j} = X_PA T_ij + (i T_{i-1,j} + j T_{i,j-1}) / 2p + (b/p)(2a S_{i+1,j} - i S_{i-1,j})
#             T_{i,j+1} = X_PB T_ij + (i T_{i-1,j} + j T_{i,j-1}) / 2p + (a/p)(2b S_{i,j+1} - j S_{i,j-1})
#             T_00 = (a - 2a^2 (X_PA^2 + 1/2p)) S_00,  T = Tx Sy Sz + Sx Ty Sz + Sx Sy Tz
#   nuclear   Theta^m_{a+1_i,b} = PA_i Theta^m - PC_i Theta^{m+1}
#                                 + a_i/2p (Theta^m_{a-1_i} - Theta^{m+1}_{a-1_i})
#                                 + b_i/2p (Theta^m_{b-1_i} - Theta^{m+1}_{b-1_i})   (same on b with PB)
#             Theta^m_00 = (2 pi / p) K_ab F_m(p R_PC^2),  V = -sum_C Z_C Theta^0
# None of this uses Hermite expansion coefficients, the reference's Laplacian-on-b form of the
# kinetic integral or its Hermite R recursion, so agreement pins those parts of the oracle.
# ---------------------------------------------------------------------------------------
def _os_overlap_1d(la, lb, PA, PB, p, s00):
    S = np.zeros((la + 1, lb + 3))
    S[0, 0] = s00
    h = 1.0 / (2 * p)
    for i in range(la + 1):
        for j in range(lb + 3):
            if i == 0 and j == 0:
                continue
            if i > 0:
                v = PA * S[i - 1, j]
                if i > 1:
                    v += (i - 1) * h * S[i - 2, j]
                if j > 0:
                    v += j * h * S[i - 1, j - 1]
            else:
                v = PB * S[i, j - 1]
                if j > 1:
                    v += (j - 1) * h * S[i, j - 2]
            S[i, j] = v
    return S


def _os_kinetic_1d(la, lb, PA, PB, a, b, p, S):
    T = np.zeros((la + 1, lb + 1))
    h = 1.0 / (2 * p)
    T[0, 0] = (a - 2 * a * a * (PA * PA + h)) * S[0, 0]
    for i in range(la + 1):
        for j in range(lb + 1):
            if i == 0 and j == 0:
                continue
            if i > 0:  # raise i from (i-1, j)
                i0 = i - 1
                v = PA * T[i0, j]
                if i0 > 0:
                    v += i0 * h * T[i0 - 1, j]
                if j > 0:
                    v += j * h * T[i0, j - 1]
                v += b / p * (2 * a * S[i, j] - (i0 * S[i0 - 1, j] if i0 > 0 else 0.0))
            else:  # raise j from (0, j-1)
                j0 = j - 1
                v = PB * T[0, j0]
                if j0 > 0:
                    v += j0 * h * T[0, j0 - 1]
                v += a / p * (2 * b * S[0, j] - (j0 * S[0, j0 - 1] if j0 > 0 else 0.0))
            T[i, j] = v
    return T


def _prim_one(ca, cb, A, B, a, b, atoms):
    """(S, T, V) of one primitive pair of Cartesian components ca, cb (unnormalized)."""
    p = a + b
    mu = a * b / p
    P = (a * A + b * B) / p
    PA, PB = P - A, P - B
    S1, T1 = [], []
    for x in range(3):
        s00 = np.sqrt(np.pi / p) * np.exp(-mu * (A[x] - B[x]) ** 2)
        S = _os_overlap_1d(ca[x], cb[x], PA[x], PB[x], p, s00)
        T = _os_kinetic_1d(ca[x], cb[x], PA[x], PB[x], a, b, p, S)
        S1.append(S[ca[x], cb[x]])
        T1.append(T[ca[x], cb[x]])
    s = S1[0] * S1[1] * S1[2]
    t = T1[0] * S1[1] * S1[2] + S1[0] * T1[1] * S1[2] + S1[0] * S1[1] * T1[2]
    Kab = np.exp(-mu * float((A - B) @ (A - B)))
    v = 0.0
    for Zc, Cc in atoms:
        PC = P - Cc
        Tb = p * float(PC @ PC)
        base = 2 * np.pi / p * Kab

        @lru_cache(maxsize=None)
        def th(av, bv, m):
            if min(av) < 0 or min(bv) < 0:
                return 0.0
            for i in range(3):
                if av[i] > 0:
                    a1 = tuple(x - (k == i) for k, x in enumerate(av))
                    a2 = tuple(x - (k == i) for k, x in enumerate(a1))
                    b1 = tuple(x - (k == i) for k, x in enumerate(bv))
                    r = PA[i] * th(a1, bv, m) - PC[i] * th(a1, bv, m + 1)
                    if a1[i] > 0:
                        r += a1[i] / (2 * p) * (th(a2, bv, m) - th(a2, bv, m + 1))
                    if bv[i] > 0:
                        r += bv[i] / (2 * p) * (th(a1, b1, m) - th(a1, b1, m + 1))
                    return r
            for i in range(3):
                if bv[i] > 0:
                    b1 = tuple(x - (k == i) for k, x in enumerate(bv))
                    b2 = tuple(x - (k == i) for k, x in enumerate(b1))
                    r = PB[i] * th(av, b1, m) - PC[i] * th(av, b1, m + 1)
                    if b1[i] > 0:
                        r += b1[i] / (2 * p) * (th(av, b2, m) - th(av, b2, m + 1))
                    return r
            return base * boys(m, Tb)

        v -= Zc * th(tuple(ca), tuple(cb), 0)
    return s, t, v


def one_electron_block(basis, sa, sb, Z, xyz):
    """(S, T, V) blocks (naoA x naoB) of normalized contracted functions of shells sa, sb."""
    from hfx.basis import cart_components
    A_, B_ = basis.shells[sa], basis.shells[sb]
    A, B = np.array(A_.center, float), np.array(B_.center, float)
    atoms = [(float(z), np.array(r, float)) for z, r in zip(Z, xyz)]
    ca_l, cb_l = cart_components(A_.l), cart_components(B_.l)
    out = np.zeros((3, len(ca_l), len(cb_l)))
    for pa in range(A_.nprim):
        for pb in range(B_.nprim):
            cc = A_.coefficients[pa] * B_.coefficients[pb]
            for ia, ca in enumerate(ca_l):
                for ib, cb in enumerate(cb_l):
                    n = A_.norm[pa * A_.nao + ia] * B_.norm[pb * B_.nao + ib]
                    s, t, v = _prim_one(ca, cb, A, B, A_.exponents[pa], B_.exponents[pb], atoms)
                    out[:, ia, ib] += n * cc * np.array([s, t, v])
    return out
