"""CPU tests pinning the oracle (a restatement of the reference's integral engine and direct
J/K loop) against mathematics that does not depend on it: mpmath Boys values, closed forms
of E and R, an independent Obara-Saika ERI evaluation, textbook values and brute force."""
import ctypes as C
import math
import os

import mpmath
import numpy as np
import pytest

import molecules
import os_reference
from hfx import lib

_f64p = C.POINTER(C.c_double)


def boys_exact(m, T):
    return os_reference.boys(m, T)


# ---------------------------------------------------------------------------------------
# Boys function (reference src/Boys.cpp:14-71)
# ---------------------------------------------------------------------------------------
@pytest.mark.parametrize("m_max", [0, 1, 2, 4, 8, 10])
def test_boys_matches_mpmath(oracle, m_max):
    Ts = [0.0, 1e-9, 1e-5, 3e-5, 0.1, 0.5, 1.0, 2.34, 2.36, 3.0, 4.2, 7.77, 12.5, 19.99, 25.0, 29.99, 30.0]
    F = (C.c_double * (m_max + 1))()
    for T in Ts:
        oracle.dll.oracle_boys(m_max, T, F)
        for m in range(m_max + 1):
            ref = boys_exact(m, T)
            assert abs(F[m] - ref) <= 2e-13 * ref, (m_max, m, T, F[m], ref)


def test_boys_large_T_is_the_reference_asymptotic_form(oracle):
    # T > 30: F_0 = sqrt(pi)/2/sqrt(T), upward (m+1/2)F_m/T without exp(-T) (Boys.cpp:16-23).
    F = (C.c_double * 9)()
    for T in (30.0001, 45.0, 300.0):
        oracle.dll.oracle_boys(8, T, F)
        f = 0.8862269254527579 / math.sqrt(T)
        for m in range(9):
            assert F[m] == pytest.approx(f, rel=1e-15)
            f = (m + 0.5) * f / T
        # the low orders are accurate there; the dropped exp(-T) shows in the high ones
        assert F[0] == pytest.approx(boys_exact(0, T), rel=1e-13)


def test_boys_series_against_mpmath(oracle):
    for m in (0, 3, 11):
        for T in (0.0, 0.7, 13.0, 29.0):
            assert oracle.dll.oracle_boys_series(m, T) == pytest.approx(boys_exact(m, T), rel=1e-13)


# ---------------------------------------------------------------------------------------
# Hermite E and R (IntegralEngine.cpp:1025-1162) against closed forms
# ---------------------------------------------------------------------------------------
def _E(oracle, l1, l2, p, PA, PB):
    n = (l1 + 1) * (l2 + 1) * (l1 + l2 + 2) // 2
    buf = (C.c_double * n)()
    oracle.dll.oracle_hermite(l1, l2, p, PA, PB, buf)
    return lambda i, j, t: buf[oracle.dll.oracle_hermite_index(l1, l2, i, j, t)]


def test_hermite_closed_forms(oracle):
    p, PA, PB = 1.7, 0.31, -0.42
    E = _E(oracle, 2, 2, p, PA, PB)
    h = 1 / (2 * p)
    assert E(0, 0, 0) == 1.0
    assert E(1, 0, 0) == pytest.approx(PA, rel=1e-15)
    assert E(1, 0, 1) == pytest.approx(h, rel=1e-15)
    assert E(0, 1, 0) == pytest.approx(PB, rel=1e-15)
    assert E(1, 1, 0) == pytest.approx(PA * PB + h, rel=1e-14)
    assert E(1, 1, 1) == pytest.approx((PA + PB) * h, rel=1e-14)
    assert E(1, 1, 2) == pytest.approx(h * h, rel=1e-14)
    assert E(2, 0, 0) == pytest.approx(PA * PA + h, rel=1e-14)
    assert E(2, 0, 1) == pytest.approx(2 * PA * h, rel=1e-14)
    assert E(2, 0, 2) == pytest.approx(h * h, rel=1e-14)
    # E^{22}_0 = <x_A^2 x_B^2> moment: PA^2 PB^2 + h(PA^2 + PB^2 + 4 PA PB) + 3h^2
    assert E(2, 2, 0) == pytest.approx(PA**2 * PB**2 + h * (PA**2 + PB**2 + 4 * PA * PB) + 3 * h * h, rel=1e-13)


def test_hermite_r_closed_forms(oracle):
    L, p = 3, 0.83
    PC = np.array([0.4, -0.7, 1.1])
    out = (C.c_double * ((L + 1) ** 3))()
    oracle.dll.oracle_hermite_r(L, p, PC.ctypes.data_as(_f64p), out)
    R = lambda t, u, v: out[t + u * (L + 1) + v * (L + 1) ** 2]
    T = p * float(PC @ PC)
    F = [boys_exact(m, T) for m in range(L + 1)]
    X, Y, Z = PC
    m2p = -2 * p
    assert R(0, 0, 0) == pytest.approx(F[0], rel=1e-12)
    assert R(1, 0, 0) == pytest.approx(X * m2p * F[1], rel=1e-12)
    assert R(0, 0, 1) == pytest.approx(Z * m2p * F[1], rel=1e-12)
    assert R(2, 0, 0) == pytest.approx(m2p * F[1] + X * X * m2p**2 * F[2], rel=1e-12)
    assert R(1, 1, 0) == pytest.approx(X * Y * m2p**2 * F[2], rel=1e-12)
    assert R(1, 1, 1) == pytest.approx(X * Y * Z * m2p**3 * F[3], rel=1e-12)
    assert R(3, 0, 0) == pytest.approx(3 * X * m2p**2 * F[2] + X**3 * m2p**3 * F[3], rel=1e-12)


# ---------------------------------------------------------------------------------------
# One-electron integrals and ERIs
# ---------------------------------------------------------------------------------------
def test_szabo_ostlund_h2(oracle):
    """Szabo & Ostlund, Modern Quantum Chemistry, H2/STO-3G at R = 1.4 a0 (Sec. 3.5.2)."""
    b = molecules.h2()
    e = lib.Engine(oracle, b)
    S, T = e.overlap(), e.kinetic()
    V = e.nuclear(b.Z, b.xyz)
    assert S[0, 1] == pytest.approx(0.6593, abs=1e-4)
    assert T[0, 0] == pytest.approx(0.7600, abs=1e-4)
    assert T[0, 1] == pytest.approx(0.2365, abs=1e-4)
    assert (T + V)[0, 0] == pytest.approx(-1.1204, abs=1e-4)
    assert (T + V)[0, 1] == pytest.approx(-0.9584, abs=1e-4)
    g = e.eri_quartets([(0, 0, 0, 0), (1, 1, 0, 0), (1, 0, 0, 0), (1, 0, 1, 0)])
    assert [x.ravel()[0] for x in g] == pytest.approx([0.7746, 0.5697, 0.4441, 0.2970], abs=1e-4)
    r = lib.scf_run(oracle, b, options=oracle.default_options(energy_tol=1e-10))
    assert r["e_total"] == pytest.approx(-1.1167, abs=1e-4)


def test_crawford_water_energy(oracle):
    """Published SCF energy of the Crawford group's H2O/STO-3G project: -74.942079928192 Eh."""
    b = molecules.crawford_h2o()
    r = lib.scf_run(oracle, b, options=oracle.default_options(energy_tol=1e-10, density_tol=1e-8,
                                                               diis_error_tol=1e-8))
    assert r["converged"] == 1
    assert r["e_nuclear"] == pytest.approx(8.002367061810450, abs=1e-12)
    assert r["e_total"] == pytest.approx(-74.942079928192, abs=1e-6)


def test_overlap_normalization(oracle):
    """Primitive normalization (Basis.cpp:61-77): single-primitive shells have S_ii = 1."""
    b = molecules.methane()
    S = lib.Engine(oracle, b).overlap()
    off = b.flat()["ao_offset"]
    for s, sh in enumerate(b.shells):
        if sh.nprim == 1:
            for c in range(sh.nao):
                i = off[s] + c
                from hfx.basis import cart_components
                (x, y, z) = cart_components(sh.l)[c]
                # <x^2|x^2> for d_xx has norm 3/3 = 1 with the (2i-1)!! factors
                assert S[i, i] == pytest.approx(1.0, rel=1e-13)
    assert np.allclose(S, S.T, atol=1e-15)


@pytest.mark.parametrize("mol", ["methane_ccpvdz", "h2o_631gss"])
def test_one_electron_against_obara_saika(oracle, mol):
    """S, T and V of every shell pair, (la, lb) through (d, d), against the independent Obara-Saika
    recurrences (reference IntegralEngine.cpp:78-142 overlap, 162-238 kinetic with its l(l-1) term
    for l >= 2 at 226-232, 259-352 nuclear attraction)."""
    b = molecules.methane() if mol == "methane_ccpvdz" else molecules.h2o("6-31g**")
    e = lib.Engine(oracle, b)
    S, T, V = e.overlap(), e.kinetic(), e.nuclear(b.Z, b.xyz)
    off = b.flat()["ao_offset"]
    seen = set()
    for sa, A in enumerate(b.shells):
        for sb, B in enumerate(b.shells[: sa + 1]):
            ref = os_reference.one_electron_block(b, sa, sb, b.Z, b.xyz)
            blk = [M[off[sa]:off[sa] + A.nao, off[sb]:off[sb] + B.nao] for M in (S, T, V)]
            for name, got, want in zip("STV", blk, ref):
                scale = max(1.0, np.abs(want).max())
                assert np.abs(got - want).max() <= 1e-12 * scale, (name, sa, sb, A.l, B.l)
            seen.add((max(A.l, B.l), min(A.l, B.l)))
    assert seen == {(0, 0), (1, 0), (1, 1), (2, 0), (2, 1), (2, 2)}


def _one_electron_mismatch(L, b):
    """Largest scaled |oracle - Obara-Saika| over the shell pairs of b, per (la, lb) class."""
    e = lib.Engine(L, b)
    S, T, V = e.overlap(), e.kinetic(), e.nuclear(b.Z, b.xyz)
    off = b.flat()["ao_offset"]
    worst = {}
    for sa, A in enumerate(b.shells):
        for sb, B in enumerate(b.shells[: sa + 1]):
            ref = os_reference.one_electron_block(b, sa, sb, b.Z, b.xyz)
            blk = [M[off[sa]:off[sa] + A.nao, off[sb]:off[sb] + B.nao] for M in (S, T, V)]
            for name, got, want in zip("STV", blk, ref):
                k = (name, max(A.l, B.l), min(A.l, B.l))
                worst[k] = max(worst.get(k, 0.0), np.abs(got - want).max() / max(1.0, np.abs(want).max()))
    return worst


@pytest.mark.parametrize("plant", ["ORACLE_PLANT_KINETIC_FACTOR", "ORACLE_PLANT_NUCLEAR_TRANSPOSE"])
def test_obara_saika_check_catches_planted_oracle_mistakes(tmp_path, plant):
    """Mutation test of the check above: the oracle built with a planted slip (the kinetic
    l(l-1) term without its factor 1/2, IntegralEngine.cpp:226-232; or the u and v indices of
    R^0_tuv transposed in the nuclear attraction, :325-340) fails the Obara-Saika comparison on
    exactly the classes the slip reaches, and still passes the others."""
    import subprocess
    here = os.path.dirname(os.path.abspath(__file__))
    odir = os.path.join(os.path.dirname(here), "oracle")
    so = str(tmp_path / "liboracle_planted.so")
    subprocess.check_call(["g++", "-std=c++20", "-fPIC", "-shared", "-fopenmp", "-O1", "-ffp-contract=off",
                           "-I" + os.path.join(os.path.dirname(here), "include"), "-D" + plant,
                           os.path.join(odir, "hfx_oracle.cpp"),
                           os.path.join(os.path.dirname(here), "hartree-fock_amd", "csrc", "hfx_scf.cpp"),
                           "-o", so])
    worst = _one_electron_mismatch(lib.Library(so), molecules.h2o("6-31g**"))
    if plant == "ORACLE_PLANT_KINETIC_FACTOR":
        hit = {k for k, v in worst.items() if v > 1e-9}
        assert hit and all(k[0] == "T" and k[1] == 2 for k in hit), hit   # only T of pairs with a d shell
    else:
        hit = {k for k, v in worst.items() if v > 1e-9}
        assert hit and all(k[0] == "V" and k[1] >= 1 for k in hit), hit   # only V, never ss
    assert all(v <= 1e-12 for k, v in worst.items() if k not in hit)


def test_ssss_closed_form(oracle):
    b = molecules.water_dimer("6-31g")
    e = lib.Engine(oracle, b)
    sh = b.shells
    s_idx = [i for i, x in enumerate(sh) if x.l == 0]
    a, bb, c, d = s_idx[3], s_idx[0], s_idx[6], s_idx[5]  # a >= b, c >= d: no reference zeros
    got = e.eri_quartets([(a, bb, c, d)])[0].ravel()[0]
    ref = 0.0
    A, B, Cc, D = (np.array(sh[i].center) for i in (a, bb, c, d))
    for pa in range(sh[a].nprim):
        for pb in range(sh[bb].nprim):
            for pc in range(sh[c].nprim):
                for pd in range(sh[d].nprim):
                    al, be, ga, de = (sh[i].exponents[k] for i, k in ((a, pa), (bb, pb), (c, pc), (d, pd)))
                    p, q = al + be, ga + de
                    P, Q = (al * A + be * B) / p, (ga * Cc + de * D) / q
                    K = math.exp(-al * be / p * float((A - B) @ (A - B)) - ga * de / q * float((Cc - D) @ (Cc - D)))
                    T = p * q / (p + q) * float((P - Q) @ (P - Q))
                    n = sh[a].norm[pa] * sh[bb].norm[pb] * sh[c].norm[pc] * sh[d].norm[pd]
                    cf = sh[a].coefficients[pa] * sh[bb].coefficients[pb] * sh[c].coefficients[pc] * sh[d].coefficients[pd]
                    ref += n * cf * 2 * math.pi**2.5 / (p * q * math.sqrt(p + q)) * K * boys_exact(0, T)
    assert got == pytest.approx(ref, rel=1e-13)


def test_eri_against_obara_saika(oracle):
    """Every angular-momentum class through dddd, against the independent Obara-Saika code."""
    b = molecules.ethane()
    e = lib.Engine(oracle, b)
    sh = b.shells
    by_l = {l: [i for i, s in enumerate(sh) if s.l == l and s.nprim <= 4] for l in (0, 1, 2)}
    rng = np.random.default_rng(3)
    quartets = []
    for la in (0, 1, 2):
        for lb in (0, 1, 2):
            for lc in (0, 1, 2):
                for ld in (0, 1, 2):
                    if la + lb + lc + ld > 6 and rng.random() < 0.6:
                        continue
                    q = [int(rng.choice(by_l[l])) for l in (la, lb, lc, ld)]
                    a_, b_ = max(q[0], q[1]), min(q[0], q[1])
                    c_, d_ = max(q[2], q[3]), min(q[2], q[3])
                    quartets.append((a_, b_, c_, d_))
    blocks = e.eri_quartets(quartets)
    off = b.flat()["ao_offset"]
    for (a, bb, c, d), blk in zip(quartets, blocks):
        ref = os_reference.contracted_eri(b, a, bb, c, d)
        for A in range(blk.shape[0]):
            for B in range(blk.shape[1]):
                for Cc in range(blk.shape[2]):
                    for D in range(blk.shape[3]):
                        i, j, k, l = off[a] + A, off[bb] + B, off[c] + Cc, off[d] + D
                        zero = j > i or l > k or (a == c and bb == d and i * (i + 1) // 2 + j < k * (k + 1) // 2 + l)
                        if zero:
                            assert blk[A, B, Cc, D] == 0.0
                        else:
                            assert blk[A, B, Cc, D] == pytest.approx(ref[A, B, Cc, D], rel=1e-10, abs=1e-13), \
                                ((a, bb, c, d), (A, B, Cc, D))


def test_schwarz_is_sqrt_of_diagonal_and_bounds(oracle):
    b = molecules.water_dimer("6-31g**")
    e = lib.Engine(oracle, b)
    Q = e.schwarz()
    G = os_reference.dense_eri(e, b)
    off = b.flat()["ao_offset"]
    nao = [s.nao for s in b.shells]
    for a in range(b.n_shells):
        for c in range(a + 1):
            ia = slice(off[a], off[a] + nao[a])
            ic = slice(off[c], off[c] + nao[c])
            diag = np.einsum("ijij->ij", G[ia, ic, ia, ic])
            assert Q[a, c] == pytest.approx(math.sqrt(diag.max()), rel=1e-14)
            assert Q[a, c] == Q[c, a]
    # Schwarz inequality over a sample of shell quartets
    rng = np.random.default_rng(5)
    for _ in range(200):
        a, bb, c, d = rng.integers(0, b.n_shells, 4)
        blk = G[off[a]:off[a] + nao[a], off[bb]:off[bb] + nao[bb], off[c]:off[c] + nao[c], off[d]:off[d] + nao[d]]
        assert np.abs(blk).max() <= Q[a, bb] * Q[c, d] * (1 + 1e-12)


def test_eri_permutational_symmetry(oracle):
    b = molecules.h2o("6-31g**")
    e = lib.Engine(oracle, b)
    G = os_reference.dense_eri(e, b)
    # blocks computed in another shell order must agree with the symmetric tensor
    off = b.flat()["ao_offset"]
    nao = [s.nao for s in b.shells]
    for (a, bb, c, d) in [(0, 3, 5, 1), (5, 5, 2, 4), (3, 1, 3, 1), (6, 2, 4, 6)]:
        blk = e.eri_quartets([(a, bb, c, d)])[0]
        if not (a >= bb and c >= d and a * (a + 1) // 2 + bb >= c * (c + 1) // 2 + d):
            continue
        ref = G[off[a]:off[a] + nao[a], off[bb]:off[bb] + nao[bb], off[c]:off[c] + nao[c], off[d]:off[d] + nao[d]]
        mask = blk != 0
        assert np.allclose(blk[mask], ref[mask], rtol=1e-13, atol=1e-15)


# ---------------------------------------------------------------------------------------
# Direct J/K (SCF.cpp:523-744) against brute force over the dense tensor
# ---------------------------------------------------------------------------------------
def _random_density(n, rng, nocc):
    C_ = rng.standard_normal((n, nocc))
    return C_ @ C_.T * 0.1


@pytest.mark.parametrize("mol", ["h2o_631gss", "oh_uhf"])
def test_jk_against_bruteforce(oracle, mol):
    rng = np.random.default_rng(11)
    b = molecules.h2o("6-31g**") if mol == "h2o_631gss" else molecules.oh_radical()
    e = lib.Engine(oracle, b)
    G = os_reference.dense_eri(e, b)
    n = b.n_aos
    Da = _random_density(n, rng, 5)
    Db = _random_density(n, rng, 4) if mol == "oh_uhf" else None
    Dt = Da + (Db if Db is not None else Da)
    J, Ka, Kb, st = e.jk(Dt, Da, Db, schwarz_thresh=0.0, density_thresh=0.0)
    Jr = np.einsum("ijkl,kl->ij", G, Dt)
    Kr = np.einsum("ijkl,jk->il", G, Da)
    assert np.abs(J - Jr).max() < 1e-12
    assert np.abs(Ka - Kr).max() < 1e-12
    if Db is not None:
        assert np.abs(Kb - np.einsum("ijkl,jk->il", G, Db)).max() < 1e-12
    nS = b.n_shells
    assert st["n_quartets"] == len([1 for a in range(nS) for bb in range(a + 1) for c in range(nS)
                                    for d in range(c + 1) if a * (a + 1) // 2 + bb >= c * (c + 1) // 2 + d])


def test_jk_linearity_and_rhf_identity(oracle):
    """J/K are linear in the density (the incremental build relies on it), and for one doubly
    occupied orbital c, c^T (J[2P] - K[P]) c = c^T J[P] c = (ii|ii) with P = c c^T."""
    rng = np.random.default_rng(2)
    b = molecules.h2o("6-31g")
    e = lib.Engine(oracle, b)
    n = b.n_aos
    D1, D2 = _random_density(n, rng, 3), _random_density(n, rng, 3)
    J1, K1, _, _ = e.jk(D1, D1, schwarz_thresh=0.0, density_thresh=0.0)
    J2, K2, _, _ = e.jk(D2, D2, schwarz_thresh=0.0, density_thresh=0.0)
    J12, K12, _, _ = e.jk(D1 + D2, D1 + D2, schwarz_thresh=0.0, density_thresh=0.0)
    assert np.abs(J12 - J1 - J2).max() < 1e-12
    assert np.abs(K12 - K1 - K2).max() < 1e-12
    c = rng.standard_normal(n)
    Dh = np.outer(c, c) * 0.05
    J, K, _, _ = e.jk(2 * Dh, Dh, schwarz_thresh=0.0, density_thresh=0.0)
    Jh, _, _, _ = e.jk(Dh, Dh, schwarz_thresh=0.0, density_thresh=0.0)
    assert c @ (J - K) @ c == pytest.approx(c @ Jh @ c, rel=1e-12)
    assert c @ K @ c == pytest.approx(c @ Jh @ c, rel=1e-12)


def test_screening_counts_match_bruteforce(oracle):
    """The survivor count equals a direct enumeration of the two shell tests."""
    rng = np.random.default_rng(4)
    b = molecules.water_dimer("6-31g")
    e = lib.Engine(oracle, b)
    Q = e.schwarz()
    n, nS = b.n_aos, b.n_shells
    D = _random_density(n, rng, 4) * 1e-6
    D[np.abs(D) < 3e-9] = 0.0
    tauS, tauD = 1e-7, 1e-9
    _, _, _, st = e.jk(2 * D, D, schwarz_thresh=tauS, density_thresh=tauD)
    off = b.flat()["ao_offset"]
    nao = [s.nao for s in b.shells]
    M = np.zeros((nS, nS))
    for x in range(nS):
        for y in range(nS):
            M[x, y] = np.abs(D[off[x]:off[x] + nao[x], off[y]:off[y] + nao[y]]).max()
    cnt = 0
    for a in range(nS):
        for bb in range(a + 1):
            for c in range(nS):
                for d in range(c + 1):
                    if a * (a + 1) // 2 + bb < c * (c + 1) // 2 + d:
                        continue
                    if Q[a, bb] * Q[c, d] < tauS:
                        continue
                    mx = max(2 * 2 * M[a, bb], 2 * 2 * M[c, d], M[a, c], M[a, d], M[bb, c], M[bb, d])
                    if mx < tauD:
                        continue
                    cnt += 1
    assert st["n_quartets"] == cnt
    assert 0 < cnt < nS**4


# ---------------------------------------------------------------------------------------
# SCF driver against an independent numpy SCF on the oracle's integrals
# ---------------------------------------------------------------------------------------
def test_device_linalg_needs_the_hip_library(oracle):
    """The host driver linked into the oracle refuses the device iteration with a clear error."""
    b = molecules.h2o("sto-3g")
    with pytest.raises(lib.HfxError, match="HIP library"):
        lib.scf_run(oracle, b, options=oracle.default_options(device_linalg=1))
    assert lib.scf_run(oracle, b)["converged"] == 1


def test_scf_against_numpy(oracle):
    import scipy.linalg as sl
    b = molecules.h2o("6-31g**")
    e = lib.Engine(oracle, b)
    S, T = e.overlap(), e.kinetic()
    V = e.nuclear(b.Z, b.xyz)
    h = T + V
    G = os_reference.dense_eri(e, b)
    nocc = b.n_electrons() // 2
    D = np.zeros_like(h)
    E_old = 0
    for _ in range(100):
        F = h + 2 * np.einsum("ijkl,kl->ij", G, D) - np.einsum("ijkl,jk->il", G, D)
        w, Cm = sl.eigh(F, S)
        D = Cm[:, :nocc] @ Cm[:, :nocc].T
        E = np.sum(D * (h + F))
        if abs(E - E_old) < 1e-12:
            break
        E_old = E
    F = h + 2 * np.einsum("ijkl,kl->ij", G, D) - np.einsum("ijkl,jk->il", G, D)
    E = np.sum(D * (h + F)) + b.nuclear_repulsion()
    # Incremental builds with density screening at 1e-10 reach a noise floor near 1e-9 Eh
    # (the reference's scheme, SCF.cpp:536-541/731); a tight pin uses tiny thresholds.
    r = lib.scf_run(oracle, b, options=oracle.default_options(energy_tol=1e-12, density_tol=1e-10,
                                                               diis_error_tol=1e-10, use_diis=0, max_iter=200,
                                                               schwarz_thresh=1e-14, density_thresh=1e-14))
    assert r["converged"] == 1
    assert r["e_total"] == pytest.approx(E, abs=1e-10)
    # orbital energies: the driver's eigensolver against scipy's generalized eigh
    assert np.abs(np.sort(r["eps_alpha"]) - w).max() < 1e-7


def test_uhf_driver_runs(oracle):
    b = molecules.oh_radical("6-31g")
    r = lib.scf_run(oracle, b, charge=0, multiplicity=2,
                    options=oracle.default_options(energy_tol=1e-9, max_iter=60))
    assert r["converged"] == 1
    assert -75.5 < r["e_total"] < -75.2


def test_model_flop_table(oracle):
    """Per-primitive-quartet model FLOPs of the factored McMurchie-Davidson count."""
    f = oracle.dll.oracle_model_flop
    assert f(0, 0, 0, 0) == 32
    assert f(1, 0, 0, 0) == 67
    assert f(1, 1, 0, 0) == 178
    assert f(1, 1, 1, 1) == 1661
    assert f(2, 1, 2, 1) == 9527
    assert f(2, 2, 2, 2) == 51836


# ---------------------------------------------------------------------------------------
# RHF nuclear gradient of the oracle: pinned by central finite differences of its own SCF
# energies and by the invariances of the exact gradient.
# ---------------------------------------------------------------------------------------
def _distorted_water(name):
    from hfx import Basis
    b = molecules.h2o(name)
    xyz = b.xyz.copy()
    xyz[0] += (0.12, -0.07, 0.05)
    xyz[2] += (-0.04, 0.09, -0.03)
    return Basis(name, b.Z, xyz)


def test_gradient_matches_finite_differences(oracle):
    import gradient_tools as gt
    b = _distorted_water("6-31g**")
    r, g, parts, _ = gt.analytic(oracle, b)
    fd = gt.finite_differences(oracle, b, coords=[(0, 0), (0, 2), (1, 1), (2, 0), (2, 2)])
    err = max(abs(fd[k] - g[k]) for k in fd)
    assert err <= 1e-6, err
    assert np.abs(g).max() > 1e-2  # a real test: the distorted geometry has sizable forces


def test_gradient_invariances(oracle):
    import gradient_tools as gt
    b = molecules.water_dimer("6-31g**")
    r, g, parts, _ = gt.analytic(oracle, b)
    # translation: every part sums to zero over atoms on its own
    for k in range(4):
        assert np.abs(parts[k].sum(0)).max() <= 1e-9, k
    # rotation: the total torque of the exact gradient vanishes
    torque = np.cross(b.xyz, g).sum(0)
    assert np.abs(torque).max() <= 1e-7
    # nuclear repulsion part against the closed form
    Z = np.asarray(b.Z, float)
    d = b.xyz[:, None, :] - b.xyz[None, :, :]
    r3 = np.linalg.norm(d, axis=-1) ** 3
    np.fill_diagonal(r3, np.inf)
    gn = -(Z[:, None, None] * Z[None, :, None] * d / r3[:, :, None]).sum(1)
    assert np.abs(parts[0] - gn).max() <= 1e-12


@pytest.mark.parametrize("mol", ["h2o_631gss", "oh_uhf"])
def test_stored_jk_against_bruteforce(oracle, mol):
    """The conventional path: the packed canonical tensor of hfx_store_eris digested by
    hfx_jk(HFX_JK_STORED) equals einsum over the dense tensor of the quartet blocks, and the
    direct build with the density test off."""
    rng = np.random.default_rng(12)
    b = molecules.h2o("6-31g**") if mol == "h2o_631gss" else molecules.oh_radical()
    e = lib.Engine(oracle, b)
    G = os_reference.dense_eri(e, b)
    n = b.n_aos
    Da = _random_density(n, rng, 5)
    Db = _random_density(n, rng, 4) if mol == "oh_uhf" else None
    Dt = Da + (Db if Db is not None else Da)
    nq, nbytes = e.store_eris(0.0)
    T = n * (n + 1) // 2
    assert nbytes == 8 * T * (T + 1) // 2
    J, Ka, Kb, st = e.jk(Dt, Da, Db, schwarz_thresh=0.0, flags=lib.HFX_JK_STORED)
    assert st["n_quartets"] == nq
    assert np.abs(J - np.einsum("ijkl,kl->ij", G, Dt)).max() < 1e-12
    assert np.abs(Ka - np.einsum("ijkl,jk->il", G, Da)).max() < 1e-12
    if Db is not None:
        assert np.abs(Kb - np.einsum("ijkl,jk->il", G, Db)).max() < 1e-12
    # with the thresholds on: the direct build without density test gives the same matrices
    e.store_eris(1e-10)
    J, Ka, Kb, _ = e.jk(Dt, Da, Db, schwarz_thresh=1e-10, flags=lib.HFX_JK_STORED)
    Jd, Kad, Kbd, _ = e.jk(Dt, Da, Db, schwarz_thresh=1e-10, density_thresh=0.0)
    assert np.abs(J - Jd).max() < 1e-12 and np.abs(Ka - Kad).max() < 1e-12
    if Db is not None:
        assert np.abs(Kb - Kbd).max() < 1e-12


def test_stored_count_and_errors(oracle):
    """Stored quartets = canonical quartets passing Q(a,b)Q(c,d) >= tau; the stored mode
    refuses to run without a store or in mixed precision."""
    b = molecules.water_dimer("6-31g")
    e = lib.Engine(oracle, b)
    Q = e.schwarz()
    nS = b.n_shells
    tau = 1e-6
    want = sum(1 for a in range(nS) for bb in range(a + 1) for c in range(nS) for d in range(c + 1)
               if a * (a + 1) // 2 + bb >= c * (c + 1) // 2 + d and Q[a, bb] * Q[c, d] >= tau)
    D = np.eye(b.n_aos) * 0.1
    with pytest.raises(lib.HfxError):
        e.jk(D, D, flags=lib.HFX_JK_STORED)
    nq, _ = e.store_eris(tau)
    assert nq == want
    with pytest.raises(lib.HfxError):
        e.jk(D, D, flags=lib.HFX_JK_STORED, precision=lib.HFX_MIXED, fp32_cut=1e-2)
    e.release_eris()
    with pytest.raises(lib.HfxError):
        e.jk(D, D, flags=lib.HFX_JK_STORED)


def test_conventional_scf_matches_direct(oracle):
    """DIRECT false (stored integrals, every build from the full density) converges to the energy
    of the dense-tensor SCF, as the direct driver does."""
    b = molecules.h2o("6-31g**")
    opts = dict(energy_tol=1e-12, density_tol=1e-10, diis_error_tol=1e-10, max_iter=100,
                schwarz_thresh=1e-14, density_thresh=1e-14)
    rc = lib.scf_run(oracle, b, options=oracle.default_options(direct=0, **opts))
    rd = lib.scf_run(oracle, b, options=oracle.default_options(direct=1, full_builds=1, **opts))
    assert rc["converged"] == 1 and rd["converged"] == 1
    assert rc["e_total"] == pytest.approx(rd["e_total"], abs=1e-10)


def test_basis_errors_are_the_references():
    """BasisSet failures (src/Basis.cpp:103-109, 161-166): an unknown basis name and an element
    the basis does not cover are errors, not empty bases."""
    from hfx import Basis
    with pytest.raises(ValueError, match="not found"):
        Basis("no-such-basis", [1], [[0.0, 0.0, 0.0]])
    with pytest.raises(ValueError, match="not found"):
        Basis("sto-3g", [92], [[0.0, 0.0, 0.0]])


@pytest.mark.parametrize("batched", [False, True])
def test_mixed_precision_switch_control_flow(oracle, batched):
    """The driver's precision stages (the oracle builds everything in FP64, so only the control
    flow differs): with switch_tol the first stage ends once |dE| < switch_tol and the FP64 stage
    keeps the DIIS history, reaching the FP64 energy in as many iterations as the FP64 SCF (+1);
    with switch_tol = 0 the first stage converges and the FP64 stage restarts."""
    mols = [molecules.water_dimer("6-31g"), molecules.h2o("6-31g**")]
    kw = dict(energy_tol=1e-8, schwarz_thresh=1e-10, density_thresh=1e-10, max_iter=60)
    run = (lambda o: lib.scf_run_batched(oracle, mols, options=o)) if batched else \
        (lambda o: [lib.scf_run(oracle, mols[0], options=o)])
    r64 = run(oracle.default_options(**kw))
    for sw in (1e-4, 0.0):
        rm = run(oracle.default_options(precision=lib.HFX_MIXED, fp32_cut=1e-2, final_fp64=1, switch_tol=sw, **kw))
        for a, b in zip(rm, r64):
            assert a["converged"] == 1
            assert a["e_total"] == pytest.approx(b["e_total"], abs=1e-7)  # both converged to dE < 1e-8
            if sw > 0:
                assert a["iterations_mixed"] < b["iterations"] - 2
                assert a["iterations"] <= b["iterations"] + 1
            else:
                assert a["iterations_mixed"] == b["iterations"]


def test_general_contraction_expansion():
    """Reading cc-pVDZ with every contraction row as a shell (zero-coefficient primitives dropped)
    gives exactly the shells of the committed cc-pvdz-expanded data, while the reference's reading
    (one row per angular momentum, src/Basis.cpp:125-159) keeps only the first row."""
    from hfx import Basis
    Z = [1, 6, 7, 8]
    xyz = np.arange(12, dtype=float).reshape(4, 3)
    a = Basis("cc-pvdz", Z, xyz, general="expand")
    b = Basis("cc-pvdz-expanded", Z, xyz)
    assert [(s.l, s.exponents, s.coefficients, s.atom) for s in a.shells] == \
        [(s.l, s.exponents, s.coefficients, s.atom) for s in b.shells]
    r = Basis("cc-pvdz", Z, xyz)
    assert r.n_aos < a.n_aos
    assert Basis("sto-3g", Z, xyz, general="expand").n_aos == Basis("sto-3g", Z, xyz).n_aos


def test_conventional_batched_scf(oracle):
    """The batched driver's conventional mode (stored tensor per molecule) reaches the same
    per-molecule energies as its direct full-build mode and as single-molecule runs."""
    mols = [molecules.h2o("6-31g"), molecules.water_dimer("6-31g")]
    kw = dict(energy_tol=1e-8, max_iter=60, schwarz_thresh=1e-12, density_thresh=1e-12)
    rc = lib.scf_run_batched(oracle, mols, options=oracle.default_options(direct=0, **kw))
    rd = lib.scf_run_batched(oracle, mols, options=oracle.default_options(direct=1, full_builds=1, **kw))
    for m, a, b in zip(mols, rc, rd):
        rs = lib.scf_run(oracle, m, options=oracle.default_options(direct=0, **kw))
        assert a["converged"] == 1 and b["converged"] == 1
        assert a["e_total"] == pytest.approx(b["e_total"], abs=1e-9)
        assert a["e_total"] == pytest.approx(rs["e_total"], abs=1e-9)
