"""The spherical-harmonic orbital space (hfx_scf_options.orbital_space with
Basis.spherical_transform): the transform against the overlap it must diagonalize, the SCF
against an independent numpy SCF in the transformed basis, and the HIP engine against the
oracle.  The reference has no spherical option (its basis is Cartesian, src/Basis.cpp:47-77);
this is the format-layer extension of the basis reader."""
import numpy as np
import pytest

import molecules
import os_reference
from hfx import lib

TIGHT = dict(energy_tol=1e-12, density_tol=1e-10, diis_error_tol=1e-10, use_diis=0, max_iter=200,
             schwarz_thresh=1e-14, density_thresh=1e-14)


def test_transform_diagonalizes_the_d_block(oracle):
    """On every d shell: T^T S T = s_xy 1 (the five functions orthogonal, with the norm of the
    shell's xy component), and the s-type contaminant xx + yy + zz is S-orthogonal to them.
    A wrong sqrt(3)/2, a wrong sign or a swapped component fails one of these."""
    b = molecules.methane()
    S = lib.Engine(oracle, b).overlap()
    T = b.spherical_transform()
    assert T.shape == (b.n_aos, b.n_aos - sum(1 for sh in b.shells if sh.l == 2))
    off = b.flat()["ao_offset"]
    col = 0
    n_d = 0
    for s, sh in enumerate(b.shells):
        if sh.l == 2:
            o = off[s]
            Sd = S[o:o + 6, o:o + 6]
            Td = T[o:o + 6, col:col + 5]
            M = Td.T @ Sd @ Td
            sxy = Sd[1, 1]
            assert np.abs(M - sxy * np.eye(5)).max() <= 1e-12 * sxy
            u = np.array([1.0, 0, 0, 1.0, 0, 1.0])  # xx + yy + zz
            assert np.abs(u @ Sd @ Td).max() <= 1e-12 * sxy
            n_d += 1
        col += 5 if sh.l == 2 else sh.nao
    assert n_d == 1 and col == T.shape[1]


def _numpy_spherical_scf(engine, b):
    import scipy.linalg as sl
    T = b.spherical_transform()
    S, h = engine.overlap(), engine.kinetic() + engine.nuclear(b.Z, b.xyz)
    G = os_reference.dense_eri(engine, b)
    Gs = np.einsum("pi,qj,rk,sl,pqrs->ijkl", T, T, T, T, G, optimize=True)
    Ss, hs = T.T @ S @ T, T.T @ h @ T
    nocc = b.n_electrons() // 2
    D = np.zeros_like(hs)
    E_old = 0.0
    for _ in range(200):
        F = hs + 2 * np.einsum("ijkl,kl->ij", Gs, D) - np.einsum("ijkl,jk->il", Gs, D)
        w, C = sl.eigh(F, Ss)
        D = C[:, :nocc] @ C[:, :nocc].T
        E = np.sum(D * (hs + F))
        if abs(E - E_old) < 1e-13:
            break
        E_old = E
    F = hs + 2 * np.einsum("ijkl,kl->ij", Gs, D) - np.einsum("ijkl,jk->il", Gs, D)
    return np.sum(D * (hs + F)) + b.nuclear_repulsion(), w


def test_spherical_scf_against_numpy(oracle):
    b = molecules.h2o("6-31g**")
    e = lib.Engine(oracle, b)
    E_np, w = _numpy_spherical_scf(e, b)
    r = lib.scf_run(oracle, b, options=oracle.default_options(**TIGHT), spherical=True)
    rc = lib.scf_run(oracle, b, options=oracle.default_options(**TIGHT))
    assert r["converged"] == 1 and r["n_mo"] == b.n_aos - 1 and rc["n_mo"] == b.n_aos
    assert r["e_total"] == pytest.approx(E_np, abs=1e-10)
    # fewer functions: the spherical energy lies above the Cartesian one
    assert r["e_total"] > rc["e_total"] + 1e-6
    m = r["n_mo"]
    assert np.abs(r["eps_alpha"][:m] - w).max() < 1e-7
    # the removed direction: zero coefficients, the marker orbital energy
    assert np.all(r["eps_alpha"][m:] == 1e6) and np.abs(r["C_alpha"][:, m:]).max() == 0.0
    # the density lies in the spherical space: D = T D_s T^T for some D_s
    T = b.spherical_transform()
    P = T @ np.linalg.pinv(T)
    assert np.abs(P @ r["D_alpha"] @ P.T - r["D_alpha"]).max() < 1e-12


def test_spherical_with_diis_and_without_d(oracle):
    """With DIIS (error P^T (FDS - SDF) P, P = X X^T S the S-projector onto the orbital space,
    which is the AO commutator itself for the full space) the run converges to the same energy;
    a basis without d shells gives exactly the Cartesian run."""
    b = molecules.h2o("6-31g**")
    ref = lib.scf_run(oracle, b, options=oracle.default_options(**TIGHT), spherical=True)
    kw = dict(energy_tol=1e-9, density_tol=1e-7, diis_error_tol=1e-7, schwarz_thresh=1e-14, density_thresh=1e-14,
              max_iter=80)
    r = lib.scf_run(oracle, b, options=oracle.default_options(**kw), spherical=True)
    assert r["converged"] == 1 and abs(r["e_total"] - ref["e_total"]) < 1e-8
    assert r["last_diis_err"] < 1e-7
    b2 = molecules.h2o("6-31g")
    rs = lib.scf_run(oracle, b2, options=oracle.default_options(**kw), spherical=True)
    rc = lib.scf_run(oracle, b2, options=oracle.default_options(**kw))
    assert rs["n_mo"] == b2.n_aos and abs(rs["e_total"] - rc["e_total"]) < 1e-10


def test_orbital_space_refusals(oracle):
    b = molecules.h2o("6-31g**")
    with pytest.raises(lib.HfxError):
        lib.scf_run_batched(oracle, [b, b], options=oracle.default_options(
            orbital_space=b.spherical_transform().ctypes.data_as(lib._f64p), n_orbital_space=24))
    o = oracle.default_options(n_orbital_space=0)
    T = np.asfortranarray(b.spherical_transform())
    o.orbital_space = T.ctypes.data_as(lib._f64p)
    with pytest.raises(lib.HfxError):
        lib.scf_run(oracle, b, options=o)


@pytest.mark.gpu
def test_spherical_gpu_matches_oracle(engine_lib, oracle):
    b = molecules.ethane()
    print("spherical ethane", b.n_aos, flush=True)
    kw = dict(energy_tol=1e-11, density_tol=1e-9, use_diis=0, max_iter=150, schwarz_thresh=1e-14,
              density_thresh=1e-14, jk_flags=lib.HFX_JK_DETERMINISTIC)
    rg = lib.scf_run(engine_lib, b, options=engine_lib.default_options(**kw), spherical=True)
    ro = lib.scf_run(oracle, b, options=oracle.default_options(**{k: v for k, v in kw.items() if k != "jk_flags"}),
                     spherical=True)
    assert rg["converged"] == ro["converged"] == 1 and rg["n_mo"] == ro["n_mo"] == b.n_aos - 2
    assert abs(rg["e_total"] - ro["e_total"]) <= 1e-9
    assert np.abs(rg["F_alpha"] - ro["F_alpha"]).max() <= 1e-8
