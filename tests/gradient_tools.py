"""Helpers of the gradient tests: tight SCF settings, (P, W) from an SCF result, central
finite differences of SCF energies, and a torch.autograd.Function over the engine's SCF
energy whose backward is the engine's analytic RHF gradient (verification harness only)."""
import numpy as np

from hfx import Basis, lib

TIGHT = dict(energy_tol=1e-11, density_tol=1e-9, full_builds=1, diis_error_tol=1e-8, schwarz_thresh=1e-13,
             density_thresh=1e-13, max_iter=150)


def densities(result, basis):
    nocc = sum(basis.Z) // 2
    P = 2.0 * result["D_alpha"]
    W = lib.energy_weighted_density(result["C_alpha"], result["eps_alpha"], nocc)
    return P, W


def scf(library, basis, **kw):
    o = dict(TIGHT)
    o.update(kw)
    r = lib.scf_run(library, basis, options=library.default_options(**o))
    assert r["converged"] == 1
    return r


def analytic(library, basis, precision=lib.HFX_FP64, fp32_cut=0.0, **kw):
    r = scf(library, basis, **kw)
    P, W = densities(r, basis)
    e = lib.Engine(library, basis)
    tau = kw.get("schwarz_thresh", TIGHT["schwarz_thresh"])
    g, parts, st = lib.rhf_gradient(e, basis, P, W, schwarz_thresh=tau, density_thresh=tau, precision=precision,
                                    fp32_cut=fp32_cut)
    e.close()
    return r, g, parts, st


def finite_differences(library, basis, coords=None, h=1e-4, **kw):
    """Central differences of the SCF energy for the (atom, xyz) pairs in coords (all by default)."""
    if coords is None:
        coords = [(A, x) for A in range(len(basis.Z)) for x in range(3)]
    out = {}
    for (A, x) in coords:
        xp, xm = basis.xyz.copy(), basis.xyz.copy()
        xp[A, x] += h
        xm[A, x] -= h
        ep = scf(library, Basis(basis.name, basis.Z, xp), **kw)["e_total"]
        em = scf(library, Basis(basis.name, basis.Z, xm), **kw)["e_total"]
        out[(A, x)] = (ep - em) / (2 * h)
    return out


def finite_differences_batched(library, basis, h=1e-4, **kw):
    """Central differences for every coordinate, all displaced molecules in one batched SCF."""
    geoms = []
    for A in range(len(basis.Z)):
        for x in range(3):
            for sgn in (1.0, -1.0):
                g = basis.xyz.copy()
                g[A, x] += sgn * h
                geoms.append(Basis(basis.name, basis.Z, g))
    o = dict(TIGHT)
    o.update(kw)
    res = lib.scf_run_batched(library, geoms, options=library.default_options(**o))
    assert all(r["converged"] == 1 for r in res)
    e = np.array([r["e_total"] for r in res]).reshape(len(basis.Z), 3, 2)
    return (e[:, :, 0] - e[:, :, 1]) / (2 * h)


def scf_energy_function(library, basis_name, Z, **kw):
    """A torch.autograd.Function: coords (n_atoms, 3) float64 -> SCF energy, with the
    engine's analytic gradient as its backward."""
    import torch

    class ScfEnergy(torch.autograd.Function):
        @staticmethod
        def forward(ctx, coords):
            b = Basis(basis_name, Z, coords.detach().cpu().numpy())
            r, g, _, _ = analytic(library, b, **kw)
            ctx.save_for_backward(torch.from_numpy(g.copy()))
            return torch.tensor(r["e_total"], dtype=torch.float64)

        @staticmethod
        def backward(ctx, grad_out):
            (g,) = ctx.saved_tensors
            return grad_out * g

    return ScfEnergy


def autograd_check():
    """Backward of the torch function against central finite differences on four coordinates
    of the water dimer (6-31G**); prints one JSON line.  Imports torch before the engine."""
    import json

    import torch

    import molecules
    b = molecules.water_dimer("6-31g**")
    library = lib.load()
    kw = dict(full_builds=0, energy_tol=1e-10, density_tol=1e-8, schwarz_thresh=1e-12, density_thresh=1e-12)
    F = scf_energy_function(library, b.name, b.Z, **kw)
    x = torch.tensor(b.xyz, dtype=torch.float64, requires_grad=True)
    E = F.apply(x)
    E.backward()
    fd = finite_differences(library, b, coords=[(0, 0), (1, 1), (3, 2), (5, 0)], **kw)
    err = max(abs(x.grad[A, c].item() - v) for (A, c), v in fd.items())
    print(json.dumps({"n": len(fd), "max_err": err}), flush=True)
