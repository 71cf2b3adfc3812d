"""Basis-set loading and the flat shell/primitive layout handed to the engine.

Restates the reference's basis reader and shell flattening:

* ``load_basis`` follows ``BasisSet::BasisSet`` / ``BasisSet::readBasis``
  (reference ``src/Basis.cpp:14-86, 89-171``): shells in atom order, then in file order;
  an entry whose ``angular_momentum`` list has several values yields one shell per value,
  in list order, sharing the exponents and taking coefficient row ``i``
  (``Basis.cpp:125-159``); extra general-contraction rows are not read
  (``Basis.cpp:145-146``).
* Cartesian order and per-primitive, per-component normalization follow
  ``Basis.cpp:47-77``: ``i`` runs from ``l`` down to 0, ``j`` from ``l-i`` down to 0,
  ``k = l-i-j``; ``N = (2a/pi)^(3/4) (4a)^(l/2) / sqrt((2i-1)!!(2j-1)!!(2k-1)!!)``.
  The contracted function is not renormalized.
* ``Basis.flat()`` is the struct-of-arrays layout of ``IntegralEngine::IntegralEngine``
  (reference ``src/IntegralEngine.cpp:13-75``, fields ``include/IntegralEngine.hpp:145-168``).
"""
import json
import math
import os
from dataclasses import dataclass, field

import numpy as np

DATA_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "basis_data")

ELEMENT_Z = {"H": 1, "He": 2, "Li": 3, "Be": 4, "B": 5, "C": 6, "N": 7, "O": 8, "F": 9, "Ne": 10}
Z_ELEMENT = {v: k for k, v in ELEMENT_Z.items()}


def dfact(n):
    """Double factorial with dfact(n<=0) = 1 (reference ``include/Utils.hpp:43-50``)."""
    if n <= 0:
        return 1
    r = 1
    i = n
    while i >= 1:
        r *= i
        i -= 2
    return r


def cart_components(l):
    """Cartesian exponents of a shell in the reference's AO order (``Basis.cpp:47-59``)."""
    out = []
    for i in range(l, -1, -1):
        for j in range(l - i, -1, -1):
            out.append((i, j, l - i - j))
    return out


def basis_file_name(name):
    """Name mangling of ``readBasis`` (``Basis.cpp:92-103``): lower case, ``*`` -> ``_st_``."""
    return name.lower().replace("*", "_st_")


@dataclass
class Shell:
    l: int
    exponents: list
    coefficients: list
    atom: int
    center: tuple
    norm: list = field(default_factory=list)  # [nprim * nao], prim-major

    @property
    def nprim(self):
        return len(self.exponents)

    @property
    def nao(self):
        return (self.l + 1) * (self.l + 2) // 2


def read_basis(name, elements, general="reference"):
    """Raw shells per element: ``{Z: [(l, exponents, coefficients), ...]}``.

    general = "reference": the reference's reading (src/Basis.cpp:125-159): an entry with angular
    momenta [l_0, l_1, ...] gives one shell per l_i with coefficient row i, so the extra rows of
    a general contraction (cc-pVDZ's s and p entries) are not read.
    general = "expand": every coefficient row of an entry becomes a shell of its own (with the
    entry's angular momentum for a single-l entry, l_i for row i of an sp-style entry), and the
    primitives whose coefficient in that row is zero are dropped."""
    if general not in ("reference", "expand"):
        raise ValueError("general must be 'reference' or 'expand'")
    path = os.path.join(DATA_DIR, basis_file_name(name) + ".json")
    if not os.path.exists(path):
        raise ValueError("Basis set '%s' not found." % name)
    with open(path) as f:
        data = json.load(f)
    out = {}
    for z in elements:
        key = str(z)
        if key not in data["elements"]:
            raise ValueError("%s not found in basis set '%s'." % (Z_ELEMENT.get(z, z), name))
        raw = []
        for sh in data["elements"][key]["electron_shells"]:
            am = sh["angular_momentum"]
            exps = [float(s) for s in sh["exponents"]]
            if general == "reference" or len(sh["coefficients"]) <= len(am):
                for i, l in enumerate(am):
                    coefs = [float(s) for s in sh["coefficients"][i]]
                    raw.append((int(l), exps, coefs))
            else:
                for i, row in enumerate(sh["coefficients"]):
                    l = int(am[i] if len(am) > 1 else am[0])
                    keep = [k for k, c in enumerate(row) if float(c) != 0.0]
                    raw.append((l, [exps[k] for k in keep], [float(row[k]) for k in keep]))
        out[z] = raw
    return out


def _norm_factors(l, exponents):
    comps = cart_components(l)
    out = []
    for a in exponents:
        for (i, j, k) in comps:
            pre1 = math.pow(2.0 * a / math.pi, 0.75)
            pre2 = math.pow(4.0 * a, (i + j + k) / 2.0)
            den = math.sqrt(float(dfact(2 * i - 1) * dfact(2 * j - 1) * dfact(2 * k - 1)))
            out.append(pre1 * pre2 / den)
    return out


class Basis:
    """A basis on a molecule: an ordered list of shells plus the flat engine layout."""

    def __init__(self, name, Z, xyz, general="reference"):
        self.name = name
        self.Z = [int(z) for z in Z]
        self.xyz = np.asarray(xyz, dtype=np.float64).reshape(-1, 3)
        raw = read_basis(name, sorted(set(self.Z)), general)
        self.shells = []
        for ia, z in enumerate(self.Z):
            for (l, exps, coefs) in raw[z]:
                sh = Shell(l=l, exponents=list(exps), coefficients=list(coefs), atom=ia,
                           center=tuple(float(c) for c in self.xyz[ia]))
                sh.norm = _norm_factors(l, exps)
                self.shells.append(sh)
        self._flat = None

    @property
    def n_shells(self):
        return len(self.shells)

    @property
    def n_aos(self):
        return sum(s.nao for s in self.shells)

    @property
    def n_prims(self):
        return sum(s.nprim for s in self.shells)

    @property
    def max_l(self):
        return max(s.l for s in self.shells)

    def flat(self):
        """Struct-of-arrays layout (``IntegralEngine.cpp:13-75``) as numpy arrays."""
        if self._flat is not None:
            return self._flat
        nS = self.n_shells
        l = np.zeros(nS, np.int32)
        n_prim = np.zeros(nS, np.int32)
        prim_off = np.zeros(nS, np.int32)
        ao_off = np.zeros(nS, np.int32)
        norm_off = np.zeros(nS, np.int32)
        centers = np.zeros((nS, 3), np.float64)
        alpha, coeff, norm, cart = [], [], [], []
        po = ao = no = 0
        for s, sh in enumerate(self.shells):
            l[s] = sh.l
            n_prim[s] = sh.nprim
            prim_off[s] = po
            ao_off[s] = ao
            norm_off[s] = no
            centers[s] = sh.center
            alpha += sh.exponents
            coeff += sh.coefficients
            norm += sh.norm
            cart += cart_components(sh.l)
            po += sh.nprim
            ao += sh.nao
            no += sh.nprim * sh.nao
        self._flat = dict(
            n_shells=nS, n_prims=po, n_aos=ao, n_norm=no,
            l=l, n_prim=n_prim, prim_offset=prim_off, ao_offset=ao_off, norm_offset=norm_off,
            centers=np.ascontiguousarray(centers.reshape(-1)),
            alpha=np.asarray(alpha, np.float64), coeff=np.asarray(coeff, np.float64),
            norm=np.asarray(norm, np.float64),
            cart=np.ascontiguousarray(np.asarray(cart, np.int32).reshape(-1)),
            atom=np.asarray([sh.atom for sh in self.shells], np.int32),
        )
        return self._flat

    def spherical_transform(self):
        """Cartesian-to-spherical transform T (n_aos x n_sph, column-major order of the shells):
        the columns of a shell are its real solid harmonics in terms of the shell's Cartesian
        functions as the engine normalizes them (every component normalized by itself,
        ``Basis.cpp:47-77``).  s and p: the Cartesian functions themselves (p in x, y, z
        order).  d, m = -2..2: xy, yz, (2zz - xx - yy)/2, xz, sqrt(3)/2 (xx - yy); with the
        components normalized alike, <xx|yy> = <xx|xx>/3, so every column has the norm of the
        shell's xy component and the five are orthogonal.  The reference has no spherical
        option; this is the orbital space of ``hfx_scf_options.orbital_space``."""
        blocks = []
        h = math.sqrt(3.0) / 2.0
        d = np.array([[0, 0, -0.5, 0, h],     # xx
                      [1, 0, 0, 0, 0],        # xy
                      [0, 0, 0, 1, 0],        # xz
                      [0, 0, -0.5, 0, -h],    # yy
                      [0, 1, 0, 0, 0],        # yz
                      [0, 0, 1, 0, 0]],       # zz
                     dtype=np.float64)
        for sh in self.shells:
            if sh.l <= 1:
                blocks.append(np.eye(sh.nao))
            elif sh.l == 2:
                blocks.append(d)
            else:
                raise ValueError("spherical_transform: l <= 2 only")
        n = self.n_aos
        m = sum(b.shape[1] for b in blocks)
        T = np.zeros((n, m))
        r = c = 0
        for b in blocks:
            T[r:r + b.shape[0], c:c + b.shape[1]] = b
            r += b.shape[0]
            c += b.shape[1]
        return T

    def nuclear_repulsion(self):
        """``Molecule::nuclearRepulsion`` (reference ``src/Molecule.cpp:75-89``)."""
        e = 0.0
        n = len(self.Z)
        for i in range(n):
            for j in range(i + 1, n):
                d = float(np.linalg.norm(self.xyz[i] - self.xyz[j]))
                e += (self.Z[i] * self.Z[j]) / d
        return e

    def n_electrons(self, charge=0):
        """``Molecule::countElectrons`` (reference ``src/Molecule.cpp:67-72``)."""
        return sum(self.Z) - charge
