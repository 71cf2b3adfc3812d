"""Seeded synthetic molecules for the benchmark and test workloads (coordinates in Bohr).

* ``h2o_sto3g``: the fixed water geometry of the H2O/STO-3G case.
* ``water_cluster_grid``: rigid waters (O-H 1.809 Bohr, HOH 104.5 deg), each uniformly
  randomly rotated (Shoemake quaternion) with its O atom on a grid, jittered by a seeded
  uniform offset per coordinate.  (H2O)16 uses a 4x2x2 grid at 5.6 Bohr spacing with
  +-0.4 Bohr jitter; (H2O)8 a 2x2x2 grid.  "Water centre" means the O atom.
* ``alkane_chain``: all-trans CnH(2n+2): C-C 2.91 Bohr, C-H 2.06 Bohr, tetrahedral angles,
  plus a seeded +-0.02 Bohr jitter on every coordinate.
* ``water_cluster_cube``: n waters with O-O >= 5.2 Bohr, placed by sequential rejection
  sampling in a cube of side 4 n^(1/3) Bohr (at most 100000 tries per water).
"""
import math

import numpy as np

OH = 1.809
HOH = math.radians(104.5)


def h2o_sto3g():
    Z = [1, 1, 8]
    xyz = np.array([[1.4327076887, -0.9610871670, 0.0],
                    [-1.4327076887, -0.9610871670, 0.0],
                    [0.0, 0.2403536480, 0.0]])
    return Z, xyz


def _monomer():
    """Rigid water in its own frame: O at the origin, H atoms in the xy plane."""
    h = HOH / 2.0
    o = np.zeros(3)
    h1 = np.array([OH * math.sin(h), OH * math.cos(h), 0.0])
    h2 = np.array([-OH * math.sin(h), OH * math.cos(h), 0.0])
    return o, h1, h2


def _random_rotation(rng):
    """Uniform random rotation matrix from a Shoemake quaternion."""
    u1, u2, u3 = rng.random(3)
    q = np.array([math.sqrt(1 - u1) * math.sin(2 * math.pi * u2),
                  math.sqrt(1 - u1) * math.cos(2 * math.pi * u2),
                  math.sqrt(u1) * math.sin(2 * math.pi * u3),
                  math.sqrt(u1) * math.cos(2 * math.pi * u3)])
    x, y, z, w = q
    return np.array([
        [1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
        [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
        [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)],
    ])


def _water_at(o_pos, rot):
    o, h1, h2 = _monomer()
    return [o_pos + rot @ o, o_pos + rot @ h1, o_pos + rot @ h2]


def water_cluster_grid(shape=(4, 2, 2), spacing=5.6, jitter=0.4, seed=1002):
    rng = np.random.default_rng(seed)
    Z, xyz = [], []
    for ix in range(shape[0]):
        for iy in range(shape[1]):
            for iz in range(shape[2]):
                centre = np.array([ix, iy, iz], float) * spacing
                centre = centre + rng.uniform(-jitter, jitter, 3)
                rot = _random_rotation(rng)
                o, h1, h2 = _water_at(centre, rot)
                Z += [8, 1, 1]
                xyz += [o, h1, h2]
    return Z, np.array(xyz)


def alkane_chain(n_carbon=20, cc=2.91, ch=2.06, jitter=0.02, seed=1003):
    """All-trans zig-zag alkane along x; the carbon backbone lies in the xy plane."""
    rng = np.random.default_rng(seed)
    tet = math.acos(-1.0 / 3.0)  # 109.47 deg
    half = (math.pi - tet) / 2.0
    dx = cc * math.cos(half)
    dy = cc * math.sin(half)
    carbons = []
    for i in range(n_carbon):
        carbons.append(np.array([i * dx, (dy / 2.0) * (1 if i % 2 == 0 else -1), 0.0]))
    Z, xyz = [], []
    for c in carbons:
        Z.append(6)
        xyz.append(c)
    # Two H per carbon, out of the backbone plane, bisecting the exterior of the C-C-C angle.
    hs = []
    for i, c in enumerate(carbons):
        side = 1.0 if i % 2 == 0 else -1.0
        # direction away from the backbone in the plane
        out = np.array([0.0, side, 0.0])
        ang = tet / 2.0
        for sz in (1.0, -1.0):
            d = out * math.cos(ang) + np.array([0.0, 0.0, sz]) * math.sin(ang)
            hs.append(c + ch * d / np.linalg.norm(d))
    # Terminal H along the chain axis continuation (one per end carbon).
    for end, sgn in ((0, -1.0), (n_carbon - 1, 1.0)):
        c = carbons[end]
        side = 1.0 if end % 2 == 0 else -1.0
        d = np.array([sgn * math.cos(half), -side * math.sin(half), 0.0])
        hs.append(c + ch * d / np.linalg.norm(d))
    for h in hs:
        Z.append(1)
        xyz.append(h)
    xyz = np.array(xyz) + rng.uniform(-jitter, jitter, (len(Z), 3))
    return Z, xyz


def water_cluster_cube(n, rng, min_oo=5.2, max_tries=2000, max_restarts=1000):
    """Sequential rejection sampling of O positions; the whole cluster restarts when one
    water cannot be placed within ``max_tries`` draws (dense clusters jam otherwise)."""
    side = 4.0 * n ** (1.0 / 3.0)
    for _r in range(max_restarts):
        os_ = []
        for _ in range(n):
            for _t in range(max_tries):
                p = rng.uniform(0.0, side, 3)
                if all(np.linalg.norm(p - q) >= min_oo for q in os_):
                    os_.append(p)
                    break
            else:
                break
        if len(os_) == n:
            break
    else:
        raise RuntimeError("rejection sampling failed to place %d waters" % n)
    Z, xyz = [], []
    for p in os_:
        rot = _random_rotation(rng)
        o, h1, h2 = _water_at(p, rot)
        Z += [8, 1, 1]
        xyz += [o, h1, h2]
    return Z, np.array(xyz)


def water_batch(n_mol=64, seed=1004, nmin=1, nmax=8):
    """Independent water clusters with n ~ U{nmin..nmax} waters each."""
    rng = np.random.default_rng(seed)
    mols = []
    for _ in range(n_mol):
        n = int(rng.integers(nmin, nmax + 1))
        mols.append(water_cluster_cube(n, rng))
    return mols


def min_distance(xyz):
    xyz = np.asarray(xyz)
    d = np.linalg.norm(xyz[:, None, :] - xyz[None, :, :], axis=-1)
    d[np.diag_indices(len(xyz))] = np.inf
    return float(d.min())
