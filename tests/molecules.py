"""Small test molecules (coordinates in Bohr) shared by the CPU and GPU tests."""
import numpy as np

from hfx import Basis, geometry


def h2():
    return Basis("sto-3g", [1, 1], [[0, 0, 0], [0, 0, 1.4]])


def h2o(name="sto-3g"):
    Z, x = geometry.h2o_sto3g()
    return Basis(name, Z, x)


def crawford_h2o():
    """Water of the Crawford group's SCF programming project (STO-3G)."""
    Z = [8, 1, 1]
    x = [[0.0, -0.143225816552, 0.0], [1.638036840407, 1.136548822547, 0.0],
         [-1.638036840407, 1.136548822547, 0.0]]
    return Basis("sto-3g", Z, x)


def water_dimer(name="6-31g**"):
    Z, x = geometry.water_cluster_grid(shape=(2, 1, 1), spacing=5.6, jitter=0.4, seed=7)
    return Basis(name, Z, x)


def methane(name="cc-pvdz-expanded"):
    d = 2.06 / np.sqrt(3.0)
    Z = [6, 1, 1, 1, 1]
    x = [[0.1, 0.05, -0.02], [d, d, d], [-d, -d, d], [-d, d, -d], [d, -d, -d]]
    return Basis(name, Z, x)


def ethane(name="cc-pvdz-expanded"):
    Z, x = geometry.alkane_chain(n_carbon=2, seed=11)
    return Basis(name, Z, x)


def oh_radical(name="6-31g**"):
    return Basis(name, [8, 1], [[0.0, 0.0, 0.0], [0.1, 0.2, 1.83]])
