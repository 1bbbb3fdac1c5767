"""Loads tests/golden/*.npz into the objects the library and the oracle take."""
import os
from types import SimpleNamespace

import numpy as np

from spring_amd.model import IFC, Crystal, Model

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name="rocksalt_3x3x3.npz"):
    d = np.load(os.path.join(HERE, name), allow_pickle=False)
    cr = Crystal(d["latvec"], d["basis_frac"], d["mass"], np.array([0, 1]), "golden")
    fc3 = IFC(3, d["atm_num3"], d["atm_indx3"], d["fc3"])
    m = Model(cr, None, fc3, None, "golden")
    qm = SimpleNamespace(mesh=d["mesh"], nq=int(d["q_int"].shape[0]), q_int=d["q_int"], q_cart=d["q_cart"],
                         irr_idx=d["irr_idx"], nq0=int(d["irr_idx"].shape[0]), pos=d["pos"],
                         multiplicity=d["multiplicity"], rot_cart=d["rot_cart"], time_rev=0)
    ph = SimpleNamespace(evec=d["evec"], omega=d["omega"], inv_omega=d["inv_omega"], nbe=d["nbe"],
                         grp_vel=d["grp_vel"], nq=int(d["omega"].shape[0]))
    return m, qm, ph, d
