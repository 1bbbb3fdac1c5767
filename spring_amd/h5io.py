"""ctypes binding of libspring_h5.so (include/spring_h5.h): the force-constant files kappa.x
reads (FC2nd_/FC3rd_/FC4th_<prefix><T>K_F.h5) and the kappa file it writes
(kappa_<prefix><T>K.h5), in the reference's layout (src/Kappa/Src/hdf5_wrap_mod.f90:40-337,
650-765; src/Kappa/Src/main.f90:192-235, 773-798).

IFC arrays here use the package's convention (model.IFC: C arrays with the Fortran index order
reversed, 0-based basis index in atm_indx); the files hold the basis index 1-based.
"""
from __future__ import annotations

import ctypes as ct
import os

import numpy as np

from .model import IFC

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libspring_h5.so")
SYMBOLS = ["spr_h5_fc_dims", "spr_h5_read_fc", "spr_h5_write_fc", "spr_h5_write_kappa", "spr_h5_read_kappa",
           "spr_h5_write_spectral", "spr_h5_read_spectral", "spr_h5_last_error"]

_lib = None


class H5Error(RuntimeError):
    pass


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise H5Error(f"{LIB_PATH} is not built: run __graft_entry__.build()")
        lib = ct.CDLL(LIB_PATH)
        lib.spr_h5_last_error.restype = ct.c_char_p
        _lib = lib
    return _lib


def _p(a):
    return a.ctypes.data_as(ct.c_void_p)


def _check(rc, what):
    if rc != 0:
        raise H5Error(f"{what}: {load().spr_h5_last_error().decode()}")


def fc_file_name(order: int, prefix: str, T: float) -> str:
    """FC2nd_/FC3rd_/FC4th_ + prefix + T (F6.1, blanks trimmed) + K_F.h5 (main.f90:182-195,235)."""
    tag = {2: "FC2nd_", 3: "FC3rd_", 4: "FC4th_"}[order]
    return f"{tag}{prefix}{T:6.1f}".replace(" ", "") + "K_F.h5"


def kappa_file_name(prefix: str, T: float) -> str:
    """kappa_ + prefix + T + K.h5 (main.f90:198)."""
    return f"kappa_{prefix}{T:6.1f}".replace(" ", "") + "K.h5"


def read_fc(path: str, order: int) -> IFC:
    L = load()
    nb, natm, lead = ct.c_int(), ct.c_int(), ct.c_int()
    _check(L.spr_h5_fc_dims(path.encode(), order, ct.byref(nb), ct.byref(natm), ct.byref(lead)), "spr_h5_fc_dims")
    nb, natm = nb.value, natm.value
    num = np.zeros(nb, np.int32)
    indx = np.zeros((nb, natm, 4), np.int32)
    fc = np.zeros((nb,) + (natm,) * (order - 1) + (3,) * order)
    _check(L.spr_h5_read_fc(path.encode(), order, nb, natm, _p(num), _p(indx), _p(fc)), "spr_h5_read_fc")
    indx[..., 3] -= 1
    return IFC(order, num, indx, fc, {"file": path})


def write_fc(path: str, ifc: IFC) -> None:
    num = np.ascontiguousarray(ifc.atm_num, np.int32)
    indx = np.ascontiguousarray(ifc.atm_indx, np.int32).copy()
    indx[..., 3] += 1
    fc = np.ascontiguousarray(ifc.fc, np.float64)
    nb, natm = indx.shape[0], indx.shape[1]
    _check(load().spr_h5_write_fc(path.encode(), ifc.order, nb, natm, _p(num), _p(indx), _p(fc)), "spr_h5_write_fc")


def write_kappa(path: str, kappa, omega_q0, lw, p3=None, mode_k=None) -> None:
    """kappa (3,3), omega_q0 and lw (Nq0, Ndof), p3 (Nq0, Ndof) or None, mode_k (Ndof, 9) or None."""
    k = np.ascontiguousarray(kappa, np.float64).reshape(9)
    w = np.ascontiguousarray(omega_q0, np.float64)
    l = np.ascontiguousarray(lw, np.float64)
    nq0, nd = w.shape
    p = None if p3 is None else np.ascontiguousarray(p3, np.float64)
    m = None if mode_k is None else np.ascontiguousarray(mode_k, np.float64)
    _check(load().spr_h5_write_kappa(path.encode(), _p(k), nd, nq0, _p(w), _p(l), None if p is None else _p(p),
                                     None if m is None else _p(m)), "spr_h5_write_kappa")


def write_spectral(path: str, spectral) -> None:
    """Adds SpectralResolvedKappa to an existing kappa file; spectral (Ndof Nq0, 11) as
    GpuContext.kappa_spectral returns it."""
    a = np.ascontiguousarray(spectral, np.float64)
    nq0_nd = a.shape[0]
    _check(load().spr_h5_write_spectral(path.encode(), nq0_nd, 1, _p(a)), "spr_h5_write_spectral")


def read_spectral(path: str, ndof: int, nq0: int):
    a = np.zeros((ndof * nq0, 11))
    _check(load().spr_h5_read_spectral(path.encode(), ndof, nq0, _p(a)), "spr_h5_read_spectral")
    return a


def read_kappa(path: str, ndof: int, nq0: int, p3: bool = False, mode_k: bool = False) -> dict:
    k = np.zeros(9)
    w = np.zeros((nq0, ndof))
    l = np.zeros((nq0, ndof))
    p = np.zeros((nq0, ndof)) if p3 else None
    m = np.zeros((ndof, 9)) if mode_k else None
    _check(load().spr_h5_read_kappa(path.encode(), ndof, nq0, _p(k), _p(w), _p(l), None if p is None else _p(p),
                                    None if m is None else _p(m)), "spr_h5_read_kappa")
    return dict(kappa=k.reshape(3, 3), omega=w, lw=l, p3=p, mode_k=m)
