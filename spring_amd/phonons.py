"""Harmonic phonons in FP64: dynamical matrix, eigensolve, group velocity, occupations.

Mirrors get_phonon / DynMatgv_SR (src/Kappa/Src/DynaMat_mod.f90:98-190, 486-709) for
non-polar crystals (no Ewald term):
  D(q)[(mu,a),(nu,b)] = sum_{N: basis(N)=nu} fc2[mu,N,a,b] e^{i q.R_N} / sqrt(m_mu m_nu)
  lambda, e = zheevd('V','U',D)     (numpy.linalg.eigh calls LAPACK ?heevd; lapack_gfort.f90:239-244)
  w = sqrt|lambda| (negative lambda beyond EPS -> -w, an imaginary mode)
  omega = w * wTHz (rad/ps);  1/omega = 1/(w wTHz);  n = 1/(exp(BEConv |w| / T) - 1)
  v = Re <e| dD/dq |e> / (2 w) * vg_ms   (m/s)
  at q = Gamma the three lowest modes get v = n = 1/omega = 0     (:629-654)

Output layout equals the reference's Phon arrays (phonon_mod.f90:33-46) with the Fortran
index order reversed: evec[q, s, row] = Evec(row, s, q); omega[q, s] = omega(s, q);
grp_vel[q, s, :] = grp_vel(:, s, q).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import constants as C
from .model import Crystal, IFC


@dataclass
class Phonons:
    evec: np.ndarray        # (nq, ndof, ndof) complex128: evec[q, s, row]
    omega: np.ndarray       # (nq, ndof) rad/ps
    inv_omega: np.ndarray   # (nq, ndof) 1/(rad/ps), 0 for Gamma acoustic
    nbe: np.ndarray         # (nq, ndof)
    grp_vel: np.ndarray     # (nq, ndof, 3) m/s
    T: float

    @property
    def nq(self):
        return self.omega.shape[0]


def dynamical_matrix(cr: Crystal, fc2: IFC, q_cart: np.ndarray, with_derivative=True):
    """D(q) and dD/dq for q (nq, 3) in 1/Angstrom; rows (mu,a), columns (nu,b)."""
    nb = cr.nb
    nq = q_cart.shape[0]
    ndof = 3 * nb
    D = np.zeros((nq, ndof, ndof), complex)
    dD = np.zeros((3, nq, ndof, ndof), complex) if with_derivative else None
    osm = cr.one_by_sqrt_mass
    for mu in range(nb):
        for n in range(int(fc2.atm_num[mu])):
            cell = fc2.atm_indx[mu, n, :3].astype(float)
            nu = int(fc2.atm_indx[mu, n, 3])
            R = cr.latvec @ cell
            ph = np.exp(1j * (q_cart @ R))                      # (nq,)
            blk = fc2.fc[mu, n] * (osm[mu] * osm[nu])           # (a, b)
            D[:, 3 * mu:3 * mu + 3, 3 * nu:3 * nu + 3] += ph[:, None, None] * blk[None]
            if with_derivative:
                for k in range(3):
                    dD[k, :, 3 * mu:3 * mu + 3, 3 * nu:3 * nu + 3] += \
                        (1j * R[k] * ph)[:, None, None] * blk[None]
    return D, dD


def solve(cr: Crystal, fc2: IFC, q_cart: np.ndarray, q_int: np.ndarray, T: float,
          chunk: int = 4096) -> Phonons:
    nq = q_cart.shape[0]
    ndof = cr.ndof
    evec = np.zeros((nq, ndof, ndof), complex)
    omega = np.zeros((nq, ndof))
    inv_omega = np.zeros((nq, ndof))
    nbe = np.zeros((nq, ndof))
    vel = np.zeros((nq, ndof, 3))
    for lo in range(0, nq, chunk):
        hi = min(nq, lo + chunk)
        D, dD = dynamical_matrix(cr, fc2, q_cart[lo:hi])
        lam, V = np.linalg.eigh(D, UPLO='U')                    # V[q, row, s]; zheevd uplo='U'
        neg = (np.abs(lam) > C.EPS) & (lam < 0.0)
        w = np.sqrt(np.abs(lam))
        w = np.where(neg, -w, w)
        gamma = np.all(q_int[lo:hi] == 0, axis=1)
        # <e|dD|e> for every mode
        ev = np.einsum('kqrc,qcs->kqrs', dD, V)
        cv = np.einsum('qrs,kqrs->qsk', np.conj(V), ev)          # (q, s, 3)
        with np.errstate(divide='ignore', invalid='ignore'):
            v = np.real(cv) / (2.0 * w[:, :, None]) * C.VG_MS
            n = 1.0 / (np.exp(C.BE_CONV * np.abs(w) / T) - 1.0)
            iw = 1.0 / (w * C.W_THZ)
        acoustic = np.zeros_like(w, dtype=bool)
        acoustic[gamma, :3] = True
        v[acoustic] = 0.0
        n[acoustic] = 0.0
        iw[acoustic] = 0.0
        evec[lo:hi] = np.transpose(V, (0, 2, 1))
        omega[lo:hi] = w * C.W_THZ
        inv_omega[lo:hi] = iw
        nbe[lo:hi] = n
        vel[lo:hi] = v
    return Phonons(evec, omega, inv_omega, nbe, vel, float(T))
