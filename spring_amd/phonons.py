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
    """D(q) and dD/dq for q (nq, 3) in 1/Angstrom; rows (mu,a), columns (nu,b).

    Every neighbour term is scattered into a dense (terms, ndof^2) table once, so D is one
    complex matrix product phases(q, terms) @ table (and dD/dq_k the same with i R_k)."""
    nb = cr.nb
    nq = q_cart.shape[0]
    ndof = 3 * nb
    osm = cr.one_by_sqrt_mass
    rows, Rs = [], []
    for mu in range(nb):
        for n in range(int(fc2.atm_num[mu])):
            cell = fc2.atm_indx[mu, n, :3].astype(float)
            nu = int(fc2.atm_indx[mu, n, 3])
            blk = np.zeros((ndof, ndof))
            blk[3 * mu:3 * mu + 3, 3 * nu:3 * nu + 3] = fc2.fc[mu, n] * (osm[mu] * osm[nu])
            rows.append(blk.ravel())
            Rs.append(cr.latvec @ cell)
    table = np.array(rows)                                    # (terms, ndof^2)
    R = np.array(Rs)                                          # (terms, 3)
    ph = np.exp(1j * (q_cart @ R.T))                          # (nq, terms)
    D = (ph @ table).reshape(nq, ndof, ndof)
    dD = None
    if with_derivative:
        dD = np.stack([((1j * ph * R[None, :, k]) @ table).reshape(nq, ndof, ndof) for k in range(3)])
    return D, dD


def _solve_chunk(cr, fc2, q_cart, q_int, T):
    D, dD = dynamical_matrix(cr, fc2, q_cart)
    lam, V = np.linalg.eigh(D, UPLO='U')                        # V[q, row, s]; zheevd uplo='U'
    neg = (np.abs(lam) > C.EPS) & (lam < 0.0)
    w = np.sqrt(np.abs(lam))
    w = np.where(neg, -w, w)
    gamma = np.all(q_int == 0, axis=1)
    # <e|dD/dq_k|e> for every mode: (q, s, k)
    cv = np.transpose((np.conj(V)[None] * (dD @ V[None])).sum(axis=2), (1, 2, 0))
    with np.errstate(divide='ignore', invalid='ignore'):
        v = np.real(cv) / (2.0 * w[:, :, None]) * C.VG_MS
        n = 1.0 / (np.exp(C.BE_CONV * np.abs(w) / T) - 1.0)
        iw = 1.0 / (w * C.W_THZ)
    acoustic = np.zeros_like(w, dtype=bool)
    acoustic[gamma, :3] = True
    v[acoustic] = 0.0
    n[acoustic] = 0.0
    iw[acoustic] = 0.0
    return np.transpose(V, (0, 2, 1)), w * C.W_THZ, iw, n, v


def solve(cr: Crystal, fc2: IFC, q_cart: np.ndarray, q_int: np.ndarray, T: float,
          chunk: int = 512, threads: int | None = None) -> Phonons:
    """Phonons of every q, in chunks spread over host threads (LAPACK releases the GIL)."""
    from concurrent.futures import ThreadPoolExecutor
    import os
    nq = q_cart.shape[0]
    ndof = cr.ndof
    evec = np.zeros((nq, ndof, ndof), complex)
    omega = np.zeros((nq, ndof))
    inv_omega = np.zeros((nq, ndof))
    nbe = np.zeros((nq, ndof))
    vel = np.zeros((nq, ndof, 3))
    los = list(range(0, nq, chunk))

    def one(lo):
        hi = min(nq, lo + chunk)
        evec[lo:hi], omega[lo:hi], inv_omega[lo:hi], nbe[lo:hi], vel[lo:hi] = \
            _solve_chunk(cr, fc2, q_cart[lo:hi], q_int[lo:hi], T)

    nt = threads or max(1, min(16, os.cpu_count() or 1, len(los)))
    if nt == 1:
        for lo in los:
            one(lo)
    else:
        with ThreadPoolExecutor(nt) as ex:
            list(ex.map(one, los))
    return Phonons(evec, omega, inv_omega, nbe, vel, float(T))
