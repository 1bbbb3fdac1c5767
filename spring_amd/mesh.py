"""Gamma-centred q-mesh, tetrahedra, crystal point group and the irreducible wedge.

Mirrors q_tetrahedron (src/Kappa/Src/TetrahedronQ_m.f90:35-331):
  * grid point i (0-based here, 1-based there) has folded integer address
    q_int in (-N/2, N/2] per axis (IndxChng_mod.f90 fold_indx) and
    i = sum_k mod(q_int_k, N_k) * prod_{j<k} N_j            (TetrahedronQ_m.f90:128)
  * q_pnt = (G / N) q_int in 1/Angstrom                       (:138)
  * six tetrahedra per mesh cube, vertex order of :172-188
  * irreducible q0: orbits of the grid under the crystal's point group (and optionally
    time reversal); the representative is the smallest grid index of each orbit, as
    spg_get_ir_reciprocal_mesh returns it (:261-313), with `pos` and `multiplicity`.

The point group is found from the structure itself (no spglib here): every integer
matrix that preserves the lattice metric and maps the basis onto itself (with some
fractional translation) is kept.
"""
from __future__ import annotations

import itertools
from dataclasses import dataclass

import numpy as np

from .model import Crystal


@dataclass
class QMesh:
    mesh: np.ndarray            # (3,) int
    q_int: np.ndarray           # (nq, 3) int32 folded addresses
    q_cart: np.ndarray          # (nq, 3) 1/Angstrom
    tetra: np.ndarray           # (6 nq, 4) int32 grid indices
    irr_idx: np.ndarray         # (nq0,) grid index of each irreducible point
    irr_q_int: np.ndarray       # (nq0, 3)
    pos: np.ndarray             # (nq,) irreducible position of every grid point
    multiplicity: np.ndarray    # (nq0,)
    rot_frac: np.ndarray        # (nsym, 3, 3) direct-space rotations, fractional
    rot_cart: np.ndarray        # (nsym, 3, 3) Cartesian rotations
    time_rev: int

    @property
    def nq(self) -> int:
        return self.q_int.shape[0]

    @property
    def nq0(self) -> int:
        return self.irr_idx.shape[0]

    def index(self, k: np.ndarray) -> np.ndarray:
        """Grid index of integer addresses k (..., 3), any representative."""
        m = self.mesh
        k = np.mod(k, m)
        return (k[..., 0] + m[0] * (k[..., 1] + m[1] * k[..., 2])).astype(np.int64)


def fold(k, mesh):
    """Fold integer addresses into (-N/2, N/2] (IndxChng_mod.f90:85-107)."""
    k = np.mod(np.asarray(k), mesh)
    return np.where(k > mesh // 2, k - mesh, k)


def point_group(cr: Crystal, symprec: float = 1e-5):
    """Rotations (fractional and Cartesian) of the crystal's point group."""
    A = cr.latvec
    metric = A.T @ A
    cands = []
    for entries in itertools.product((-1, 0, 1), repeat=9):
        R = np.array(entries, float).reshape(3, 3)
        if abs(abs(np.linalg.det(R)) - 1.0) > 1e-9:
            continue
        if np.max(np.abs(R.T @ metric @ R - metric)) > symprec * np.max(np.abs(metric)) * 10:
            continue
        cands.append(R)
    frac = cr.basis_frac
    spec = cr.species
    rots = []
    for R in cands:
        rp = frac @ R.T
        ok_any = False
        for j in range(cr.nb):
            if spec[j] != spec[0]:
                continue
            t = frac[j] - rp[0]
            mapped = rp + t
            ok = True
            for i in range(cr.nb):
                d = mapped[i] - frac
                d = d - np.round(d)
                dist = np.linalg.norm(d @ A.T, axis=1)
                hit = (dist < 1e-3 * 10) & (spec == spec[i])
                if not hit.any():
                    ok = False
                    break
            if ok:
                ok_any = True
                break
        if ok_any:
            rots.append(R.astype(np.int64))
    rot_frac = np.array(rots)
    rot_cart = np.array([A @ R @ np.linalg.inv(A) for R in rot_frac])
    return rot_frac, rot_cart


def make_mesh(cr: Crystal, mesh, time_rev: int = 0, symmetry: bool = True) -> QMesh:
    mesh = np.asarray(mesh, np.int64)
    n1, n2, n3 = (int(x) for x in mesh)
    nq = n1 * n2 * n3
    ii = np.arange(nq)
    raw = np.stack([ii % n1, (ii // n1) % n2, ii // (n1 * n2)], axis=1)
    q_int = fold(raw, mesh).astype(np.int32)
    q_cart = (q_int / mesh) @ cr.G.T
    # tetrahedra: cube corners (TetrahedronQ_m.f90:103-110) and six tetrahedra (:172-188)
    box = np.array([[0, 0, 0], [1, 0, 0], [0, 0, 1], [0, 1, 1],
                    [1, 0, 1], [1, 1, 1], [1, 1, 0], [0, 1, 0]])
    tets = np.array([[0, 1, 2, 3], [2, 3, 4, 1], [5, 3, 4, 1],
                     [5, 3, 6, 1], [6, 3, 7, 1], [0, 3, 7, 1]])
    # loop order of the reference: z outer, y, x inner -> cube index == grid index of raw
    corner = fold(raw[:, None, :] + box[None, :, :], mesh)
    corner_idx = (np.mod(corner[..., 0], n1) + n1 * (np.mod(corner[..., 1], n2)
                  + n2 * np.mod(corner[..., 2], n3)))
    tetra = corner_idx[:, tets].reshape(nq * 6, 4).astype(np.int32)

    if symmetry:
        rot_frac, rot_cart = point_group(cr)
    else:
        rot_frac = np.eye(3, dtype=np.int64)[None]
        rot_cart = np.eye(3)[None]
    # reciprocal-space action on integer addresses: k -> R^T k  (orbits are those of the
    # whole group, so R^T and (R^-1)^T give the same partition)
    ops = [R.T for R in rot_frac]
    if time_rev:
        ops = ops + [-R.T for R in rot_frac]
    rep = ii.copy()
    for Rt in ops:
        k2 = raw @ Rt.T
        # the mesh must be invariant; fractions of N would signal a bad mesh
        j = (np.mod(k2[:, 0], n1) + n1 * (np.mod(k2[:, 1], n2) + n2 * np.mod(k2[:, 2], n3)))
        rep = np.minimum(rep, j)
    # iterate to the orbit minimum (closure)
    while True:
        new = rep[rep]
        if np.array_equal(new, rep):
            break
        rep = new
    irr_idx = np.unique(rep)
    pos = np.searchsorted(irr_idx, rep).astype(np.int32)
    mult = np.bincount(pos, minlength=irr_idx.size).astype(np.int32)
    return QMesh(mesh, q_int, q_cart, tetra, irr_idx.astype(np.int32),
                 q_int[irr_idx].copy(), pos, mult, rot_frac, rot_cart, time_rev)


def pair_partner(qm: QMesh, i0_grid: int) -> np.ndarray:
    """q2 index for every q1 with q0 + q1 + q2 = G (QuasiMomentumConsrv3rd_m.f90:86-106).

    The reference picks G from {0,+-N}^3 so that q2 lands inside the folded zone; the grid
    index of q2 does not depend on that choice."""
    k0 = qm.q_int[i0_grid].astype(np.int64)
    return qm.index(-k0[None, :] - qm.q_int.astype(np.int64))


def pair_list(qm: QMesh, i0_grid: int):
    """Pair-permutation reduced list: (q1, q2) with q1 <= q2, weight 2 if q1 < q2 else 1.

    ScatterProbAllq1 fills W for q1 and its transpose for q2 from one evaluation
    (ThirdOrder_m.f90:134,150-162); the Gaussian summand is symmetric under the swap, so
    one evaluation with weight 2 is the same sum.  Works for even meshes, where the
    reference's Nq/2+1 pair-count assertion (QuasiMomentumConsrv3rd_m.f90:148-153) fails."""
    q2 = pair_partner(qm, i0_grid)
    q1 = np.arange(qm.nq)
    keep = q1 <= q2
    w = np.where(q1[keep] == q2[keep], 1.0, 2.0)
    return q1[keep].astype(np.int32), q2[keep].astype(np.int32), w
