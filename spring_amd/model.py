"""Seeded synthetic crystals and interatomic force constants (IFCs).

The scattering path consumes a crystal (`Crystal`, the fields of the reference's
`cell` type, src/Kappa/Src/unit_mod.f90:9-38) and FC2/FC3/FC4 tensors in the layout
the reference reads from its HDF5 files (`FC3type`, src/Kappa/Src/FC3_mod.f90:30-99;
`FC4type`, FC4_mod.f90:30-97):

    atm_num[mu]                      number of neighbour sites of basis atom mu (self first)
    atm_indx[mu, N, :] = (c1,c2,c3,nu)   lattice cell of site N and its basis atom (0-based nu)
    fc3[mu, N1, N2, a, b, g]         Fortran fC(g,b,a,N2,N1,mu): identical bytes
    fc4[mu, N1, N2, N3, a, b, g, d]  Fortran fC(d,g,b,a,N3,N2,N1,mu)
    fc2[mu, N, a, b]                 Fortran fC(1,b,a,N,mu)

Arrays are C-ordered with the Fortran index order reversed, so their bytes are the
reference's column-major arrays.  Basis indices are 0-based here (1-based in Fortran).

Model (what the benchmark configurations specify):
  * FC2: longitudinal / transverse springs per neighbour shell; each shell constant gets
    one seeded +-5 % factor (shell-wise, so the point group of the lattice is kept);
    the acoustic sum rule (ASR) fixes the self term.
  * FC3 / FC4: seeded N(0, s) entries on every cluster whose pairwise distances all lie
    within the cutoff shell, then the projection onto {index-permutation symmetric}
    intersected with {ASR: sum over the last site = 0}, by alternating orthogonal
    projections until the residual is below 1e-13 of the tensor norm.
"""
from __future__ import annotations

import itertools
from dataclasses import dataclass, field

import numpy as np


@dataclass
class Crystal:
    latvec: np.ndarray          # (3,3): latvec[:, i] is lattice vector i (Angstrom)
    basis_frac: np.ndarray      # (nb, 3) fractional positions
    mass: np.ndarray            # (nb,) amu
    species: np.ndarray         # (nb,) int species id (for symmetry)
    name: str = "crystal"

    @property
    def nb(self) -> int:
        return self.mass.shape[0]

    @property
    def ndof(self) -> int:
        return 3 * self.nb

    @property
    def vol(self) -> float:
        a = self.latvec
        return abs(float(np.dot(a[:, 0], np.cross(a[:, 1], a[:, 2]))))

    @property
    def G(self) -> np.ndarray:
        """Reciprocal vectors as columns, G^T A = 2 pi I (unit_mod.f90:199-207)."""
        return 2.0 * np.pi * np.linalg.inv(self.latvec).T

    @property
    def tau(self) -> np.ndarray:
        """Cartesian basis positions, (nb, 3)."""
        return self.basis_frac @ self.latvec.T

    @property
    def one_by_sqrt_mass(self) -> np.ndarray:
        return 1.0 / np.sqrt(self.mass)


@dataclass
class IFC:
    order: int
    atm_num: np.ndarray         # (nb,) int32
    atm_indx: np.ndarray        # (nb, natm_max, 4) int32: cell (3) + basis (0-based)
    fc: np.ndarray              # order-dependent layout, see module docstring
    meta: dict = field(default_factory=dict)

    @property
    def natm_max(self) -> int:
        return self.atm_indx.shape[1]


# ----------------------------------------------------------------------------------------
# crystals


def rocksalt_primitive(a: float = 5.64, masses=(22.99, 35.45), name="rocksalt-2") -> Crystal:
    """fcc primitive cell, cation at 0, anion at (1/2,1/2,1/2) (= a/2 (1,1,1) ~ a/2 x)."""
    lat = 0.5 * a * np.array([[0.0, 1.0, 1.0], [1.0, 0.0, 1.0], [1.0, 1.0, 0.0]]).T
    frac = np.array([[0.0, 0.0, 0.0], [0.5, 0.5, 0.5]])
    return Crystal(lat, frac, np.array(masses, float), np.array([0, 1]), name)


def rocksalt_supercell(n: int, a: float = 4.21, masses=(24.3, 16.0), jitter: float = 0.02,
                       seed: int = 3) -> Crystal:
    """2n-atom cell: n x 1 x 1 supercell of the rocksalt primitive cell (n cations, n anions),
    every atom displaced by a seeded uniform jitter of at most `jitter` Angstrom per axis."""
    prim = rocksalt_primitive(a, masses)
    lat = prim.latvec.copy()
    lat[:, 0] *= float(n)
    rng = np.random.default_rng(seed)
    cart = []
    spec = []
    for i in range(n):
        for b in range(2):
            r = prim.latvec @ (prim.basis_frac[b] + np.array([i, 0, 0]))
            cart.append(r)
            spec.append(b)
    cart = np.array(cart) + rng.uniform(-jitter, jitter, size=(2 * n, 3))
    frac = cart @ np.linalg.inv(lat).T
    mass = np.array([masses[s] for s in spec], float)
    return Crystal(lat, frac, mass, np.array(spec), f"rocksalt-{2 * n}")


def rocksalt_supercell10(a: float = 4.21, masses=(24.3, 16.0), jitter: float = 0.02,
                         seed: int = 3) -> Crystal:
    """10-atom cell: 5x1x1 supercell of the rocksalt primitive cell (5 cations, 5 anions)."""
    return rocksalt_supercell(5, a, masses, jitter, seed)


# ----------------------------------------------------------------------------------------
# neighbour lists


def _neighbours(cr: Crystal, rcut: float, tol: float = 1e-3):
    """For every basis atom mu: sites (cell, nu) within rcut, self first, then by distance."""
    nb = cr.nb
    nmax = int(np.ceil(rcut / min(np.linalg.norm(cr.latvec, axis=0)))) + 2
    rng = range(-nmax, nmax + 1)
    cells = np.array(list(itertools.product(rng, rng, rng)), dtype=np.int64)
    tau = cr.tau
    out = []
    for mu in range(nb):
        lst = []
        for nu in range(nb):
            r = cells @ cr.latvec.T + tau[nu] - tau[mu]
            d = np.linalg.norm(r, axis=1)
            for c, dd in zip(cells[d <= rcut + tol], d[d <= rcut + tol]):
                lst.append((round(float(dd), 6), tuple(int(x) for x in c), nu))
        lst.sort(key=lambda t: (t[0], t[2], t[1]))
        out.append([(c, nu) for _, c, nu in lst])
    return out


def _site_pos(cr: Crystal, c, nu):
    return cr.latvec @ np.asarray(c, float) + cr.tau[nu]


def _pack_sites(nbrs):
    nb = len(nbrs)
    natm = max(len(l) for l in nbrs)
    atm_num = np.array([len(l) for l in nbrs], np.int32)
    atm_indx = np.zeros((nb, natm, 4), np.int32)
    for mu, lst in enumerate(nbrs):
        for n, (c, nu) in enumerate(lst):
            atm_indx[mu, n, :3] = c
            atm_indx[mu, n, 3] = nu
    return atm_num, atm_indx


# ----------------------------------------------------------------------------------------
# FC2 spring model


def spring_fc2(cr: Crystal, shells, seed: int = 1, perturb: float = 0.05,
               nominal: Crystal | None = None) -> IFC:
    """shells = [(r_shell, k_L, k_T), ...] in Angstrom and eV/A^2.

    A pair (mu, site) belongs to the shell whose radius matches its distance in `nominal`
    (the undistorted crystal; defaults to `cr`), and its spring acts along the actual bond.
    """
    nominal = nominal if nominal is not None else cr
    rng = np.random.default_rng(seed)
    fac = 1.0 + perturb * rng.uniform(-1.0, 1.0, size=(len(shells), 2))
    radii = np.array([s[0] for s in shells])
    rcut = radii.max()
    nbrs = _neighbours(nominal, rcut)
    atm_num, atm_indx = _pack_sites(nbrs)
    fc = np.zeros((cr.nb, atm_indx.shape[1], 3, 3))
    for mu, lst in enumerate(nbrs):
        self_term = np.zeros((3, 3))
        for n, (c, nu) in enumerate(lst):
            if n == 0:
                continue
            dn = np.linalg.norm(_site_pos(nominal, c, nu) - _site_pos(nominal, (0, 0, 0), mu))
            sh = int(np.argmin(np.abs(radii - dn)))
            kl = shells[sh][1] * fac[sh, 0]
            kt = shells[sh][2] * fac[sh, 1]
            r = _site_pos(cr, c, nu) - _site_pos(cr, (0, 0, 0), mu)
            u = r / np.linalg.norm(r)
            k = kl * np.outer(u, u) + kt * (np.eye(3) - np.outer(u, u))
            fc[mu, n] = -k            # fc[mu, N, a, b] = Phi(mu a ; N b)
            self_term += k
        fc[mu, 0] = self_term
    return IFC(2, atm_num, atm_indx, fc, {"shells": shells, "factors": fac})


# ----------------------------------------------------------------------------------------
# anharmonic IFCs: random + permutation symmetry + ASR


def _clusters(cr: Crystal, nbrs, order: int, rcut: float, tol: float = 1e-3):
    """All (mu, N1..N_{order-1}) whose sites are pairwise within rcut."""
    out = []
    for mu, lst in enumerate(nbrs):
        pos = [_site_pos(cr, c, nu) for c, nu in lst]
        m = len(lst)
        for combo in itertools.product(range(m), repeat=order - 1):
            ok = True
            for i in range(order - 1):
                for j in range(i + 1, order - 1):
                    if np.linalg.norm(pos[combo[i]] - pos[combo[j]]) > rcut + tol:
                        ok = False
                        break
                if not ok:
                    break
            if ok:
                out.append((mu,) + combo)
    return out


def random_anharmonic(cr: Crystal, order: int, rcut: float, sigma: float, seed: int,
                      max_iter: int = 2000, nominal: Crystal | None = None) -> IFC:
    """Seeded N(0, sigma) IFC of `order` (3 or 4) on pairwise-`rcut` clusters, projected
    onto permutation symmetry + ASR.  Cluster membership uses `nominal` (undistorted)."""
    nominal = nominal if nominal is not None else cr
    nbrs = _neighbours(nominal, rcut)
    atm_num, atm_indx = _pack_sites(nbrs)
    clus = _clusters(nominal, nbrs, order, rcut)
    nc = len(clus)
    # site of each cluster member as (cell, nu) tuples; the first member is (0, mu)
    key = {}
    sites = []
    for t, cl in enumerate(clus):
        mu = cl[0]
        s = [((0, 0, 0), mu)] + [nbrs[mu][n] for n in cl[1:]]
        sites.append(s)
        key[tuple(s)] = t
    # permutation maps: cluster t under permutation p is cluster perm_idx[p][t]
    perms = list(itertools.permutations(range(order)))
    perm_idx = np.zeros((len(perms), nc), np.int64)
    for pi, p in enumerate(perms):
        for t, s in enumerate(sites):
            ps = [s[i] for i in p]
            c0 = ps[0][0]
            shifted = tuple((tuple(c[k] - c0[k] for k in range(3)), nu) for c, nu in ps)
            perm_idx[pi, t] = key[shifted]
    # ASR groups: clusters sharing all but the last member
    grp_key = {}
    grp = np.zeros(nc, np.int64)
    for t, s in enumerate(sites):
        g = tuple(s[:-1])
        grp[t] = grp_key.setdefault(g, len(grp_key))
    ngrp = len(grp_key)
    cnt = np.bincount(grp, minlength=ngrp).astype(float)

    rng = np.random.default_rng(seed)
    shape = (nc,) + (3,) * order
    phi = rng.normal(0.0, sigma, size=shape)
    norm0 = np.linalg.norm(phi)

    def sym(x):
        acc = np.zeros_like(x)
        for pi, p in enumerate(perms):
            # value at cluster t with component order (c_0..c_{o-1}) equals the permuted
            # cluster's tensor with axes permuted accordingly
            inv = np.argsort(p)
            acc += np.transpose(x[perm_idx[pi]], (0,) + tuple(1 + i for i in inv))
        return acc / len(perms)

    def asr(x):
        s = np.zeros((ngrp,) + (3,) * order)
        np.add.at(s, grp, x)
        return x - (s / cnt.reshape((-1,) + (1,) * order))[grp]

    it = 0
    for it in range(max_iter):
        phi = asr(sym(phi))
        r1 = np.linalg.norm(sym(phi) - phi)
        if r1 < 1e-13 * norm0:
            break
    # pack into the reference layout
    natm = atm_indx.shape[1]
    if order == 3:
        fc = np.zeros((cr.nb, natm, natm, 3, 3, 3))
        for t, cl in enumerate(clus):
            fc[cl[0], cl[1], cl[2]] = phi[t]
    else:
        fc = np.zeros((cr.nb, natm, natm, natm, 3, 3, 3, 3))
        for t, cl in enumerate(clus):
            fc[cl[0], cl[1], cl[2], cl[3]] = phi[t]
    return IFC(order, atm_num, atm_indx, fc,
               {"clusters": nc, "iterations": it + 1, "rcut": rcut, "sigma": sigma,
                "seed": seed})


# ----------------------------------------------------------------------------------------
# the benchmark models


@dataclass
class Model:
    crystal: Crystal
    fc2: IFC
    fc3: IFC | None
    fc4: IFC | None
    name: str


def rocksalt2(with_fc4: bool = False, seed: int = 7) -> Model:
    """2-atom rocksalt (a=5.64 A, 22.99/35.45 amu): FC2 springs 1.5/0.2 (1st shell) and
    0.3/0.05 eV/A^2 (2nd shell), FC3 N(0,1) eV/A^3 within the 2nd shell, and optionally
    FC4 N(0,2) eV/A^4 within the 1st shell."""
    cr = rocksalt_primitive()
    a = 5.64
    r1, r2 = a / 2.0, a / np.sqrt(2.0)
    fc2 = spring_fc2(cr, [(r1, 1.5, 0.2), (r2, 0.3, 0.05)], seed=seed)
    fc3 = random_anharmonic(cr, 3, r2, 1.0, seed + 1)
    fc4 = random_anharmonic(cr, 4, r1, 2.0, seed + 2) if with_fc4 else None
    return Model(cr, fc2, fc3, fc4, "rocksalt2")


def rocksalt10(seed: int = 11) -> Model:
    """10-atom cell (5 x 24.3 + 5 x 16.0 amu, jittered 5x1x1 rocksalt supercell,
    a = 4.21 A): same spring model, FC3 N(0,0.5) eV/A^3 within the 3rd shell."""
    a = 4.21
    cr = rocksalt_supercell10(a=a, seed=seed)
    nominal = rocksalt_supercell10(a=a, jitter=0.0)
    r1, r2, r3 = a / 2.0, a / np.sqrt(2.0), a * np.sqrt(3.0) / 2.0
    fc2 = spring_fc2(cr, [(r1, 1.5, 0.2), (r2, 0.3, 0.05)], seed=seed, nominal=nominal)
    fc3 = random_anharmonic(cr, 3, r3, 0.5, seed + 1, nominal=nominal)
    return Model(cr, fc2, fc3, None, "rocksalt10")


def fcc1(with_fc4: bool = True, seed: int = 31) -> Model:
    """Monatomic fcc (a = 3.61 A, 63.55 amu, three branches): nearest-neighbour springs 1.2/0.1
    eV/A^2, FC3 N(0,1) eV/A^3 and FC4 N(0,2) eV/A^4 within the nearest-neighbour shell."""
    a = 3.61
    lat = 0.5 * a * np.array([[0.0, 1.0, 1.0], [1.0, 0.0, 1.0], [1.0, 1.0, 0.0]]).T
    cr = Crystal(lat, np.zeros((1, 3)), np.array([63.55]), np.array([0]), "fcc-1")
    r1 = a / np.sqrt(2.0)
    fc2 = spring_fc2(cr, [(r1, 1.2, 0.1)], seed=seed)
    fc3 = random_anharmonic(cr, 3, r1, 1.0, seed + 1)
    fc4 = random_anharmonic(cr, 4, r1, 2.0, seed + 2) if with_fc4 else None
    return Model(cr, fc2, fc3, fc4, "fcc1")


def rocksalt_chain(n: int, seed: int = 21) -> Model:
    """2n-atom jittered n x 1 x 1 rocksalt supercell (a = 4.21 A, 24.3/16.0 amu): the
    rocksalt10 spring model with FC3 N(0,0.5) eV/A^3 within the 2nd shell (a smaller cluster
    set, for tests of the intermediate cell sizes)."""
    a = 4.21
    cr = rocksalt_supercell(n, a=a, seed=seed)
    nominal = rocksalt_supercell(n, a=a, jitter=0.0)
    r1, r2 = a / 2.0, a / np.sqrt(2.0)
    fc2 = spring_fc2(cr, [(r1, 1.5, 0.2), (r2, 0.3, 0.05)], seed=seed, nominal=nominal)
    fc3 = random_anharmonic(cr, 3, r2, 0.5, seed + 1, nominal=nominal)
    return Model(cr, fc2, fc3, None, f"rocksalt{2 * n}")
