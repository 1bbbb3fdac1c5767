"""Oracle of the four-phonon tetrahedron linewidth (TetrahedronIntegration4th,
Tetrahedron4th_sep.f90:37-425), pinned without the reference's code:

* orc_w4_rows: the Gaussian four-phonon sum rebuilt from the exported W4 equals orc_lw4_gauss
  (pinned on its own in test_oracle.py), which fixes the W4 layout and q3 = -(q0+q1+q2).
* orc_lw4_tetra_from_w: the linear tetrahedron integral of a function linear on each
  tetrahedron is exact, so over the energy Mq it has the moments of a linear simplex,
      int dE  I(E)   = sum_t  mean_v F_v / 6
      int dE E I(E)  = sum_t (sum_v M_v F_v + sum_v M_v sum_v F_v) / 20 / 6
  (6 = tetrahedra per grid cell).  With three bands far apart the three processes occupy
  disjoint energy ranges, so each is checked on its own, with its weight (1/6, 1/2, 1/2), its
  M and its F; random W and n per grid point make the check see which q2 and which
  q3 = -(q0+q1+q2) every vertex reads.  The energy integrals are exact (3-point Gauss-Legendre
  between consecutive vertex energies: the integrand is a cubic there).
"""
import numpy as np
import pytest

from oracle import oracle as O
from tests.conftest import problem


def _q3(qm, i0, q1, q2):
    return qm.index(-(qm.q_int[i0] + qm.q_int[q1] + qm.q_int[q2]).astype(np.int64)[None, :])[0]


def test_w4_rows_rebuild_the_gaussian_sum(rocksalt):
    qm, ph = problem(rocksalt, (3, 3, 3))
    inp = O.Inputs(rocksalt, qm, ph, which=4)
    sigma = 2.0
    for i0 in (int(qm.irr_idx[1]), 0):
        W = O.w4_rows(inp, i0, 0, qm.nq)
        w, n = ph.omega, ph.nbe
        lw = np.zeros(inp.nd)
        for q1 in range(qm.nq):
            for q2 in range(qm.nq):
                q3 = _q3(qm, i0, q1, q2)
                w1, w2, w3 = (w[q1][:, None, None], w[q2][None, :, None], w[q3][None, None, :])
                n1, n2, n3 = (n[q1][:, None, None], n[q2][None, :, None], n[q3][None, None, :])
                for s0 in range(3 if i0 == 0 else 0, inp.nd):
                    d = lambda x: np.exp(-0.5 * (x / sigma) ** 2) / (sigma * np.sqrt(2 * np.pi))
                    n0 = n[i0][s0]
                    t = W[q1, q2, s0] * (n1 * n2 * n3 / n0 * d(w[i0][s0] - w1 - w2 - w3) / 6
                                         + 0.5 * (1 + n1) * n2 * n3 / n0 * d(w[i0][s0] + w1 - w2 - w3)
                                         + 0.5 * (1 + n1) * (1 + n2) * n3 / n0 * d(w[i0][s0] + w1 + w2 - w3))
                    lw[s0] += np.nansum(t)
        ref = O.lw4_gauss(inp, i0, sigma)
        assert np.allclose(lw, ref, rtol=1e-11, atol=1e-14 * np.abs(ref).max())
        if i0 == 0:
            assert np.all(W[:, :, :3] == 0.0)


def _synthetic(qm, seed=3):
    """Three bands: a in [1, 2], b in [5, 6], c in [20, 21] rad/ps at every grid point; W of
    one q1 nonzero for (s1, s2, s3) = (c, b, a) only, random per (q2, s0); random n."""
    rng = np.random.default_rng(seed)
    nq, nd = qm.nq, 3
    omega = np.empty((nq, nd))
    omega[:, 0] = rng.uniform(1, 2, nq)
    omega[:, 1] = rng.uniform(5, 6, nq)
    omega[:, 2] = rng.uniform(20, 21, nq)
    nbe = rng.uniform(0.1, 2.0, (nq, nd))
    W = np.zeros((1, nq, nd, nd, nd, nd))
    W[0, :, :, 2, 1, 0] = rng.uniform(0.5, 1.5, (nq, nd))
    return omega, nbe, W


def _moments(qm, i0, q1, W, omega, nbe, n0):
    """Expected (zeroth, first) moments of each process from the simplex identities."""
    t = qm.tetra
    q3 = np.array([_q3(qm, i0, q1, q2) for q2 in range(qm.nq)])
    w1, nn1 = omega[q1, 2], nbe[q1, 2]
    w2, nn2 = omega[t, 1], nbe[t, 1]                       # (6 nq, 4)
    w3, nn3 = omega[q3[t], 0], nbe[q3[t], 0]
    out = []
    for s0 in range(3):
        Wv = W[0, t, s0, 2, 1, 0]
        procs = [(1 / 6, w1 + w2 + w3, Wv * (nn1 * nn2 * nn3 / n0[s0])),
                 (0.5, w2 + w3 - w1, Wv * ((1 + nn1) * nn2 * nn3 / n0[s0])),
                 (0.5, w3 - w1 - w2, Wv * ((1 + nn1) * (1 + nn2) * nn3 / n0[s0]))]
        row = []
        for c, M, F in procs:
            m0 = c * np.sum(F.sum(1) / 4) / 6
            m1 = c * np.sum(((M * F).sum(1) + M.sum(1) * F.sum(1)) / 20) / 6
            row.append((m0, m1, np.unique(M)))
        out.append(row)
    return out


def _integrate(f, nodes):
    """int f over [nodes[0], nodes[-1]] by 3-point Gauss-Legendre on every sub-interval."""
    x, wt = np.polynomial.legendre.leggauss(3)
    tot0 = tot1 = 0.0
    for a, b in zip(nodes[:-1], nodes[1:]):
        if b <= a:
            continue
        for xi, wi in zip(x, wt):
            e = 0.5 * (a + b) + 0.5 * (b - a) * xi
            v = f(e)
            tot0 = tot0 + 0.5 * (b - a) * wi * v
            tot1 = tot1 + 0.5 * (b - a) * wi * e * v
    return tot0, tot1


def test_lw4_tetra_moments_of_each_process():
    from spring_amd import model
    qm, _ = problem(model.rocksalt2(), (3, 3, 3))
    omega, nbe, W = _synthetic(qm)
    i0, q1 = 5, 11
    n0 = np.array([0.7, 1.3, 0.4])
    lw = lambda E: O.lw4_tetra_from_w(qm.mesh, qm.q_int, qm.tetra, i0, q1, W, omega, nbe,
                                      omega0=np.full(3, E), nbe0=n0)
    exp = _moments(qm, i0, q1, W, omega, nbe, n0)
    for p in range(3):
        Ms = exp[0][p][2]
        nodes = np.unique(np.concatenate([Ms, [Ms.min() - 0.5, Ms.max() + 0.5]]))
        # the three processes are apart: every other process is zero on this range
        others = [exp[0][k][2] for k in range(3) if k != p]
        assert all(o.max() < nodes[0] or o.min() > nodes[-1] for o in others)
        got0, got1 = _integrate(lw, nodes)
        for s0 in range(3):
            m0, m1, _ = exp[s0][p]
            assert got0[s0] == pytest.approx(m0, rel=1e-9), (p, s0)
            assert got1[s0] == pytest.approx(m1, rel=1e-9), (p, s0)


def test_lw4_tetra_degenerate_vertices_and_gamma():
    from spring_amd import model
    qm, _ = problem(model.rocksalt2(), (3, 3, 3))
    nq, nd = qm.nq, 3
    rng = np.random.default_rng(1)
    W = rng.uniform(0.5, 1.5, (1, nq, nd, nd, nd, nd))
    nbe = rng.uniform(0.1, 2.0, (nq, nd))
    # flat bands: every vertex of every tetrahedron has the same M; Mq on it is a 0/0 case
    omega = np.tile([1.0, 2.0, 4.0], (nq, 1))
    for E in (3.0, 7.0, 5.0, 6.0, 1.0):
        lw = O.lw4_tetra_from_w(qm.mesh, qm.q_int, qm.tetra, 4, 2, W, omega, nbe,
                                omega0=np.full(nd, E), nbe0=np.ones(nd))
        assert np.all(np.isfinite(lw))
    # q0 = Gamma: the three lowest s0 are skipped
    omega, nbe, W = _synthetic(qm)
    lw = O.lw4_tetra_from_w(qm.mesh, qm.q_int, qm.tetra, 0, 3, W, omega, nbe,
                            omega0=np.full(3, 27.0), nbe0=np.ones(3))
    assert np.all(lw == 0.0)
    lw = O.lw4_tetra_from_w(qm.mesh, qm.q_int, qm.tetra, 1, 3, W, omega, nbe,
                            omega0=np.full(3, 27.0), nbe0=np.ones(3))
    assert np.all(lw > 0.0)


def test_lw4_tetra_end_to_end_is_the_sum_over_q1_chunks(rocksalt):
    qm, ph = problem(rocksalt, (3, 3, 3))
    inp = O.Inputs(rocksalt, qm, ph, which=4)
    i0 = int(qm.irr_idx[2])
    a = O.lw4_tetra(inp, qm, i0, threads=4, chunk=5)
    W = O.w4_rows(inp, i0, 0, qm.nq)
    b = O.lw4_tetra_from_w(qm.mesh, qm.q_int, qm.tetra, i0, 0, W, ph.omega, ph.nbe)
    assert np.allclose(a, b, rtol=1e-12, atol=1e-15 * np.abs(b).max())
    assert np.all(b >= 0.0) and b.max() > 0.0
