"""The four-phonon process terms of the oracle, pinned by physics rather than by restating
them (processes and weights of TetrahedronIntegration4th, Tetrahedron4th_sep.f90:242-266 and
:322-343, as orc_lw4_gauss and orc_lw4_tetra_from_w apply them).

Two facts fix every term independently of how the reference writes it:

* Linearised Fermi golden rule.  With every mode except lambda0 in equilibrium, the relaxation
  rate of lambda0 from a process is the derivative of its net rate with respect to n0, i.e. a
  difference of occupation products:
      A  lambda0 -> 1 + 2 + 3         (1+n1)(1+n2)(1+n3) - n1 n2 n3
      B  lambda0 + 1 -> 2 + 3         n1 (1+n2)(1+n3) - (1+n1) n2 n3
      C  lambda0 + 1 + 2 -> 3         n1 n2 (1+n3) - (1+n1)(1+n2) n3
  These equal the oracle's compact factors n1n2n3/n0, (1+n1)n2n3/n0, (1+n1)(1+n2)n3/n0 only
  on the energy shell with Bose-Einstein occupations, so a misplaced (1 + n) changes the sum.
* Counting.  Each distinct physical process counts once, identical bosons with the factor
  1/m! (m equal modes): a sum over multisets {1,2,3} (A), over 1 and multisets {2,3} (B), over
  multisets {1,2} and 3 (C).  The reference sums over ordered (q1, q2) and all branches and
  weights the three processes 1/6, 1/2, 1/2; swapping 1/6 and 1/2 breaks the equality.

Gaussian: flat integer bands make every branch quartet exactly on shell or at least one unit
away, and sigma = 0.01 turns the Gaussian into exactly c0 on shell and exactly 0 off shell.
Tetrahedron: bands flat to 1e-7 (so every vertex is on shell to that order); the integral of
the linewidth over an energy window is the grid sum of the process terms whose energy lies in
the window (the simplex moment identity), compared with the same multiset sum.
"""
import numpy as np
import pytest

from oracle import oracle as O
from tests.conftest import problem


def _q3_table(qm, i0):
    """q3 = -(q0 + q1 + q2) for every (q1, q2) (the reference's Allq1q2q3(3, :, :))."""
    k = qm.q_int.astype(np.int64)
    s = -(k[i0][None, None, :] + k[:, None, :] + k[None, :, :])
    return qm.index(s.reshape(-1, 3)).reshape(qm.nq, qm.nq)


def _bose(w, beta):
    return 1.0 / np.expm1(beta * w)


def _physical_terms(W0, q3, w, n, nd, onshell):
    """Sum over distinct physical processes for one s0.

    W0[q1, q2, s1, s2, s3] (q3 implied), w[q, s], n[q, s] per grid point, onshell(p, M, w0)
    the energy condition of process p on the array M.  Returns the A, B, C sums of
    W x (textbook occupation difference) x (1/m! for identical modes)."""
    nq = W0.shape[0]
    q1 = np.arange(nq)[:, None, None, None, None]
    q2 = np.arange(nq)[None, :, None, None, None]
    qq3 = q3[:, :, None, None, None]
    s1 = np.arange(nd)[None, None, :, None, None]
    s2 = np.arange(nd)[None, None, None, :, None]
    s3 = np.arange(nd)[None, None, None, None, :]
    shape = W0.shape
    l1 = np.broadcast_to(q1 * nd + s1, shape)
    l2 = np.broadcast_to(q2 * nd + s2, shape)
    l3 = np.broadcast_to(qq3 * nd + s3, shape)
    w1, w2, w3 = (np.broadcast_to(w[q, s], shape) for q, s in ((q1, s1), (q2, s2), (qq3, s3)))
    n1, n2, n3 = (np.broadcast_to(n[q, s], shape) for q, s in ((q1, s1), (q2, s2), (qq3, s3)))

    def pair_sym(a, b):
        return np.where(a == b, 0.5, 1.0)

    # A: multisets {1, 2, 3}, canonical l1 <= l2 <= l3
    canon = (l1 <= l2) & (l2 <= l3)
    symA = np.where((l1 == l2) & (l2 == l3), 1.0 / 6.0, np.where((l1 == l2) | (l2 == l3), 0.5, 1.0))
    dA = (1 + n1) * (1 + n2) * (1 + n3) - n1 * n2 * n3
    A = np.sum(np.where(canon & onshell(0, w1 + w2 + w3), W0 * dA * symA, 0.0))
    # B: 1 free, multisets {2, 3}
    dB = n1 * (1 + n2) * (1 + n3) - (1 + n1) * n2 * n3
    B = np.sum(np.where((l2 <= l3) & onshell(1, w2 + w3 - w1), W0 * dB * pair_sym(l2, l3), 0.0))
    # C: multisets {1, 2}, 3 free
    dC = n1 * n2 * (1 + n3) - (1 + n1) * (1 + n2) * n3
    C = np.sum(np.where((l1 <= l2) & onshell(2, w3 - w1 - w2), W0 * dC * pair_sym(l1, l2), 0.0))
    return A, B, C


def _flat_inputs(rocksalt, qm, ph, bands, beta):
    inp = O.Inputs(rocksalt, qm, ph, which=4)
    nq, nd = qm.nq, inp.nd
    w = np.tile(np.asarray(bands, float), (nq, 1))
    n = _bose(w, beta)
    inp.omega, inp.p_omega = O._p(w, np.float64)
    inp.inv_omega, inp.p_inv_omega = O._p(1.0 / w, np.float64)
    inp.nbe, inp.p_nbe = O._p(n, np.float64)
    return inp, w, n


@pytest.mark.parametrize("bands", [(1, 2, 3, 5, 8, 13), (2, 3, 5, 7, 11, 13), (3, 4, 5, 6, 7, 30)])
def test_lw4_gauss_equals_the_sum_over_physical_processes(rocksalt, bands):
    qm, ph = problem(rocksalt, (3, 3, 3))
    beta = 0.37
    inp, w, n = _flat_inputs(rocksalt, qm, ph, bands, beta)
    nd, sigma = inp.nd, 0.01
    c0 = 1.0 / (sigma * np.sqrt(2.0 * np.pi))
    seen = np.zeros(3)
    for i0 in (int(qm.irr_idx[1]), int(qm.irr_idx[3])):
        lw = O.lw4_gauss(inp, i0, sigma)
        W = O.w4_rows(inp, i0, 0, qm.nq)
        q3 = _q3_table(qm, i0)
        exp = np.zeros(nd)
        swapped = np.zeros(nd)
        for s0 in range(nd):
            w0 = w[i0, s0]
            A, B, C = _physical_terms(W[:, :, s0], q3, w, n, nd, lambda p, M: M == w0)
            seen += np.abs([A, B, C]) > 0
            exp[s0] = c0 * (A + B + C)
            # the reference's weights with 1/6 and 1/2 swapped: A x 3, B and C x 1/3
            swapped[s0] = c0 * (3 * A + (B + C) / 3)
        assert np.all(exp > 0)
        assert lw == pytest.approx(exp, rel=1e-12)
        assert np.max(np.abs(swapped - lw) / lw) > 0.1
    # the band sets between them give every process on-shell terms (the last has B only)
    assert seen[1] > 0
    if bands != (3, 4, 5, 6, 7, 30):
        assert seen[0] > 0 and seen[2] > 0


def test_on_shell_identities_of_the_compact_factors():
    """The compact factors equal the occupation differences on the shell and only there: a
    (1 + n) moved to another mode does not."""
    rng = np.random.default_rng(5)
    beta = 0.8
    for _ in range(200):
        w1, w2 = rng.uniform(0.1, 5.0, 2)
        # A: w0 = w1 + w2 + w3;  B: w0 + w1 = w2 + w3;  C: w0 + w1 + w2 = w3
        w3 = rng.uniform(0.1, 5.0)
        cases = {
            "A": (w1 + w2 + w3, w1, w2, w3),
            "B": (w2 + w3 - w1, w1, w2, w3),
            "C": (w3 - w1 - w2, w1, w2, w3),
        }
        for p, (w0, a, b, c) in cases.items():
            if w0 <= 0.05:
                continue
            n0, n1, n2, n3 = (_bose(x, beta) for x in (w0, a, b, c))
            compact = {"A": n1 * n2 * n3 / n0, "B": (1 + n1) * n2 * n3 / n0,
                       "C": (1 + n1) * (1 + n2) * n3 / n0}[p]
            diff = {"A": (1 + n1) * (1 + n2) * (1 + n3) - n1 * n2 * n3,
                    "B": n1 * (1 + n2) * (1 + n3) - (1 + n1) * n2 * n3,
                    "C": n1 * n2 * (1 + n3) - (1 + n1) * (1 + n2) * n3}[p]
            assert compact == pytest.approx(diff, rel=1e-10)
            moved = {"A": (1 + n1) * n2 * n3 / n0, "B": n1 * (1 + n2) * n3 / n0,
                     "C": (1 + n1) * n2 * (1 + n3) / n0}[p]
            assert abs(moved - diff) > 1e-3 * abs(diff)


def _integrate(f, nodes):
    """int f over [nodes[0], nodes[-1]] by 3-point Gauss-Legendre on every sub-interval (exact
    for the piecewise cubic tetrahedron linewidth between consecutive vertex energies)."""
    x, wt = np.polynomial.legendre.leggauss(3)
    tot = 0.0
    for a, b in zip(nodes[:-1], nodes[1:]):
        if b <= a:
            continue
        for xi, wi in zip(x, wt):
            tot = tot + 0.5 * (b - a) * wi * f(0.5 * (a + b) + 0.5 * (b - a) * xi)
    return tot


def test_lw4_tetra_window_integral_equals_the_sum_over_physical_processes():
    from spring_amd import model
    qm, _ = problem(model.rocksalt2(), (2, 2, 2))
    nq, nd = qm.nq, 3
    rng = np.random.default_rng(17)
    bands, eps, beta = np.array([1.0, 3.0, 9.0]), 1e-7, 0.2
    w = bands[None, :] + eps * rng.uniform(0.0, 1.0, (nq, nd))
    n = _bose(w, beta)
    i0 = 5
    q3 = _q3_table(qm, i0)
    # a W symmetric under permutations of (1, 2, 3) (the symmetry of |V4|^2), per s0
    u = rng.uniform(0.5, 1.5, (nq, nd))
    v = rng.uniform(0.5, 1.5, (nq, nd))
    c = rng.uniform(0.5, 1.5, nd)
    W = np.zeros((nq, nq, nd, nd, nd, nd))
    for q1 in range(nq):
        for q2 in range(nq):
            a, b, d = u[q1][:, None, None], u[q2][None, :, None], u[q3[q1, q2]][None, None, :]
            e, f, g = v[q1][:, None, None], v[q2][None, :, None], v[q3[q1, q2]][None, None, :]
            sym = a * b * d + (e + f + g) ** 2
            W[q1, q2] = c[:, None, None, None] * sym[None]
    # every process energy of every vertex, to place the integration nodes
    M_all = []
    for q1 in range(nq):
        for q2 in range(nq):
            a = w[q1][:, None, None]
            b = w[q2][None, :, None]
            d = w[q3[q1, q2]][None, None, :]
            M_all += [(a + b + d).ravel(), (b + d - a).ravel(), (d - a - b).ravel()]
    M_all = np.unique(np.concatenate(M_all))
    for E in (5.0, 7.0, 11.0, 13.0, 15.0):
        lo, hi = E - 1e-5, E + 1e-5
        nodes = np.concatenate([[lo], M_all[(M_all > lo) & (M_all < hi)], [hi]])
        n0 = np.full(nd, _bose(E, beta))
        lw = lambda x: O.lw4_tetra_from_w(qm.mesh, qm.q_int, qm.tetra, i0, 0, W, w, n,
                                          omega0=np.full(nd, x), nbe0=n0)
        got = _integrate(lw, nodes)
        inwin = lambda p, M: (M > lo) & (M < hi)
        exp = np.zeros(nd)
        swapped = np.zeros(nd)
        for s0 in range(nd):
            A, B, C = _physical_terms(W[:, :, s0], q3, w, n, nd, inwin)
            exp[s0] = A + B + C
            swapped[s0] = 3 * A + (B + C) / 3
        assert np.all(exp > 0), E
        assert got == pytest.approx(exp, rel=2e-6), E
        assert np.max(np.abs(swapped - got) / got) > 0.1, E
