"""Pinning of the FP64 CPU oracle (runs without a GPU).

None of the reference's tests exercise kappa.x or LineWidth.x, so each oracle function is
pinned by closed forms, library routines, brute force and invariants chosen so that a
dropped term, wrong sign, wrong index or transposed operand fails at least one check.
"""
import numpy as np
import pytest
import scipy.constants as sc
from scipy.stats import norm

from spring_amd import constants as C, mesh
from tests.conftest import problem
from oracle import oracle as O


# ---------------------------------------------------------------- constants (SI derivations)
def test_constants_from_si():
    hbar, e, amu, kb = sc.hbar, sc.e, sc.atomic_mass, sc.k
    # sqrt(eV / (A^2 amu)) in rad/ps
    assert C.W_THZ == pytest.approx(np.sqrt(e / (1e-20 * amu)) * 1e-12, rel=2e-5)
    # hbar / k_B in K ps
    assert C.BE_CONV / C.W_THZ == pytest.approx(hbar / kb * 1e12, rel=2e-5)
    # W = pi/16 hbar |V|^2/(w w w): V in eV/A^3/amu^1.5, w in rad/ps, W in 1/ps^2
    assert C.SCTR3_THZ2 == pytest.approx(np.pi / 16 * hbar * e ** 2 / amu ** 3, rel=1e-4)
    assert C.SCTR4_THZ2 == pytest.approx(np.pi / 8 * hbar ** 2 * e ** 2 / amu ** 4 * 1e8, rel=1e-4)
    # kappa prefactor: hbar^2/k_B with w in 1/ps, tau in ps, V in A^3 -> W/mK
    assert C.KW_M_K == pytest.approx(hbar ** 2 / kb * 1e42, rel=1e-6)


# ---------------------------------------------------------------- V3
@pytest.fixture(scope="module")
def tiny(rocksalt):
    qm, ph = problem(rocksalt, (3, 3, 3))
    return rocksalt, qm, ph, O.Inputs(rocksalt, qm, ph)


def _pair(qm, i0, q1):
    q2 = int(mesh.pair_partner(qm, i0)[q1])
    return q2


def test_v3_literal_vs_factorised(tiny):
    m, qm, ph, inp = tiny
    rng = np.random.default_rng(0)
    for i0, q1 in [(1, 5), (4, 0), (0, 13), (7, 7)]:
        q2 = _pair(qm, i0, q1)
        chk = [int(i0 == 0), int(q1 == 0), int(q2 == 0)]
        # random (non-unitary) complex "eigenvectors" exercise every index path
        e = [rng.normal(size=(6, 6)) + 1j * rng.normal(size=(6, 6)) for _ in range(3)]
        a = O.scatter_mat_el(inp, qm.q_cart[q1], qm.q_cart[q2], chk, *e, literal=True)
        b = O.scatter_mat_el(inp, qm.q_cart[q1], qm.q_cart[q2], chk, *e, literal=False)
        assert np.abs(a - b).max() < 1e-13 * np.abs(a).max()


def test_v3_single_term_closed_form(rocksalt):
    """One nonzero FC3 element: V(s0,s1,s2) = f p1 p2 m e0 e1 e2 (no conjugation)."""
    import copy
    m = copy.deepcopy(rocksalt)
    fc = np.zeros_like(m.fc3.fc)
    mu, n1, n2, a, b, g = 1, 3, 7, 2, 0, 1
    fc[mu, n1, n2, a, b, g] = 1.7
    m.fc3.fc = fc
    qm, ph = problem(m, (3, 3, 3))
    inp = O.Inputs(m, qm, ph)
    rng = np.random.default_rng(1)
    e = [rng.normal(size=(6, 6)) + 1j * rng.normal(size=(6, 6)) for _ in range(3)]
    q1, q2 = qm.q_cart[5], qm.q_cart[11]
    V = O.scatter_mat_el(inp, q1, q2, [0, 0, 0], *e)
    cr = m.crystal
    nu, eta = m.fc3.atm_indx[mu, n1, 3], m.fc3.atm_indx[mu, n2, 3]
    R1 = cr.latvec @ m.fc3.atm_indx[mu, n1, :3]
    R2 = cr.latvec @ m.fc3.atm_indx[mu, n2, :3]
    pref = 1.7 * np.exp(1j * (q1 @ R1)) * np.exp(1j * (q2 @ R2)) / np.sqrt(
        cr.mass[mu] * cr.mass[nu] * cr.mass[eta])
    # e[k][s][row]
    expect = pref * np.einsum('i,j,k->ijk', e[0][:, 3 * mu + a], e[1][:, 3 * nu + b], e[2][:, 3 * eta + g])
    assert np.allclose(V, expect, rtol=1e-13, atol=1e-15)


def test_v3_gauge_and_linearity(tiny):
    m, qm, ph, inp = tiny
    i0, q1 = 2, 9
    q2 = _pair(qm, i0, q1)
    e0, e1, e2 = ph.evec[i0], ph.evec[q1].copy(), ph.evec[q2]
    V = O.scatter_mat_el(inp, qm.q_cart[q1], qm.q_cart[q2], [0, 0, 0], e0, e1, e2)
    e1g = e1 * np.exp(1j * np.linspace(0.3, 2.0, 6))[:, None]
    Vg = O.scatter_mat_el(inp, qm.q_cart[q1], qm.q_cart[q2], [0, 0, 0], e0, e1g, e2)
    assert np.allclose(np.abs(V) ** 2, np.abs(Vg) ** 2, rtol=1e-12, atol=1e-16)
    inp2 = O.Inputs(m, qm, ph)
    inp2.fc = inp.fc * 3.0
    inp2.p_fc = inp2.fc.ctypes.data_as(inp.p_fc.__class__)
    V3 = O.scatter_mat_el(inp2, qm.q_cart[q1], qm.q_cart[q2], [0, 0, 0], e0, e1, e2)
    assert np.allclose(V3, 3.0 * V, rtol=1e-13, atol=1e-16)


def test_v3_time_reversal_and_pair_swap(tiny):
    m, qm, ph, inp = tiny
    i0, q1 = 3, 10
    q2 = _pair(qm, i0, q1)
    e0, e1, e2 = ph.evec[i0], ph.evec[q1], ph.evec[q2]
    V = O.scatter_mat_el(inp, qm.q_cart[q1], qm.q_cart[q2], [0, 0, 0], e0, e1, e2)
    # time reversal: (-q, e*) gives V*
    Vt = O.scatter_mat_el(inp, -qm.q_cart[q1], -qm.q_cart[q2], [0, 0, 0], e0.conj(), e1.conj(), e2.conj())
    assert np.allclose(Vt, V.conj(), rtol=1e-12, atol=1e-14)
    # permutation-symmetric FC3: swapping (q1,e1) and (q2,e2) transposes s1 <-> s2
    Vs = O.scatter_mat_el(inp, qm.q_cart[q2], qm.q_cart[q1], [0, 0, 0], e0, e2, e1)
    assert np.allclose(Vs, np.transpose(V, (0, 2, 1)), rtol=1e-10, atol=1e-13)


def test_v3_asr_acoustic_gamma(tiny):
    """q2 = Gamma with a uniform translation for e2: sum_N2 Phi = 0 -> V = 0."""
    m, qm, ph, inp = tiny
    i0, q1 = 4, 0
    # q1 = -q0  =>  q2 = 0
    q1 = int(qm.index(-qm.q_int[i0].astype(np.int64)))
    q2 = _pair(qm, i0, q1)
    assert q2 == 0
    cr = m.crystal
    e2 = np.zeros((6, 6), complex)
    for s, u in enumerate(np.eye(3)):
        vec = np.concatenate([np.sqrt(cr.mass[k]) * u for k in range(2)])
        e2[s] = vec / np.linalg.norm(vec)
    V = O.scatter_mat_el(inp, qm.q_cart[q1], qm.q_cart[q2], [0, 0, 0], ph.evec[i0], ph.evec[q1], e2)
    Vref = O.scatter_mat_el(inp, qm.q_cart[q1], qm.q_cart[q2], [0, 0, 0], ph.evec[i0], ph.evec[q1], ph.evec[q2])
    assert np.abs(V[:, :, :3]).max() < 1e-12 * np.abs(Vref).max()


def test_scatter_prob_transpose_and_factor(tiny):
    m, qm, ph, inp = tiny
    rng = np.random.default_rng(2)
    phi = rng.normal(size=(6, 6, 6)) + 1j * rng.normal(size=(6, 6, 6))
    iw = [rng.uniform(0.1, 1.0, 6) for _ in range(3)]
    W, Wp = O.scatter_prob(6, [0, 0, 0], *iw, phi, 2.5)
    expect = np.abs(phi) ** 2 * np.einsum('i,j,k->ijk', *iw) * 2.5
    assert np.allclose(W, expect, rtol=1e-14)
    assert np.array_equal(Wp, np.transpose(W, (0, 2, 1)))
    W0, _ = O.scatter_prob(6, [1, 0, 0], *iw, phi, 2.5)
    assert np.all(W0[:3] == 0) and np.allclose(W0[3:], expect[3:])


# ---------------------------------------------------------------- Gaussian delta stage
def test_delta_stage_vs_scipy_norm():
    rng = np.random.default_rng(3)
    nd = 6
    W = rng.uniform(0, 1, (nd, nd, nd))
    w0, w1, w2 = (rng.uniform(0, 3, nd) for _ in range(3))
    n1, n2 = rng.uniform(0, 2, nd), rng.uniform(0, 2, nd)
    sigma = 0.4
    lw = O.lw3_from_w(W, w0, w1, w2, n1, n2, sigma)
    d = lambda x: norm.pdf(x, scale=sigma)
    exp = np.zeros(nd)
    for s0 in range(nd):
        for s1 in range(nd):
            for s2 in range(nd):
                exp[s0] += W[s0, s1, s2] * ((n1[s1] + n2[s2] + 1) * d(w0[s0] - w1[s1] - w2[s2])
                                            + (n1[s1] - n2[s2]) * (d(w0[s0] + w1[s1] - w2[s2])
                                                                   - d(w0[s0] - w1[s1] + w2[s2])))
    assert np.allclose(lw, exp, rtol=1e-13)


def test_lw3_pair_swap_symmetry(tiny):
    """The summand is symmetric under (q1,s1) <-> (q2,s2); the pair-permutation reduced
    sum (q1 <= q2, weight 2) equals the full sum."""
    m, qm, ph, inp = tiny
    for i0 in (1, 3):
        full, W = O.lw3_gauss(inp, int(qm.irr_idx[i0]), 0.1, want_w=True)
        q1s, q2s, wts = mesh.pair_list(qm, int(qm.irr_idx[i0]))
        red = np.zeros(6)
        ig = int(qm.irr_idx[i0])
        for q1, q2, w in zip(q1s, q2s, wts):
            red += w * O.lw3_from_w(W[q1], ph.omega[ig], ph.omega[q1], ph.omega[q2],
                                    ph.nbe[q1], ph.nbe[q2], 0.1)
        assert np.allclose(red, full, rtol=1e-12, atol=1e-300)


def test_lw3_gauss_pairs_equals_full_sum(tiny, rocksalt):
    """orc_lw3_gauss_pairs (each reduced pair evaluated once, the transposed W standing for the
    swapped pair, as ScatterProbAllq1 stores it) equals orc_lw3_gauss over every q1: at Gamma and
    off it, on an odd mesh (3^3) and an even one (4^3, q1 == q2 pairs), with a branch rule, and
    with the list split in two, to 1e-10 (measured 1.6e-12: the full sum evaluates V3 of the
    swapped ordering on its own, and the (n1 - n2) terms cancel).  Dropping the swapped ordering
    moves the sum by order one."""
    m, qm, ph, inp = tiny
    q4, p4 = problem(rocksalt, (4, 4, 4))
    i4 = O.Inputs(rocksalt, q4, p4)
    nonzero = 0
    for q, inpx in ((qm, inp), (q4, i4)):
        for i0 in (0, 1, q.nq0 - 1):
            g = int(q.irr_idx[i0])
            q1s, q2s, _ = mesh.pair_list(q, g)
            pairs = np.stack([q1s, q2s], axis=1)
            for pol in (None, (0, 5, 0, 2, 0, 5)):
                full = O.lw3_gauss(inpx, g, 0.1, pol=pol)
                red = O.lw3_gauss_pairs(inpx, g, 0.1, pairs, pol=pol)
                half = O.lw3_gauss_pairs(inpx, g, 0.1, pairs[::2], pol=pol) + \
                    O.lw3_gauss_pairs(inpx, g, 0.1, pairs[1::2], pol=pol)
                scale = np.abs(full).max()
                nonzero += scale > 0
                assert np.allclose(red, full, rtol=1e-10, atol=1e-14 * scale)
                assert np.allclose(half, full, rtol=1e-10, atol=1e-14 * scale)
    assert nonzero >= 9
    # the swapped ordering matters: counting each pair once (no transpose) gives a different sum
    g = int(q4.irr_idx[1])
    q1s, q2s, w = mesh.pair_list(q4, g)
    once = sum(O.lw3_gauss(i4, g, 0.1, int(a), int(a) + 1) for a in q1s)
    assert not np.allclose(once, O.lw3_gauss(i4, g, 0.1), rtol=1e-3)


def test_lw3_fc_scaling(tiny):
    m, qm, ph, inp = tiny
    lw = O.lw3_gauss(inp, int(qm.irr_idx[2]), 0.1)
    inp2 = O.Inputs(m, qm, ph)
    inp2.fc = inp.fc * 2.0
    inp2.p_fc = inp2.fc.ctypes.data_as(inp.p_fc.__class__)
    assert np.allclose(O.lw3_gauss(inp2, int(qm.irr_idx[2]), 0.1), 4.0 * lw, rtol=1e-12)


# ---------------------------------------------------------------- V4
def test_v4_literal_vs_factorised(rocksalt):
    qm, ph = problem(rocksalt, (2, 2, 2))
    inp = O.Inputs(rocksalt, qm, ph, which=4)
    rng = np.random.default_rng(4)
    e = [rng.normal(size=(6, 6)) + 1j * rng.normal(size=(6, 6)) for _ in range(4)]
    q = qm.q_cart
    for (a, b, c) in [(1, 2, 3), (0, 5, 7), (6, 6, 1)]:
        A = O.scatter_mat_el4(inp, q[a], q[b], q[c], [0, 0, 0, 0], *e, literal=True)
        B = O.scatter_mat_el4(inp, q[a], q[b], q[c], [0, 0, 0, 0], *e, literal=False)
        assert np.abs(A - B).max() < 1e-13 * np.abs(A).max()


def test_v4_gauge_and_permutation(rocksalt):
    qm, ph = problem(rocksalt, (3, 3, 3))
    inp = O.Inputs(rocksalt, qm, ph, which=4)
    q = qm.q_cart
    i0, q1, q2 = 4, 7, 12
    q3 = int(qm.index(-(qm.q_int[i0] + qm.q_int[q1] + qm.q_int[q2]).astype(np.int64)))
    ev = ph.evec
    V = O.scatter_mat_el4(inp, q[q1], q[q2], q[q3], [0, 0, 0, 0], ev[i0], ev[q1], ev[q2], ev[q3])
    # swapping q2 <-> q3 transposes s2 <-> s3 (permutation-symmetric FC4)
    Vs = O.scatter_mat_el4(inp, q[q1], q[q3], q[q2], [0, 0, 0, 0], ev[i0], ev[q1], ev[q3], ev[q2])
    assert np.abs(Vs - np.transpose(V, (0, 1, 3, 2))).max() < 1e-10 * np.abs(V).max()


# ---------------------------------------------------------------- kappa
def test_kappa_symm_equals_full_sum(rocksalt):
    """kappa_noIter over the full mesh == kappa_noIter_symm over the irreducible wedge when
    every star member carries the rotated velocity of its representative (built here
    explicitly); cubic point group -> isotropic diagonal tensor."""
    import copy
    qm, ph = problem(rocksalt, (6, 6, 6))
    ph = copy.deepcopy(ph)
    rng = np.random.default_rng(5)
    N = qm.mesh
    Rinvt = [np.round(np.linalg.inv(R).T).astype(int) for R in qm.rot_frac]
    for ir, rep in enumerate(qm.irr_idx):
        k = qm.q_int[rep]
        little = [r for r, M in enumerate(Rinvt) if np.all(np.mod(M @ k - k, N) == 0)]
        v = rng.normal(size=(ph.omega.shape[1], 3)) * 1000.0
        v = np.mean([v @ qm.rot_cart[r].T for r in little], axis=0)
        for q in np.nonzero(qm.pos == ir)[0]:
            r = next(r for r, M in enumerate(Rinvt) if np.all(np.mod(M @ k - qm.q_int[q], N) == 0))
            ph.grp_vel[q] = v @ qm.rot_cart[r].T
            ph.omega[q] = ph.omega[rep]
            ph.nbe[q] = ph.nbe[rep]
    lw_irr = 0.05 + 0.01 * np.abs(ph.omega[qm.irr_idx]) ** 0.5
    vol = rocksalt.crystal.vol
    kf = O.kappa_rta(qm, ph, lw_irr, 300.0, vol)
    ks = O.kappa_rta_symm(qm, ph, lw_irr, 300.0, vol)
    assert np.allclose(kf, ks, rtol=1e-12, atol=1e-12 * kf[0, 0])
    assert np.allclose(kf, kf.T, rtol=1e-12)
    assert np.allclose(np.diag(kf), kf[0, 0], rtol=1e-10)
    assert abs(kf[0, 1]) < 1e-10 * kf[0, 0]
    # tau = 0.5/lw: doubling lw halves kappa
    assert np.allclose(O.kappa_rta(qm, ph, 2 * lw_irr, 300.0, vol), kf / 2, rtol=1e-13)


def test_kappa_single_mode_closed_form():
    """One mode on a 1-point mesh: kappa_ab = kW/(T^2 V) w^2 n(n+1) (0.5/lw) v_a v_b."""
    class QM:
        q_int = np.array([[1, 0, 0]], np.int32)
        pos = np.array([0], np.int32)
        nq = 1
    class PH:
        omega = np.array([[0, 0, 0, 0, 0, 20.0]])
        nbe = np.array([[0, 0, 0, 0, 0, 0.7]])
        grp_vel = np.zeros((1, 6, 3))
    PH.grp_vel[0, 5] = [1000.0, -500.0, 250.0]
    lw = np.ones((1, 6)) * 1e30
    lw[0, 5] = 0.25
    k = O.kappa_rta(QM, PH, lw, 300.0, 40.0)
    v = PH.grp_vel[0, 5]
    expect = C.KW_M_K / (300.0 ** 2 * 40.0) * 400.0 * 0.7 * 1.7 * 2.0 * np.outer(v, v)
    assert np.allclose(k, expect, rtol=1e-13, atol=1e-300)


# ---------------------------------------------------------------- tetrahedron method
def test_tetra_weights_closed_forms():
    """Sum of vertex weights is 1; with F = M (linear), sum I F = Mqs (the iso-surface value);
    g integrates to 1 over [M_min, M_max]; a constant F passes through unchanged."""
    rng = np.random.default_rng(7)
    for _ in range(200):
        M = rng.uniform(-1, 3, 4)
        lo, hi = M.min(), M.max()
        x = rng.uniform(lo, hi)
        enter, g, I, gif = O.tetra_weights(x, M, np.full(4, 2.5))
        assert enter == 1
        assert abs(I.sum() - 1.0) < 1e-10
        assert abs(gif - g * 2.5 / 6.0) < 1e-10 * max(1.0, abs(g))
        order = np.argsort(M, kind="stable")
        _, _, I2, _ = O.tetra_weights(x, M, M)
        assert abs(np.dot(I2, np.sort(M)) - x) < 1e-9 * max(1.0, abs(x))
    M = np.array([0.1, 0.7, 1.3, 2.0])
    xs = np.linspace(0.1, 2.0, 20001)
    gs = np.array([O.tetra_weights(x, M, np.ones(4))[1] for x in xs])
    assert abs(np.trapezoid(gs, xs) - 1.0) < 1e-6
    assert O.tetra_weights(2.5, M, np.ones(4))[0] == 0


def test_lw3_tetra_agrees_with_gaussian_limit(rocksalt):
    """The tetrahedron linewidth and the Gaussian-delta linewidth approximate the same BZ
    integral: summed over branches and a spread of q0 on a 12^3 mesh (sigma = 0.6 rad/ps) they
    agree to a few per cent.  A flipped sign or a dropped difference process (each about 30-60 %
    of the total here) fails this."""
    from tests.conftest import problem
    qm, ph = problem(rocksalt, (12, 12, 12))
    inp = O.Inputs(rocksalt, qm, ph)
    tt = gg = 0.0
    for qi in range(1, qm.nq0, 12):
        ig = int(qm.irr_idx[qi])
        _, W = O.lw3_gauss(inp, ig, 1.0, want_w=True)
        lt, p3 = O.lw3_tetra(qm, ph, ig, W)
        assert np.all(p3 >= 0.0)
        tt += lt.sum()
        gg += O.lw3_gauss(inp, ig, 0.6).sum()
    assert abs(tt - gg) / gg < 0.05
