"""Branch selection rules in the oracle (PolarizationStrtEnd, Helper_m.f90:310-353; the ranges of
CalculateLwTetrahedron's loops, ThirdOrder_m.f90:207-214, of TetrahedronIntegration4th's,
Tetrahedron4th_sep.f90:213-224, and of kappa_noIter's mode loop, ThermalConductivity.f90:106-111).

Pinned without restating the loops: acoustic (A) and optical (O) branches partition every branch
(D), so each sum restricted on one slot to A plus the same restricted to O is the unrestricted
sum; and a W with a single nonzero branch triplet shows which slot each range restricts."""
import numpy as np
import pytest

from oracle import oracle as O
from tests.conftest import problem


def _rules(order, nd, slot, letter):
    r = ["D"] * order
    r[slot] = letter
    return O.polarization(order, nd, r)


def test_polarization_letters():
    assert list(O.polarization(3, 6, "AOD")) == [0, 2, 3, 5, 0, 5]
    assert list(O.polarization(4, 30, "oaxD")) == [3, 29, 0, 2, 0, 29, 0, 29]


def test_lw3_from_w_slots_and_partition():
    rng = np.random.default_rng(0)
    nd = 6
    w0, w1, w2 = (rng.uniform(1, 20, nd) for _ in range(3))
    n1, n2 = rng.uniform(0.1, 3, nd), rng.uniform(0.1, 3, nd)
    sigma = 30.0                       # every Gaussian well above zero
    W = np.zeros((nd, nd, nd))
    W[4, 1, 5] = 1.0                   # only (s0, s1, s2) = (4, 1, 5)
    full = O.lw3_from_w(W, w0, w1, w2, n1, n2, sigma)
    assert full[4] != 0 and np.count_nonzero(full) == 1
    for slot, letter, keeps in ((1, "A", True), (1, "O", False), (2, "A", False), (2, "O", True),
                                (0, "O", True), (0, "A", False)):
        got = O.lw3_from_w(W, w0, w1, w2, n1, n2, sigma, pol=_rules(3, nd, slot, letter))
        assert np.array_equal(got, full if keeps else np.zeros(nd)), (slot, letter)
    W = rng.uniform(0, 1, (nd, nd, nd))
    full = O.lw3_from_w(W, w0, w1, w2, n1, n2, sigma)
    for slot in range(3):
        a = O.lw3_from_w(W, w0, w1, w2, n1, n2, sigma, pol=_rules(3, nd, slot, "A"))
        o = O.lw3_from_w(W, w0, w1, w2, n1, n2, sigma, pol=_rules(3, nd, slot, "O"))
        assert np.allclose(a + o, full, rtol=1e-13, atol=0), slot


def test_lw3_tetra_partition(rocksalt):
    qm, ph = problem(rocksalt, (4, 4, 4))
    inp = O.Inputs(rocksalt, qm, ph)
    i0 = int(qm.irr_idx[3])
    _, W = O.lw3_gauss(inp, i0, 1.0, want_w=True)
    lw, p3 = O.lw3_tetra(qm, ph, i0, W)
    assert lw.max() > 0
    for slot in range(3):
        la, pa = O.lw3_tetra(qm, ph, i0, W, pol=_rules(3, 6, slot, "A"))
        lo, po = O.lw3_tetra(qm, ph, i0, W, pol=_rules(3, 6, slot, "O"))
        assert np.allclose(la + lo, lw, rtol=1e-12, atol=1e-14 * lw.max()), slot
        assert np.allclose(pa + po, p3, rtol=1e-12, atol=1e-14 * p3.max()), slot
    # s0 restricted to A: the optical rows are zero and the acoustic ones unchanged
    la, pa = O.lw3_tetra(qm, ph, i0, W, pol=_rules(3, 6, 0, "A"))
    assert np.array_equal(la[3:], np.zeros(3)) and np.array_equal(la[:3], lw[:3])


def test_lw3_gauss_partition(rocksalt):
    qm, ph = problem(rocksalt, (4, 4, 4))
    inp = O.Inputs(rocksalt, qm, ph)
    i0 = int(qm.irr_idx[2])
    full = O.lw3_gauss(inp, i0, 0.3)
    for slot in (1, 2):
        a = O.lw3_gauss(inp, i0, 0.3, pol=_rules(3, 6, slot, "A"))
        o = O.lw3_gauss(inp, i0, 0.3, pol=_rules(3, 6, slot, "O"))
        assert np.allclose(a + o, full, rtol=1e-12, atol=1e-14 * np.abs(full).max()), slot


def test_lw4_partition(rocksalt):
    qm, ph = problem(rocksalt, (3, 3, 3))
    inp = O.Inputs(rocksalt, qm, ph, which=4)
    i0 = int(qm.irr_idx[1])
    full = O.lw4_gauss(inp, i0, 2.0)
    W = O.w4_rows(inp, i0, 0, qm.nq)
    tfull = O.lw4_tetra_from_w(qm.mesh, qm.q_int, qm.tetra, i0, 0, W, ph.omega, ph.nbe)
    assert full.max() > 0 and tfull.max() > 0
    for slot in range(4):
        a = O.lw4_gauss(inp, i0, 2.0, pol=_rules(4, 6, slot, "A"))
        o = O.lw4_gauss(inp, i0, 2.0, pol=_rules(4, 6, slot, "O"))
        assert np.allclose(a + o, full, rtol=1e-12, atol=1e-14 * full.max()), slot
        ta = O.lw4_tetra_from_w(qm.mesh, qm.q_int, qm.tetra, i0, 0, W, ph.omega, ph.nbe,
                                pol=_rules(4, 6, slot, "A"))
        to = O.lw4_tetra_from_w(qm.mesh, qm.q_int, qm.tetra, i0, 0, W, ph.omega, ph.nbe,
                                pol=_rules(4, 6, slot, "O"))
        assert np.allclose(ta + to, tfull, rtol=1e-12, atol=1e-14 * tfull.max()), slot


@pytest.mark.parametrize("symm", [False, True])
def test_kappa_partition(rocksalt, symm):
    qm, ph = problem(rocksalt, (4, 4, 4))
    lw = np.random.default_rng(2).uniform(0.5, 1.5, (qm.nq0, 6))
    f = O.kappa_rta_symm if symm else O.kappa_rta
    vol = rocksalt.crystal.vol
    k = f(qm, ph, lw, 300.0, vol)
    ka = f(qm, ph, lw, 300.0, vol, srange=[0, 2])
    ko = f(qm, ph, lw, 300.0, vol, srange=[3, 5])
    assert np.allclose(ka + ko, k, rtol=1e-12, atol=1e-14 * np.abs(k).max())
    assert np.abs(ka).max() > 0 and np.abs(ko).max() > 0
