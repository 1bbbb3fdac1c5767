"""Pins the oracle's iterative-BTE restatement (src/Kappa/Src_iterative) before it judges the GPU.

- 1/tau equals twice the tetrahedron linewidth of the Kappa path (orc_lw3_tetra, itself pinned
  in test_oracle.py): the + process of the iterative code is the M2 = w2 - w1 process of
  CalculateLwTetrahedron, its - process is M1 = w1 + w2 with half weight, the M3 term of the
  linewidth equals minus the M2 term by the q1 <-> q2 relabelling, and Sctr3THz2 of the
  iterative code is 4x the linewidth one: 1/tau = 4 (A2 + A1/2) = 2 (A1 + 2 A2) = 2 lw.
- With W+ = W- = 1 and F = a w on the full grid the vertex weights reproduce Mqs:
  dF(q, s) = a w(q, s) 1/tau(q, s) exactly.  A wrong partner point, a sign, a dropped F term
  or a wrong process weight fails it.
- dF is linear in F; F(-q) = -F(q) is honoured through the map.
- kappa of F0 (first iteration) equals the full-mesh RTA kappa of the Kappa path with the
  tetrahedron linewidths, which pins F0 and the kappa formula against orc_kappa_rta.
"""
import numpy as np
import pytest

from oracle import oracle as O
from tests.conftest import problem


@pytest.fixture(scope="module")
def bte4(rocksalt):
    qm, ph = problem(rocksalt, (4, 4, 4))
    inp = O.Inputs(rocksalt, qm, ph)
    irr, map_irr, map_sign, deg = O.tr_irreducible(qm)
    W = np.array([O.lw3_gauss(inp, int(g), 1.0, want_w=True)[1] for g in irr])
    Wp, Wm = O.bte_w_pm(qm, ph, irr, W)
    return dict(qm=qm, ph=ph, inp=inp, irr=irr, map_irr=map_irr, map_sign=map_sign, deg=deg, W=W,
                Wp=Wp, Wm=Wm, model=rocksalt)


def test_tr_irreducible_counts(bte4):
    qm = bte4["qm"]
    irr, deg = bte4["irr"], bte4["deg"]
    # 4^3 grid: 8 points with q = -q, the other 56 in 28 pairs
    assert len(irr) == 36 and deg.sum() == qm.nq and (deg == 1).sum() == 8
    neg = qm.index(-qm.q_int.astype(np.int64))
    for g in range(qm.nq):
        k = bte4["map_irr"][g]
        assert irr[k] in (g, neg[g])
        assert bte4["map_sign"][g] == (1 if irr[k] == g else -1)


def test_tau_inv_is_twice_the_tetrahedron_linewidth(bte4):
    qm, ph = bte4["qm"], bte4["ph"]
    itau = O.bte_sums(qm, ph, bte4["irr"], bte4["map_irr"], bte4["map_sign"], bte4["Wp"], bte4["Wm"])
    for k, g in enumerate(bte4["irr"]):
        lw, _ = O.lw3_tetra(qm, ph, int(g), bte4["W"][k])
        sel = slice(3, None) if g == 0 else slice(None)
        ref = 2.0 * lw[sel]
        assert np.allclose(itau[k][sel], ref, rtol=1e-9, atol=1e-12 * np.abs(ref).max())


def test_delta_f_reproduces_mqs_with_unit_weights(bte4):
    qm, ph = bte4["qm"], bte4["ph"]
    nq, nd = qm.nq, ph.omega.shape[1]
    full = np.arange(nq, dtype=np.int32)
    ones = np.ones((nq, nq, nd, nd, nd))
    a = np.array([0.3, -1.1, 0.7])
    F = ph.omega[:, :, None] * a[None, None, :]
    itau = O.bte_sums(qm, ph, full, full, np.ones(nq, np.int32), ones, ones)
    dF = O.bte_sums(qm, ph, full, full, np.ones(nq, np.int32), ones, ones, F)
    ref = ph.omega[:, :, None] * a[None, None, :] * itau[:, :, None]
    assert np.abs(itau).max() > 0
    assert np.allclose(dF, ref, rtol=1e-10, atol=1e-12 * np.abs(ref).max())


def test_delta_f_is_linear_and_odd(bte4):
    qm, ph = bte4["qm"], bte4["ph"]
    args = (qm, ph, bte4["irr"], bte4["map_irr"], bte4["map_sign"], bte4["Wp"], bte4["Wm"])
    rng = np.random.default_rng(3)
    F = rng.normal(size=(len(bte4["irr"]), ph.omega.shape[1], 3))
    G = rng.normal(size=F.shape)
    d = O.bte_sums(*args, 2.0 * F - 0.5 * G)
    ref = 2.0 * O.bte_sums(*args, F) - 0.5 * O.bte_sums(*args, G)
    assert np.allclose(d, ref, rtol=1e-10, atol=1e-12 * np.abs(ref).max())
    assert np.all(O.bte_sums(*args, np.zeros_like(F)) == 0.0)


def test_kappa_of_f0_is_the_rta_kappa(bte4, rocksalt):
    qm, ph = bte4["qm"], bte4["ph"]
    T = 300.0
    vol = rocksalt.crystal.vol
    k, k0, it, F = O.bte_solve(qm, ph, vol, T, bte4["irr"], bte4["map_irr"], bte4["map_sign"], bte4["deg"],
                               bte4["Wp"], bte4["Wm"], eps_k=1e-9, max_iter=200)
    # full-mesh RTA kappa of the Kappa path (orc_kappa_rta) with each grid point's linewidth the
    # tetrahedron linewidth of its time-reversal representative (the tetrahedra break the point
    # group, so the per-q rates are not symmetrised here, as in the iterative code)
    # group velocities of degenerate branches are not odd in q point by point, so every grid point
    # takes its representative's w, n and (sign-flipped) v, as the iterative code does
    lw = np.array([O.lw3_tetra(qm, ph, int(g), bte4["W"][k])[0] for k, g in enumerate(bte4["irr"])])
    lw[bte4["irr"] == 0, :3] = 1.0    # skipped modes: any finite value
    rep = bte4["irr"][bte4["map_irr"]]
    tr = type("TR", (), dict(q_int=qm.q_int, pos=bte4["map_irr"], nq=qm.nq))
    phr = type("PH", (), dict(omega=ph.omega[rep], nbe=ph.nbe[rep],
                              grp_vel=ph.grp_vel[rep] * bte4["map_sign"][:, None, None]))
    krta = O.kappa_rta(tr, phr, lw, T, vol)
    assert np.allclose(k0, krta, rtol=1e-8, atol=1e-10 * np.abs(krta).max())
    # the iteration converges and moves kappa away from RTA
    assert it < 200 and np.all(np.isfinite(k))
    assert abs(np.trace(k) - np.trace(k0)) > 1e-6 * abs(np.trace(k0))
