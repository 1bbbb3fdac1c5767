"""The oracle's off-mesh q0 linewidth (LineWidth.x, src/LineWidth/Src/Linewidth_m.f90:48-417),
pinned without restating it: for a q0 on the grid the G search must land every q2 on the grid
point -(q0 + q1), so with q2's phonons taken from the grid the W rows and both linewidths must
equal the grid path's (lw3_gauss, lw3_tetra, pinned elsewhere); the non-zero-centred grid of a
grid q0 must give the same numbers; and a q0 shifted by a reciprocal lattice vector must give the
same linewidth (the G search reaches q2 in either case)."""
import numpy as np

from oracle import oracle as O
from tests.conftest import problem


def _ph0(ph, i):
    return dict(evec=ph.evec[i], omega=ph.omega[i], inv_omega=ph.inv_omega[i])


def _grid_ph2(qm, ph, i0):
    """q2 = -(q0 + q1) of every q1 taken from the grid (same eigenvector bases as the grid path)."""
    partner = qm.index(-qm.q_int[i0][None, :].astype(np.int64) - qm.q_int.astype(np.int64))
    return dict(evec=ph.evec[partner], omega=ph.omega[partner], inv_omega=ph.inv_omega[partner],
                nbe=ph.nbe[partner])


def test_grid_q0_matches_grid_path(rocksalt):
    qm, ph = problem(rocksalt, (4, 4, 4))
    inp = O.Inputs(rocksalt, qm, ph)
    for i0 in (int(qm.irr_idx[3]), int(qm.irr_idx[5])):
        q0 = qm.q_cart[i0]
        _, q2c, exists = O.offmesh_q2(rocksalt, qm, q0, True)
        assert exists.all()
        # every q2 is -(q0 + q1) up to a reciprocal lattice vector
        x = np.linalg.solve(rocksalt.crystal.G, (q2c + q0 + qm.q_cart).T).T
        assert np.allclose(x, np.round(x), atol=1e-12)
        ph2 = _grid_ph2(qm, ph, i0)
        lw_g, W, _ = O.lw3_offmesh(rocksalt, qm, ph, q0, _ph0(ph, i0), 300.0, "gauss", 0.3, True, ph2)
        ref_g, W_ref = O.lw3_gauss(inp, i0, 0.3, want_w=True)
        assert np.allclose(W, W_ref, rtol=0, atol=1e-12 * W_ref.max())
        assert np.allclose(lw_g, ref_g, rtol=1e-11, atol=1e-13 * ref_g.max())
        lw_t, _, _ = O.lw3_offmesh(rocksalt, qm, ph, q0, _ph0(ph, i0), 300.0, "tetra", 0.3, True, ph2)
        ref_t, _ = O.lw3_tetra(qm, ph, i0, W_ref)
        assert ref_t.max() > 0
        assert np.allclose(lw_t, ref_t, rtol=1e-11, atol=1e-13 * ref_t.max())
        # the non-zero-centred grid (q1 in [0, 1) reduced coordinates, q2 with 0 <= x <= 1)
        lw_n, _, ex_n = O.lw3_offmesh(rocksalt, qm, ph, q0, _ph0(ph, i0), 300.0, "tetra", 0.3, False, ph2)
        assert ex_n.all()
        assert np.allclose(lw_n, ref_t, rtol=1e-11, atol=1e-13 * ref_t.max())


def test_offgrid_q0_shift_by_G(rocksalt):
    from spring_amd import phonons
    qm, ph = problem(rocksalt, (4, 4, 4))
    G = rocksalt.crystal.G
    q0 = G @ np.array([0.31, -0.12, 0.07])
    p0 = phonons.solve(rocksalt.crystal, rocksalt.fc2, q0[None, :], np.ones((1, 3), np.int32), 300.0)
    ph0 = dict(evec=p0.evec[0], omega=p0.omega[0], inv_omega=p0.inv_omega[0])
    lw, W, ex = O.lw3_offmesh(rocksalt, qm, ph, q0, ph0, 300.0, "tetra")
    assert ex.all() and lw.max() > 0
    lw_s, W_s, _ = O.lw3_offmesh(rocksalt, qm, ph, q0 + G[:, 0], ph0, 300.0, "tetra")
    assert np.allclose(W_s, W, rtol=0, atol=1e-12 * W.max())
    assert np.allclose(lw_s, lw, rtol=1e-11, atol=1e-13 * lw.max())
    # a generic q0 has no grid partner: the q2 differ from every grid point
    _, q2c, _ = O.offmesh_q2(rocksalt, qm, q0, True)
    x = np.linalg.solve(G, q2c.T).T * qm.mesh
    assert np.abs(x - np.round(x)).min() > 1e-3


def test_q2_search_order_and_boundary(rocksalt):
    """The first G of (0, 1, -1, 2, -2)^3 wins: a q2 on the zone boundary (x = 1/2) keeps the
    G found first, not its image at -1/2."""
    qm, _ = problem(rocksalt, (4, 4, 4))
    G = rocksalt.crystal.G
    q0 = G @ np.array([0.5, 0.0, 0.0])
    q1c, q2c, ex = O.offmesh_q2(rocksalt, qm, q0, True)
    assert ex.all()
    x = np.linalg.solve(G, q2c.T).T
    assert np.all(np.abs(x) <= 0.5 + 1e-12)
    i = int(np.flatnonzero(np.all(qm.q_int == 0, axis=1))[0])   # q1 = 0: q2 = G - q0
    # G = 0 gives x = -1/2 (inside), found before G = b1 (x = +1/2)
    assert np.allclose(x[i], [-0.5, 0.0, 0.0], atol=1e-12)


def test_path_points_closed_form():
    """q_points_highsymm on a cubic cell (G = 2 pi I): the points of a segment are k/count of the
    way from its start (start excluded, end included), distances are cumulative arc length, a
    segment's zero-centred flag follows the rule on its two ends, a zero-count segment adds no
    point and keeps the distance, and the first node's count is not used."""
    G = 2.0 * np.pi * np.eye(3)
    nodes = [[0.0, 0.0, 0.0, 7],            # count of the first node: ignored
             [0.0, 0.5, 0.5, 4],            # Gamma -> X: |start| <= 1/2 -> zero-centred
             [0.6, 0.7, 0.8, 2],            # X -> generic: |X| <= 1/2 -> zero-centred
             [0.9, 0.6, 0.7, 3],            # both ends outside [-1/2, 1/2], no negative: not
             [0.9, 0.6, 0.7, 0],            # zero-count segment
             [0.9, -0.6, 0.7, 1]]           # a negative coordinate -> zero-centred
    q, d, zc = O.q_points_highsymm(nodes, G)
    assert q.shape == (4 + 2 + 3 + 1, 3)
    x = q / (2.0 * np.pi)
    assert np.allclose(x[:4], [[0, k / 8, k / 8] for k in (1, 2, 3, 4)], atol=1e-15)
    a, b = np.array([0.0, 0.5, 0.5]), np.array([0.6, 0.7, 0.8])
    assert np.allclose(x[4:6], [a + (b - a) * k / 2 for k in (1, 2)], atol=1e-14)
    c = np.array([0.9, 0.6, 0.7])
    assert np.allclose(x[6:9], [b + (c - b) * k / 3 for k in (1, 2, 3)], atol=1e-14)
    assert np.allclose(x[9], [0.9, -0.6, 0.7], atol=1e-14)
    L = 2.0 * np.pi * np.array([np.linalg.norm([0, 0.5, 0.5]), np.linalg.norm(b - a), np.linalg.norm(c - b), 1.2])
    cum = np.cumsum(L)
    assert np.allclose(d[[3, 5, 8, 9]], cum, rtol=1e-14)
    assert np.allclose(d[:4], L[0] * np.arange(1, 5) / 4, rtol=1e-14)
    assert np.all(np.diff(d) > 0)
    assert zc.tolist() == [True] * 6 + [False] * 3 + [True]
