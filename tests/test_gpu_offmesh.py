"""LineWidth.x's off-mesh q0 path on the GPU (spr_lw3_offmesh) against the oracle's restatement
(pinned in test_offmesh_oracle.py): W of every grid q1 and the tetrahedron and Gaussian
linewidths, zero-centred and not, for 6 and 30 branches; and for a q0 on the grid the Gaussian
linewidth equals the grid path's.  The oracle takes q2's eigenvectors from the same GPU solve, so
that for a generic q0 the W rows agree one by one.  Where q2 has degenerate branches (a q0 on a
symmetry line) the basis of a degenerate subspace follows the last bits of q2, which the two sides
round differently: there W summed over each degenerate q2 subspace and the Gaussian linewidth
(both basis-free) are compared, and the tetrahedron sum, which weighs the branches of a subspace
by their neighbours' frequencies, is not."""
import numpy as np
import pytest

from tests.conftest import problem, rel_err

pytestmark = pytest.mark.gpu


def _dict(p, i=None):
    sl = (lambda a: a) if i is None else (lambda a: a[i])
    return dict(evec=sl(p.evec), omega=sl(p.omega), inv_omega=sl(p.inv_omega), nbe=sl(p.nbe))


def _group_sum(W, w2, tol=1e-7):
    """W[q1][s0][s1][s2] summed over the s2 of each degenerate subspace of q2 (w2[q1][s2])."""
    out = np.zeros_like(W)
    for i in range(W.shape[0]):
        lab = np.zeros(W.shape[3], int)
        for s in range(1, W.shape[3]):
            same = abs(w2[i, s] - w2[i, s - 1]) <= tol * max(1.0, abs(w2[i, s]))
            lab[s] = lab[s - 1] if same else lab[s - 1] + 1
        for g in range(lab.max() + 1):
            out[i, :, :, g] = W[i][:, :, lab == g].sum(axis=2)
    return out


def _check(m, qm, ph, ctx, q0, zero_centered, delta, sigma=0.1, exact=True):
    from spring_amd.lib import get_phonon
    from oracle import oracle as O
    cr = m.crystal
    p0 = _dict(get_phonon(cr, m.fc2, q0[None, :], np.ones((1, 3), np.int32), 300.0), 0)
    _, q2c, exists = O.offmesh_q2(m, qm, q0, zero_centered)
    g2 = np.all(np.abs(q2c) < 1e-10, axis=1)
    ph2 = _dict(get_phonon(cr, m.fc2, q2c, np.repeat(np.where(g2, 0, 1)[:, None], 3, axis=1), 300.0))
    ref, W_ref, _ = O.lw3_offmesh(m, qm, ph, q0, p0, 300.0, delta, sigma, zero_centered, ph2)
    lw, W = ctx.lw3_offmesh(m, q0, p0, 300.0, delta=delta, sigma=sigma, zero_centered=zero_centered, want_w=True)
    if not exact:
        W, W_ref = _group_sum(W, ph2["omega"]), _group_sum(W_ref, ph2["omega"])
    assert np.abs(W - W_ref).max() <= 1e-11 * W_ref.max(), (delta, zero_centered)
    assert ref.max() > 0
    if exact or delta == "gauss":
        assert rel_err(lw, ref, 1e-9) < 1e-10, (delta, zero_centered, lw, ref)
    return exists


@pytest.mark.parametrize("zero_centered", [True, False])
def test_offmesh_2atom(rocksalt, zero_centered):
    from spring_amd.lib import GpuContext
    qm, ph = problem(rocksalt, (4, 4, 4))
    ctx = GpuContext(rocksalt, qm, ph)
    G = rocksalt.crystal.G
    for frac, exact in (([0.31, -0.12, 0.07], True), ([0.125, 0.25, 0.0], False)):
        q0 = G @ np.array(frac)
        print("offmesh 2-atom", frac, zero_centered, flush=True)
        for delta in ("tetra", "gauss"):
            _check(rocksalt, qm, ph, ctx, q0, zero_centered, delta, exact=exact)
    ctx.close()


def test_offmesh_grid_q0_gauss(rocksalt):
    from spring_amd.lib import GpuContext
    qm, ph = problem(rocksalt, (4, 4, 4))
    ctx = GpuContext(rocksalt, qm, ph)
    grid = ctx.lw3_all(0.3, "fp64")
    for k in (1, 3, 5):
        i0 = int(qm.irr_idx[k])
        p0 = _dict(ph, i0)
        lw = ctx.lw3_offmesh(rocksalt, qm.q_cart[i0], p0, 300.0, delta="gauss", sigma=0.3)
        assert rel_err(lw, grid[k], 1e-9) < 1e-10, k
    ctx.close()


def test_offmesh_10atom(rocksalt10):
    from spring_amd.lib import GpuContext
    qm, ph = problem(rocksalt10, (2, 2, 2))
    ctx = GpuContext(rocksalt10, qm, ph)
    q0 = rocksalt10.crystal.G @ np.array([0.21, 0.13, -0.34])
    for delta in ("tetra", "gauss"):
        print("offmesh 10-atom", delta, flush=True)
        _check(rocksalt10, qm, ph, ctx, q0, True, delta, sigma=0.3)
    ctx.close()
