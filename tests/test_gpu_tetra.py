"""Tetrahedron-method three-phonon linewidths (CalculateLwTetrahedron) on the GPU against the
FP64 oracle, which integrates the oracle's own W of every q1 with the same method."""
import numpy as np
import pytest

from tests.conftest import problem, rel_err

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dims", [(4, 4, 4), (5, 5, 5)])
def test_lw3_tetra_parity_2atom(rocksalt, dims):
    from spring_amd.lib import GpuContext
    from oracle import oracle as O
    qm, ph = problem(rocksalt, dims)
    ctx = GpuContext(rocksalt, qm, ph)
    inp = O.Inputs(rocksalt, qm, ph)
    lw_all, p3_all = ctx.lw3_tetra_all()
    for qi in range(qm.nq0):
        ig = int(qm.irr_idx[qi])
        _, W = O.lw3_gauss(inp, ig, 1.0, want_w=True)
        lw, p3 = O.lw3_tetra(qm, ph, ig, W)
        assert rel_err(lw_all[qi], lw, 1e-9) < 1e-10, qi
        assert rel_err(p3_all[qi], p3, 1e-9) < 1e-10, qi
    one, p3one = ctx.lw3_tetra(2)
    assert np.array_equal(one, lw_all[2]) and np.array_equal(p3one, p3_all[2])


def test_lw3_tetra_10atom(rocksalt10):
    from spring_amd.lib import GpuContext
    from oracle import oracle as O
    qm, ph = problem(rocksalt10, (2, 2, 2))
    ctx = GpuContext(rocksalt10, qm, ph)
    inp = O.Inputs(rocksalt10, qm, ph)
    qi = 1
    ig = int(qm.irr_idx[qi])
    _, W = O.lw3_gauss(inp, ig, 1.0, want_w=True)
    lw, p3 = O.lw3_tetra(qm, ph, ig, W)
    g_lw, g_p3 = ctx.lw3_tetra(qi)
    assert rel_err(g_lw, lw, 1e-9) < 1e-10
    assert rel_err(g_p3, p3, 1e-9) < 1e-10
