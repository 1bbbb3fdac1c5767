"""A one-atom cell (monatomic fcc, three branches) through every kernel built for Ndof = 3: the
small-cell three-phonon kernel in its Gaussian and W-output modes (the tetrahedron path), the
generic kernel, the three-phonon tetrahedron sums, and the four-phonon Gaussian and tetrahedron
sums -- each against the oracle."""
import numpy as np
import pytest

from tests.conftest import problem, rel_err

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def fcc():
    from spring_amd import model
    return model.fcc1(with_fc4=True)


@pytest.mark.parametrize("dims", [(3, 3, 3), (4, 4, 4)])
@pytest.mark.parametrize("kernel", ["small", "generic"])
def test_one_atom_lw3(fcc, dims, kernel, monkeypatch):
    from spring_amd.lib import GpuContext
    from oracle import oracle as O
    monkeypatch.setenv("SPR_KERNEL", kernel)
    qm, ph = problem(fcc, dims)
    ctx = GpuContext(fcc, qm, ph)
    inp = O.Inputs(fcc, qm, ph)
    ref = O.lw3_gauss_all(inp, qm.irr_idx, 0.5, threads=8)
    assert ref.max() > 0
    print("one-atom lw3", dims, kernel, flush=True)
    assert rel_err(ctx.lw3_all(0.5, "fp64"), ref, 1e-9) < 1e-10
    assert rel_err(ctx.lw3_all(0.5, "fp32"), ref, 1e-6) < 1e-4
    ctx.close()


@pytest.mark.parametrize("dims", [(3, 3, 3), (4, 4, 4)])
def test_one_atom_lw3_tetra(fcc, dims):
    from spring_amd.lib import GpuContext
    from oracle import oracle as O
    qm, ph = problem(fcc, dims)
    ctx = GpuContext(fcc, qm, ph)
    inp = O.Inputs(fcc, qm, ph)
    lw_all, p3_all = ctx.lw3_tetra_all()
    for qi in range(qm.nq0):
        ig = int(qm.irr_idx[qi])
        _, W = O.lw3_gauss(inp, ig, 1.0, want_w=True)
        lw, p3 = O.lw3_tetra(qm, ph, ig, W)
        assert rel_err(lw_all[qi], lw, 1e-9) < 1e-10, qi
        assert rel_err(p3_all[qi], p3, 1e-9) < 1e-10, qi
    assert lw_all.max() > 0
    ctx.close()


@pytest.mark.parametrize("dims", [(3, 3, 3), (4, 4, 4)])
def test_one_atom_lw4(fcc, dims):
    from spring_amd.lib import GpuContext
    from oracle import oracle as O
    qm, ph = problem(fcc, dims)
    ctx = GpuContext(fcc, qm, ph)
    inp = O.Inputs(fcc, qm, ph, which=4)
    ref = O.lw4_gauss_all(inp, qm.irr_idx, 2.0, threads=8)
    assert ref.max() > 0
    print("one-atom lw4", dims, flush=True)
    assert rel_err(ctx.lw4_all(2.0, "fp64"), ref, 1e-9) < 1e-10
    tl = ctx.lw4_tetra_all()
    for qi in range(min(qm.nq0, 4)):
        tref = O.lw4_tetra(inp, qm, int(qm.irr_idx[qi]), threads=8, chunk=4)
        assert rel_err(tl[qi], tref, 1e-9) < 1e-10, qi
    ctx.close()


def test_one_atom_bte(fcc):
    """Iterative BTE of the one-atom cell on a 4^3 mesh: the W store from the small-cell kernel's
    Ndof = 3 W mode over every q1 of each time-reversal representative; 1/tau and dF of a random
    F against the oracle's tetrahedron sums (dF held normwise, and elementwise above 1e-4 of its
    largest component, where a random F leaves no cancellation to rounding)."""
    from spring_amd.lib import GpuContext
    from oracle import oracle as O
    qm, ph = problem(fcc, (4, 4, 4))
    ctx = GpuContext(fcc, qm, ph)
    inp = O.Inputs(fcc, qm, ph)
    irr, map_irr, map_sign, deg = O.tr_irreducible(qm)
    W = np.array([O.lw3_gauss(inp, int(g), 1.0, want_w=True)[1] for g in irr])
    Wp, Wm = O.bte_w_pm(qm, ph, irr, W)
    g_irr, g_deg = ctx.bte_irr()
    assert np.array_equal(g_irr, irr) and np.array_equal(g_deg, deg)
    itau = ctx.bte_tau_inv()
    ref = O.bte_sums(qm, ph, irr, map_irr, map_sign, Wp, Wm)
    assert np.abs(ref).max() > 0
    assert rel_err(itau, ref, 1e-9) < 1e-10
    F = np.random.default_rng(3).normal(size=(len(irr), ph.omega.shape[1], 3))
    dF = ctx.bte_delta_f(F)
    refF = O.bte_sums(qm, ph, irr, map_irr, map_sign, Wp, Wm, F)
    en = np.max(np.abs(dF - refF)) / np.max(np.abs(refF))
    print(f"one-atom BTE: 1/tau {rel_err(itau, ref, 1e-9):.2e}, dF normwise {en:.2e}", flush=True)
    assert en < 1e-12 and rel_err(dF, refF, 1e-4) < 1e-10
    ctx.close()
