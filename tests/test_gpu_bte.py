"""GPU parity of the iterative Boltzmann transport solution (kappa_iter.x) against the FP64
CPU oracle: the time-reversal irreducible set, 1/tau, one dF for a random F, and the whole
self-consistent loop (kappa, kappa of F0, iteration count)."""
import numpy as np
import pytest

from oracle import oracle as O
from tests.conftest import problem, rel_err

pytestmark = pytest.mark.gpu


def _setup(model, dims):
    from spring_amd.lib import GpuContext
    qm, ph = problem(model, dims)
    ctx = GpuContext(model, qm, ph)
    inp = O.Inputs(model, qm, ph)
    irr, map_irr, map_sign, deg = O.tr_irreducible(qm)
    W = np.array([O.lw3_gauss(inp, int(g), 1.0, want_w=True)[1] for g in irr])
    Wp, Wm = O.bte_w_pm(qm, ph, irr, W)
    return ctx, qm, ph, (irr, map_irr, map_sign, deg, Wp, Wm)


@pytest.mark.parametrize("dims", [(4, 4, 4), (5, 5, 5)])
def test_bte_tau_inv_and_delta_f(rocksalt, dims):
    ctx, qm, ph, (irr, map_irr, map_sign, deg, Wp, Wm) = _setup(rocksalt, dims)
    g_irr, g_deg = ctx.bte_irr()
    assert np.array_equal(g_irr, irr) and np.array_equal(g_deg, deg)
    itau = ctx.bte_tau_inv()
    ref = O.bte_sums(qm, ph, irr, map_irr, map_sign, Wp, Wm)
    e = rel_err(itau, ref, 1e-9)
    print(f"{dims} 1/tau max rel err {e:.2e}", flush=True)
    assert e < 1e-10
    # a random F makes some dF components cancel to near zero: those are held to the sum's
    # rounding (normwise 1e-12), the others elementwise to 1e-10
    F = np.random.default_rng(7).normal(size=(len(irr), ph.omega.shape[1], 3))
    dF = ctx.bte_delta_f(F)
    refF = O.bte_sums(qm, ph, irr, map_irr, map_sign, Wp, Wm, F)
    en = np.max(np.abs(dF - refF)) / np.max(np.abs(refF))
    e = rel_err(dF, refF, 1e-4)
    print(f"{dims} dF normwise rel err {en:.2e}, elementwise (above 1e-4 of max) {e:.2e}", flush=True)
    assert en < 1e-12 and e < 1e-10
    # the store of W (the small-cell kernel's W mode over every q1) against the generic kernel's
    from spring_amd.lib import GpuContext
    import os
    os.environ["SPR_KERNEL"] = "generic"
    try:
        cg = GpuContext(rocksalt, qm, ph)
        assert rel_err(itau, cg.bte_tau_inv(), 1e-9) < 1e-12
        dF_g = cg.bte_delta_f(F)
        assert np.max(np.abs(dF - dF_g)) / np.max(np.abs(dF_g)) < 1e-12
        cg.close()
    finally:
        del os.environ["SPR_KERNEL"]


def test_bte_solve(rocksalt):
    ctx, qm, ph, (irr, map_irr, map_sign, deg, Wp, Wm) = _setup(rocksalt, (4, 4, 4))
    T, vol = 300.0, rocksalt.crystal.vol
    out = ctx.bte_solve(T, eps_k=1e-6, max_iter=100)
    k, k0, it, F = O.bte_solve(qm, ph, vol, T, irr, map_irr, map_sign, deg, Wp, Wm, 1e-6, 100)
    print("bte kappa", np.diag(out["kappa"]), "iterations", out["iterations"], flush=True)
    _check_solve(out, k, k0, it, F)


def _check_solve(out, k, k0, it, F):
    # the collision matrix is summed with LDS atomics, so its last bits (and, when the eps_k test
    # is met by a hair, the iteration count) can vary between runs; then kappa is the neighbouring
    # iterate, which differs from the oracle's by less than eps_k
    assert abs(out["iterations"] - it) <= 1
    assert rel_err(out["kappa0"], k0, 1e-6) < 1e-10
    if out["iterations"] == it:
        assert rel_err(out["kappa"], k, 1e-6) < 1e-9
        assert rel_err(out["F"], F, 1e-6) < 1e-8
    else:
        assert np.abs(out["kappa"] - k).max() < 1e-6


def test_bte_fallback_path(rocksalt, monkeypatch):
    """SPR_BTE_NO_MATRIX=1: 1/tau from the tetrahedron sums (k_bte<0>) and the solve with the
    per-iteration sums and the host F update -- the path of meshes whose collision matrix does
    not fit -- against the oracle and against the matrix path."""
    ctx_m, qm, ph, (irr, map_irr, map_sign, deg, Wp, Wm) = _setup(rocksalt, (4, 4, 4))
    T, vol = 300.0, rocksalt.crystal.vol
    itau_m = ctx_m.bte_tau_inv()
    out_m = ctx_m.bte_solve(T, eps_k=1e-6, max_iter=100)
    ctx_m.close()
    monkeypatch.setenv("SPR_BTE_NO_MATRIX", "1")
    from spring_amd.lib import GpuContext
    ctx = GpuContext(rocksalt, qm, ph)
    itau = ctx.bte_tau_inv()
    ref = O.bte_sums(qm, ph, irr, map_irr, map_sign, Wp, Wm)
    assert rel_err(itau, ref, 1e-9) < 1e-10
    assert rel_err(itau, itau_m, 1e-9) < 1e-12
    out = ctx.bte_solve(T, eps_k=1e-6, max_iter=100)
    k, k0, it, F = O.bte_solve(qm, ph, vol, T, irr, map_irr, map_sign, deg, Wp, Wm, 1e-6, 100)
    print("fallback kappa", np.diag(out["kappa"]), "iterations", out["iterations"], flush=True)
    _check_solve(out, k, k0, it, F)
    if out["iterations"] == out_m["iterations"]:
        assert rel_err(out["kappa"], out_m["kappa"], 1e-6) < 1e-9
    ctx.close()


def test_bte_matrix_core_w_store_4atom(monkeypatch):
    """4-atom cell (Ndof 12), 3^3 mesh: the W store of cells with three or more atoms comes from
    the matrix-core kernel's W mode over every q1 of each representative (each ordering evaluated
    on its own).  1/tau and dF against the oracle as for the 2-atom cell, the same against the
    generic-kernel store (SPR_KERNEL=generic) to rounding, and the whole loop against the
    oracle's."""
    from spring_amd import model
    m = model.rocksalt_chain(2)
    ctx, qm, ph, (irr, map_irr, map_sign, deg, Wp, Wm) = _setup(m, (3, 3, 3))
    g_irr, g_deg = ctx.bte_irr()
    assert np.array_equal(g_irr, irr) and np.array_equal(g_deg, deg)
    itau = ctx.bte_tau_inv()
    ref = O.bte_sums(qm, ph, irr, map_irr, map_sign, Wp, Wm)
    e = rel_err(itau, ref, 1e-9)
    print(f"4-atom 1/tau max rel err {e:.2e}", flush=True)
    assert e < 1e-10
    F = np.random.default_rng(5).normal(size=(len(irr), ph.omega.shape[1], 3))
    dF = ctx.bte_delta_f(F)
    refF = O.bte_sums(qm, ph, irr, map_irr, map_sign, Wp, Wm, F)
    en = np.max(np.abs(dF - refF)) / np.max(np.abs(refF))
    e = rel_err(dF, refF, 1e-4)
    print(f"4-atom dF normwise rel err {en:.2e}, elementwise (above 1e-4 of max) {e:.2e}", flush=True)
    assert en < 1e-12 and e < 1e-10
    T, vol = 300.0, m.crystal.vol
    out = ctx.bte_solve(T, eps_k=1e-6, max_iter=100)
    k, k0, it, Fs = O.bte_solve(qm, ph, vol, T, irr, map_irr, map_sign, deg, Wp, Wm, 1e-6, 100)
    _check_solve(out, k, k0, it, Fs)
    from spring_amd.lib import GpuContext
    monkeypatch.setenv("SPR_KERNEL", "generic")
    cg = GpuContext(m, qm, ph)
    itau_g = cg.bte_tau_inv()
    dF_g = cg.bte_delta_f(F)
    monkeypatch.delenv("SPR_KERNEL")
    assert rel_err(itau, itau_g, 1e-9) < 1e-12
    assert np.max(np.abs(dF - dF_g)) / np.max(np.abs(dF_g)) < 1e-12
    cg.close()
    ctx.close()
