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


def test_full_size_tetra_2atom_16(rocksalt):
    """The reference's default delta at the 2-atom 16^3 size (145 irreducible q0, W of 16 q0 per
    launch): two seeded q0 of the all-q0 call equal the oracle's tetrahedron sums over its own W
    of all 4096 q1, and every linewidth and P3 is finite and non-negative in total."""
    from spring_amd.lib import GpuContext
    from oracle import oracle as O
    qm, ph = problem(rocksalt, (16, 16, 16))
    ctx = GpuContext(rocksalt, qm, ph)
    inp = O.Inputs(rocksalt, qm, ph)
    print("tetra 16^3: all q0 on the GPU", flush=True)
    lw_all, p3_all = ctx.lw3_tetra_all()
    assert np.all(np.isfinite(lw_all)) and np.all(np.isfinite(p3_all))
    assert np.all(lw_all.sum(axis=1) > 0.0) and np.all(p3_all >= 0.0)
    rng = np.random.default_rng(16)
    for qi in rng.choice(np.arange(1, qm.nq0), size=2, replace=False):
        ig = int(qm.irr_idx[qi])
        print(f"tetra 16^3: oracle q0 #{qi}", flush=True)
        _, W = O.lw3_gauss(inp, ig, 1.0, want_w=True)
        lw, p3 = O.lw3_tetra(qm, ph, ig, W)
        e1, e2 = rel_err(lw_all[qi], lw, 1e-9), rel_err(p3_all[qi], p3, 1e-9)
        print(f"tetra 16^3 q0 #{qi}: lw {e1:.2e} p3 {e2:.2e}", flush=True)
        assert e1 < 1e-10 and e2 < 1e-10


@pytest.mark.parametrize("cell", ["rocksalt4", "rocksalt10"])
def test_lw3_tetra_mfma_w_matches_generic(cell, rocksalt10, monkeypatch):
    """Cells of 3+ atoms take W from the matrix-core kernel's W mode (pair-reduced, both
    orderings written); the generic W path (SPR_KERNEL=generic) evaluates every q1 on its own.
    Every q0's tetrahedron linewidth and P3 agree, and equal the oracle's on sampled q0."""
    from spring_amd import model
    from spring_amd.lib import GpuContext
    from oracle import oracle as O
    m = rocksalt10 if cell == "rocksalt10" else model.rocksalt_chain(2)
    qm, ph = problem(m, (3, 3, 3))
    ctx = GpuContext(m, qm, ph)
    lw_m, p3_m = ctx.lw3_tetra_all()
    ctx.close()
    monkeypatch.setenv("SPR_KERNEL", "generic")
    ctx = GpuContext(m, qm, ph)
    lw_g, p3_g = ctx.lw3_tetra_all()
    ctx.close()
    assert lw_m.max() > 0
    assert rel_err(lw_m, lw_g, 1e-9) < 1e-11 and rel_err(p3_m, p3_g, 1e-9) < 1e-11
    inp = O.Inputs(m, qm, ph)
    for qi in (1, qm.nq0 - 1):
        ig = int(qm.irr_idx[qi])
        _, W = O.lw3_gauss(inp, ig, 1.0, want_w=True)
        lw, p3 = O.lw3_tetra(qm, ph, ig, W)
        assert rel_err(lw_m[qi], lw, 1e-9) < 1e-10 and rel_err(p3_m[qi], p3, 1e-9) < 1e-10, qi


def test_full_size_tetra_10atom_20(rocksalt10):
    """The reference's default delta on the 10-atom 20^3 workload (time reversal): one seeded
    q0's tetrahedron linewidth and P3 (W of all 8000 q1 from the matrix-core W mode, 1.7 GB)
    equal the oracle's tetrahedron sums over its own W."""
    from concurrent.futures import ThreadPoolExecutor
    from spring_amd.lib import GpuContext
    from oracle import oracle as O
    qm, ph = problem(rocksalt10, (20, 20, 20), time_rev=1)
    ctx = GpuContext(rocksalt10, qm, ph)
    inp = O.Inputs(rocksalt10, qm, ph)
    qi = int(np.random.default_rng(20).integers(1, qm.nq0))
    ig = int(qm.irr_idx[qi])
    g_lw, g_p3 = ctx.lw3_tetra(qi)
    ctx.close()
    chunks = [(lo, min(qm.nq, lo + 250)) for lo in range(0, qm.nq, 250)]
    print(f"tetra 20^3 10-atom: oracle W of q0 #{qi}", flush=True)
    with ThreadPoolExecutor(16) as ex:
        parts = list(ex.map(lambda c: O.lw3_gauss(inp, ig, 1.0, c[0], c[1], want_w=True)[1], chunks))
    W = np.concatenate(parts)
    lw, p3 = O.lw3_tetra(qm, ph, ig, W)
    e1, e2 = rel_err(g_lw, lw, 1e-9), rel_err(g_p3, p3, 1e-9)
    print(f"tetra 20^3 10-atom q0 #{qi}: lw {e1:.2e} p3 {e2:.2e}", flush=True)
    assert lw.max() > 0 and e1 < 1e-10 and e2 < 1e-10
