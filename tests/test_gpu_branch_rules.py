"""Branch selection rules (spr_set_branch_rules; SelectionRule3/4 -> Polarization3rd/4th) on the
GPU against the oracle's ranges (pinned in test_branch_rules_oracle.py): every three-phonon kernel
(small-cell, matrix-core in FP64 and FP32, generic), the tetrahedron sums, the four-phonon Gaussian
and tetrahedron sums, and the kappa mode range; plus the spectrally resolved kappa."""
import numpy as np
import pytest

from tests.conftest import problem, rel_err

pytestmark = pytest.mark.gpu

RULES3 = ["AOD", "DOA", "OAA", "DDA"]


def _pol(order, nd, rules):
    from oracle import oracle as O
    return O.polarization(order, nd, rules)


@pytest.mark.parametrize("kernel", ["small", "mfma", "generic"])
def test_lw3_gauss_rules_2atom(rocksalt, kernel, monkeypatch):
    from spring_amd.lib import GpuContext
    from oracle import oracle as O
    monkeypatch.setenv("SPR_KERNEL", kernel)
    qm, ph = problem(rocksalt, (4, 4, 4))
    ctx = GpuContext(rocksalt, qm, ph)
    inp = O.Inputs(rocksalt, qm, ph)
    base = ctx.lw3_all(0.3, "fp64")
    for rules in RULES3:
        print("lw3 rules", kernel, rules, flush=True)
        ctx.set_branch_rules(3, rules)
        lw = ctx.lw3_all(0.3, "fp64")
        ref = O.lw3_gauss_all(inp, qm.irr_idx, 0.3, threads=8, pol=_pol(3, 6, rules))
        assert rel_err(lw, ref, 1e-9) < 1e-10, rules
    ctx.set_branch_rules(3, "DDD")
    assert np.array_equal(ctx.lw3_all(0.3, "fp64"), base)
    ctx.close()


def test_lw3_gauss_rules_10atom(rocksalt10):
    from spring_amd.lib import GpuContext
    from oracle import oracle as O
    qm, ph = problem(rocksalt10, (2, 2, 2))
    ctx = GpuContext(rocksalt10, qm, ph)
    inp = O.Inputs(rocksalt10, qm, ph)
    for rules in ("AOD", "OAA"):
        ctx.set_branch_rules(3, rules)
        ref = O.lw3_gauss_all(inp, qm.irr_idx, 0.1, threads=8, pol=_pol(3, 30, rules))
        assert rel_err(ctx.lw3_all(0.1, "fp64"), ref, 1e-9) < 1e-10, rules
        err = rel_err(ctx.lw3_all(0.1, "fp32"), ref, 1e-6)
        assert err < 1e-4, (rules, err)
        # the per-pair terms follow the same rule
        q0 = np.array([1, 1, 2], np.int32)
        q1 = np.array([0, 3, 5], np.int32)
        terms = ctx.lw3_pairs(q0, q1, 0.1, "fp64")
        assert np.all(terms[:, 3:] == 0.0) if rules[0] == "A" else np.all(terms[:, :3] == 0.0)
    ctx.close()


def test_lw3_tetra_rules(rocksalt, rocksalt10):
    from spring_amd.lib import GpuContext
    from oracle import oracle as O
    for m, dims, nd in ((rocksalt, (4, 4, 4), 6), (rocksalt10, (2, 2, 2), 30)):
        qm, ph = problem(m, dims)
        ctx = GpuContext(m, qm, ph)
        inp = O.Inputs(m, qm, ph)
        for rules in ("AOD", "DOA"):
            ctx.set_branch_rules(3, rules)
            lw_all, p3_all = ctx.lw3_tetra_all()
            for qi in range(min(qm.nq0, 4)):
                ig = int(qm.irr_idx[qi])
                _, W = O.lw3_gauss(inp, ig, 1.0, want_w=True)
                lw, p3 = O.lw3_tetra(qm, ph, ig, W, pol=_pol(3, nd, rules))
                assert rel_err(lw_all[qi], lw, 1e-9) < 1e-10, (nd, rules, qi)
                assert rel_err(p3_all[qi], p3, 1e-9) < 1e-10, (nd, rules, qi)
        ctx.close()


def test_lw4_rules(rocksalt):
    from spring_amd.lib import GpuContext
    from oracle import oracle as O
    qm, ph = problem(rocksalt, (3, 3, 3))
    ctx = GpuContext(rocksalt, qm, ph)
    inp = O.Inputs(rocksalt, qm, ph, which=4)
    for rules in ("AODD", "DAOA"):
        pol = _pol(4, 6, rules)
        ctx.set_branch_rules(4, rules)
        ref = O.lw4_gauss_all(inp, qm.irr_idx, 2.0, threads=8, pol=pol)
        assert rel_err(ctx.lw4_all(2.0, "fp64"), ref, 1e-9) < 1e-10, rules
        assert rel_err(ctx.lw4_all(2.0, "fp32"), ref, 1e-6) < 1e-4, rules
        tref = np.array([O.lw4_tetra(inp, qm, int(qm.irr_idx[i]), threads=8, chunk=4, pol=pol) for i in (0, 1)])
        tl = np.array([ctx.lw4_tetra(i) for i in (0, 1)])
        assert rel_err(tl, tref, 1e-9) < 1e-10, rules
    ctx.close()


def test_kappa_mode_range_and_spectral(rocksalt):
    from spring_amd import model
    from spring_amd.lib import GpuContext
    from oracle import oracle as O
    m = model.rocksalt2()                   # no FC4: the range is the three-phonon rule's
    qm, ph = problem(m, (4, 4, 4))
    ctx = GpuContext(m, qm, ph)
    lw = np.random.default_rng(4).uniform(0.5, 1.5, (qm.nq0, 6))
    vol = m.crystal.vol
    for symm, f in ((True, O.kappa_rta_symm), (False, O.kappa_rta)):
        k_all, _ = ctx.kappa(lw, 300.0, symm=symm)
        ctx.set_branch_rules(3, "ODD")
        k_o, _ = ctx.kappa(lw, 300.0, symm=symm)
        close = lambda a, b: np.allclose(a, b, rtol=1e-12, atol=1e-14 * np.abs(b).max())
        assert close(k_o, f(qm, ph, lw, 300.0, vol, srange=[3, 5]))
        ctx.set_branch_rules(3, "ADD")
        k_a, _, spec = ctx.kappa_spectral(lw, 300.0, symm=symm)
        assert close(k_a, f(qm, ph, lw, 300.0, vol, srange=[0, 2]))
        assert np.allclose(k_a + k_o, k_all, rtol=1e-12, atol=1e-14 * np.abs(k_all).max())
        # Spectral_k: (s, omega of the irreducible q0, nine terms); summed / Nq it is kappa
        assert spec.shape == (6 * qm.nq0, 11)
        s_col = np.repeat(np.arange(1, 7), qm.nq0)
        assert np.array_equal(spec[:, 0], s_col)
        assert np.array_equal(spec[:, 1], ph.omega[qm.irr_idx].T.ravel())
        assert np.allclose(spec[:, 2:].sum(0) / qm.nq, k_a.ravel(), rtol=1e-12, atol=1e-14 * np.abs(k_a).max())
        assert np.all(spec[qm.nq0 * 3:, 2:] == 0.0)          # optical modes outside the rule
        ctx.set_branch_rules(3, "DDD")
    ctx.close()
