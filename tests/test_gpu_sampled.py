"""spr_lw3_sampled (the importance-sampled linewidth estimate: weighted sums of q1 terms per
irreducible q0, the samples grouped into 64-pair items that run on the full-sum kernels with the
weight applied per pair) against the per-pair terms of spr_lw3_pairs and the full sums."""
import numpy as np
import pytest

from tests.conftest import problem, rel_err

pytestmark = pytest.mark.gpu


def _normwise(a, b):
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


def test_sampled_unit_weights_equal_the_full_sum_2atom(rocksalt):
    """Every q1 of every irreducible q0 with weight 1 (the ordered terms t(q1)) is the full
    linewidth: 2-atom 4^3 (small-cell kernel), FP64 and FP32; shuffled sample order."""
    from spring_amd.lib import GpuContext
    qm, ph = problem(rocksalt, (4, 4, 4))
    ctx = GpuContext(rocksalt, qm, ph)
    q0 = np.repeat(np.arange(qm.nq0), qm.nq)
    q1 = np.tile(np.arange(qm.nq), qm.nq0)
    perm = np.random.default_rng(1).permutation(q0.size)
    for prec, tol in (("fp64", 1e-10), ("fp32", 1e-5)):
        # FP64 per mode (each ordering of a pair is its own sample here, where the full sum
        # evaluates the pair once, so the two agree to rounding); FP32 normwise
        full = ctx.lw3_all(0.1, prec)
        est = ctx.lw3_sampled(q0[perm], q1[perm], np.ones(q0.size), 0.1, prec)
        e = rel_err(est, full, 1e-9) if prec == "fp64" else _normwise(est, full)
        print(f"2-atom 4^3 {prec}: unit-weight sampled sum vs full sum {e:.2e}", flush=True)
        assert e < tol, prec
    ctx.close()


def test_sampled_weights_10atom(rocksalt10):
    """10-atom 3^3 with time reversal: random (q0, q1) samples, some q0 with none, random
    weights.  FP64 (per-atom matrix-core kernel) equals the weighted per-pair terms to 1e-12; FP32
    and split BF16 run on the s2-per-wave kernel and agree with the FP64 sums within FP32 / BF16
    rounding; unit weights over every q1 give the full FP64 sum; the q0 without samples are zero."""
    from spring_amd.lib import GpuContext, SprError
    qm, ph = problem(rocksalt10, (3, 3, 3), time_rev=1)
    ctx = GpuContext(rocksalt10, qm, ph)
    rng = np.random.default_rng(7)
    keep = rng.permutation(qm.nq0)[: qm.nq0 - 3]
    n = 700
    q0 = rng.choice(keep, size=n)
    q1 = rng.integers(0, qm.nq, size=n)
    w = rng.uniform(0.2, 3.0, size=n)
    terms = ctx.lw3_pairs(q0, q1, 0.1, "fp64")
    ref = np.zeros((qm.nq0, ctx.nd))
    np.add.at(ref, q0, terms * w[:, None])
    est = ctx.lw3_sampled(q0, q1, w, 0.1, "fp64")
    assert ctx.last_launch()[0].startswith("mm_f64")
    e = rel_err(est, ref, 1e-9)
    print(f"10-atom FP64 weighted sampled sums vs per-pair terms {e:.2e}", flush=True)
    assert e < 1e-12
    missing = np.setdiff1d(np.arange(qm.nq0), keep)
    assert np.all(est[missing] == 0.0) and np.abs(est).max() > 0
    for prec, tol in (("fp32", 1e-5), ("bf16", 1e-3)):
        e32 = ctx.lw3_sampled(q0, q1, w, 0.1, prec)
        assert ctx.last_launch()[0].startswith("mx_"), ctx.last_launch()
        en = _normwise(e32, est)
        print(f"10-atom {prec} weighted sampled sums vs FP64 {en:.2e}", flush=True)
        assert en < tol, prec
    qa = np.repeat(np.arange(qm.nq0), qm.nq)
    qb = np.tile(np.arange(qm.nq), qm.nq0)
    full = ctx.lw3_all(0.1, "fp64")
    assert rel_err(ctx.lw3_sampled(qa, qb, np.ones(qa.size), 0.1, "fp64"), full, 1e-9) < 1e-10
    with pytest.raises(SprError):
        ctx.lw3_sampled(q0[:3], np.array([0, 1, qm.nq]), w[:3], 0.1, "fp64")
    with pytest.raises(SprError):
        ctx.lw3_sampled(q0[:3], q1[:3], np.array([1.0, np.nan, 1.0]), 0.1, "fp64")
    ctx.close()
