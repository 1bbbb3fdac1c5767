"""spr_lw3_pairs_noalloc (the per-pair terms with no device allocation) against spr_lw3_pairs, the
entry point it is a variant of: the same kernels on the same pairs, so the terms must be bit for
bit equal.  Two calls in a row on each context, the second with fewer pairs, so that a stale
scratch region or a wrong offset into it shows; both through the C ABI."""
import numpy as np
import pytest

from tests.conftest import problem

pytestmark = pytest.mark.gpu


def _pairs(qm, n, seed):
    rng = np.random.default_rng(seed)
    q0 = rng.integers(0, qm.nq0, n).astype(np.int32)
    q1 = rng.integers(0, qm.nq, n).astype(np.int32)
    return q0, q1


@pytest.mark.parametrize("which, dims, precs", [
    ("rocksalt10", (3, 3, 3), ("fp64", "fp32", "bf16")),   # matrix-core kernel
    ("rocksalt", (4, 4, 4), ("fp64", "fp32")),             # small-cell kernel
])
def test_noalloc_pairs_equal_the_allocating_entry(request, which, dims, precs):
    from spring_amd.lib import GpuContext
    m = request.getfixturevalue(which)
    qm, ph = problem(m, dims)
    ctx = GpuContext(m, qm, ph)
    for prec in precs:
        for n, seed in ((700, 1), (93, 2)):
            q0, q1 = _pairs(qm, n, seed)
            ref = ctx.lw3_pairs(q0, q1, 0.1, prec)
            got = ctx.lw3_pairs_noalloc(q0, q1, 0.1, prec)
            assert got.shape == (n, m.crystal.ndof)
            assert np.abs(ref).max() > 0
            assert np.array_equal(got, ref), (which, prec, n)
    ctx.close()


def test_noalloc_pairs_limit(rocksalt):
    from spring_amd.lib import GpuContext, SprError
    qm, ph = problem(rocksalt, (4, 4, 4))
    ctx = GpuContext(rocksalt, qm, ph)
    q0, q1 = _pairs(qm, 4097, 3)
    with pytest.raises(SprError, match="scratch limit of 4096"):
        ctx.lw3_pairs_noalloc(q0, q1, 0.1, "fp64")
    with pytest.raises(SprError, match="sigma"):
        ctx.lw3_pairs_noalloc(q0[:4], q1[:4], 0.0, "fp64")
    q0, q1 = _pairs(qm, 4096, 4)
    out = ctx.lw3_pairs_noalloc(q0, q1, 0.1, "fp64")
    assert np.array_equal(out, ctx.lw3_pairs(q0, q1, 0.1, "fp64"))
    ctx.close()
