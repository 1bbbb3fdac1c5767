"""Mixed-precision MFMA three-phonon kernel (FP32 and BF16) against the FP64 oracle."""
import numpy as np
import pytest

from tests.conftest import problem, rel_err

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ten(rocksalt10):
    from spring_amd.lib import GpuContext
    from oracle import oracle as O
    qm, ph = problem(rocksalt10, (2, 2, 2))
    ctx = GpuContext(rocksalt10, qm, ph)
    ref = O.lw3_gauss_all(O.Inputs(rocksalt10, qm, ph), qm.irr_idx, 0.1, threads=8)
    return qm, ph, ctx, ref


def _sig(ref):
    return np.abs(ref) > 1e-6 * np.abs(ref).max()


def test_mfma_fp32_10atom(ten):
    qm, ph, ctx, ref = ten
    lw = ctx.lw3_all(0.1, "fp32")
    err = np.abs(lw - ref)[_sig(ref)] / np.abs(ref)[_sig(ref)]
    print("fp32 mfma max rel", err.max(), flush=True)
    assert err.max() < 1e-3


def test_mfma_bf16_10atom(ten):
    qm, ph, ctx, ref = ten
    lw = ctx.lw3_all(0.1, "bf16")
    s = _sig(ref)
    err = np.abs(lw - ref)[s] / np.abs(ref)[s]
    print("bf16 mfma max/mean rel", err.max(), err.mean(), flush=True)
    assert err.mean() < 2e-2 and err.max() < 2e-1
    # summed over modes (what kappa sees) the BF16 error averages down
    assert abs(lw.sum() - ref.sum()) / ref.sum() < 1e-2


def test_mfma_fp32_2atom_matches_fp64(rocksalt, monkeypatch):
    from spring_amd.lib import GpuContext
    qm, ph = problem(rocksalt, (4, 4, 4))
    ctx = GpuContext(rocksalt, qm, ph)
    ref = ctx.lw3_all(0.1, "fp64")
    monkeypatch.setenv("SPR_KERNEL", "mfma")
    lw = ctx.lw3_all(0.1, "fp32")
    bf = ctx.lw3_all(0.1, "bf16")
    monkeypatch.delenv("SPR_KERNEL")
    s = _sig(ref)
    assert (np.abs(lw - ref)[s] / np.abs(ref)[s]).max() < 1e-3
    assert abs(bf.sum() - ref.sum()) / ref.sum() < 2e-2
