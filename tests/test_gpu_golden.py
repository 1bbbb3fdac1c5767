"""GPU kernels against the committed golden fixture (literal-loop results)."""
import numpy as np
import pytest

from tests.golden_fixture import load

pytestmark = pytest.mark.gpu


def test_gpu_reproduces_golden(monkeypatch):
    from spring_amd.lib import GpuContext
    m, qm, ph, d = load()
    ctx = GpuContext(m, qm, ph)
    sig = float(d["sigma"])
    for kern in ("small", "generic"):
        monkeypatch.setenv("SPR_KERNEL", kern)
        lw = ctx.lw3_all(sig, "fp64")
        assert np.allclose(lw, d["lw"], rtol=1e-10, atol=1e-12 * np.abs(d["lw"]).max()), kern
    monkeypatch.delenv("SPR_KERNEL")
    W = ctx.v3_w(1)
    assert np.allclose(W, d["W_q0_1"], rtol=1e-10, atol=1e-12 * np.abs(W).max())
