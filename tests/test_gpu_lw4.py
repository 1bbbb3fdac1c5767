"""GPU parity of the four-phonon Gaussian path against the FP64 CPU oracle."""
import numpy as np
import pytest

from tests.conftest import problem, rel_err

pytestmark = pytest.mark.gpu


def test_lw4_parity_fp64_and_fp32(rocksalt):
    from spring_amd.lib import GpuContext
    from oracle import oracle as O
    qm, ph = problem(rocksalt, (3, 3, 3))
    ctx = GpuContext(rocksalt, qm, ph)
    inp = O.Inputs(rocksalt, qm, ph, which=4)
    lw = ctx.lw4_all(0.1, "fp64")
    ref = O.lw4_gauss_all(inp, qm.irr_idx, 0.1, threads=8)
    print("lw4 max rel", rel_err(lw, ref, 1e-9), flush=True)
    assert rel_err(lw, ref, 1e-9) < 1e-10
    lw32 = ctx.lw4_all(0.1, "fp32")
    assert rel_err(lw32, ref, 1e-6) < 1e-4
    # single-q0 entry point agrees with the batched one
    assert np.allclose(ctx.lw4(1, 0.1, "fp64"), lw[1], rtol=1e-12, atol=0)


def test_lw4_even_mesh(rocksalt):
    from spring_amd.lib import GpuContext
    from oracle import oracle as O
    qm, ph = problem(rocksalt, (4, 4, 4))
    ctx = GpuContext(rocksalt, qm, ph)
    inp = O.Inputs(rocksalt, qm, ph, which=4)
    sel = [0, 3, qm.nq0 - 1]
    lw = ctx.lw4_all(0.1, "fp64")[sel]
    ref = np.array([O.lw4_gauss(inp, int(qm.irr_idx[i]), 0.1) for i in sel])
    assert rel_err(lw, ref, 1e-9) < 1e-10


def test_lw4_fp32_matrix_cores_vs_fp64_and_vector(rocksalt, monkeypatch):
    """FP32 runs on the matrix-core kernel; it agrees with the FP64 path and with the FP32
    vector kernel (SPR_KERNEL=generic) within the FP32 tolerance."""
    from spring_amd.lib import GpuContext
    qm, ph = problem(rocksalt, (4, 4, 4))
    ctx = GpuContext(rocksalt, qm, ph)
    lw64 = ctx.lw4_all(0.1, "fp64")
    lw32 = ctx.lw4_all(0.1, "fp32")
    monkeypatch.setenv("SPR_KERNEL", "generic")
    lw32v = ctx.lw4_all(0.1, "fp32")
    e64, ev = rel_err(lw32, lw64, 1e-6), rel_err(lw32, lw32v, 1e-6)
    print("lw4 fp32 matrix cores: vs fp64", e64, "vs fp32 vector", ev, flush=True)
    assert e64 < 1e-4 and ev < 1e-4
    assert np.all(np.isfinite(lw32))


def test_lw4_full_size_10(rocksalt):
    """The four-phonon Gaussian workload at its full size (2-atom with FC4, 10^3 mesh, sigma 0.1):
    one seeded irreducible q0 of the all-q0 call in FP64 against the oracle over all 10^6 (q1, q2),
    threaded over q1 chunks."""
    from concurrent.futures import ThreadPoolExecutor
    from spring_amd.lib import GpuContext
    from oracle import oracle as O
    qm, ph = problem(rocksalt, (10, 10, 10))
    ctx = GpuContext(rocksalt, qm, ph)
    print("lw4 10^3: all q0 on the GPU (FP64)", flush=True)
    lw = ctx.lw4_all(0.1, "fp64")
    qi = int(np.random.default_rng(10).integers(1, qm.nq0))
    ig = int(qm.irr_idx[qi])
    inp = O.Inputs(rocksalt, qm, ph, which=4)
    print(f"lw4 10^3: oracle q0 #{qi}", flush=True)
    spans = [(lo, min(qm.nq, lo + 25)) for lo in range(0, qm.nq, 25)]
    with ThreadPoolExecutor(16) as ex:
        parts = list(ex.map(lambda sp: O.lw4_gauss(inp, ig, 0.1, sp[0], sp[1]), spans))
    ref = np.sum(parts, axis=0)
    err = rel_err(lw[qi], ref, 1e-9)
    print(f"lw4 10^3 q0 #{qi}: {err:.2e}", flush=True)
    assert err < 1e-10
    # the FP32 matrix-core kernel (k_lw4_mm) on every irreducible q0 of the same workload against
    # the FP64 linewidths just checked; and kappa of the summed three- and four-phonon linewidths
    lw32 = ctx.lw4_all(0.1, "fp32")
    e32 = rel_err(lw32, lw, 1e-6)
    lw3 = ctx.lw3_all(0.1, "fp64")
    k64, _ = ctx.kappa(lw3 + 0.5 * lw, 300.0, symm=True)
    k32, _ = ctx.kappa(lw3 + 0.5 * lw32, 300.0, symm=True)
    ek = float(np.max(np.abs(k32 - k64)) / np.max(np.abs(k64)))
    print(f"lw4 10^3 FP32 vs FP64: linewidths {e32:.2e}, kappa {ek:.2e}", flush=True)
    assert e32 < 1e-4 and ek < 1e-6
