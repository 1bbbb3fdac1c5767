"""GPU parity of the four-phonon tetrahedron path (spr_lw4_tetra, TetrahedronIntegration4th)
against the FP64 CPU oracle, which is pinned in test_lw4_tetra_oracle.py.  Tolerance 1e-10
relative (FP64 with a different summation order)."""
import numpy as np
import pytest

from tests.conftest import problem, rel_err

pytestmark = pytest.mark.gpu


def _ref(rocksalt, qm, ph, sel):
    from oracle import oracle as O
    inp = O.Inputs(rocksalt, qm, ph, which=4)
    return np.array([O.lw4_tetra(inp, qm, int(qm.irr_idx[i]), threads=8, chunk=4) for i in sel])


def test_lw4_tetra_parity_3(rocksalt):
    from spring_amd.lib import GpuContext
    qm, ph = problem(rocksalt, (3, 3, 3))
    ctx = GpuContext(rocksalt, qm, ph)
    print("lw4_tetra 3^3 all q0", flush=True)
    lw = ctx.lw4_tetra_all()
    ref = _ref(rocksalt, qm, ph, range(qm.nq0))
    err = rel_err(lw, ref, 1e-9)
    print("lw4_tetra 3^3 max rel", err, flush=True)
    assert err < 1e-10
    assert np.all(lw[0, :3] == 0.0) and lw.max() > 0.0
    assert np.allclose(ctx.lw4_tetra(1), lw[1], rtol=1e-12, atol=0)


def test_lw4_tetra_parity_4_chunked(rocksalt, monkeypatch):
    """Even mesh (zone-boundary points), W4 held a few q1 rows at a time (several chunks)."""
    from spring_amd.lib import GpuContext
    qm, ph = problem(rocksalt, (4, 4, 4))
    ctx = GpuContext(rocksalt, qm, ph)
    sel = [0, 2, qm.nq0 - 1]
    print("lw4_tetra 4^3 whole", flush=True)
    whole = np.array([ctx.lw4_tetra(i) for i in sel])
    monkeypatch.setenv("SPR_W4_CHUNK_MB", "2")      # 64 q x 1296 x 8 B = 0.66 MB a row: 3 rows
    print("lw4_tetra 4^3 chunked", flush=True)
    chunked = np.array([ctx.lw4_tetra(i) for i in sel])
    ref = _ref(rocksalt, qm, ph, sel)
    e1, e2 = rel_err(whole, ref, 1e-9), rel_err(chunked, ref, 1e-9)
    print("lw4_tetra 4^3 max rel whole", e1, "chunked", e2, flush=True)
    assert e1 < 1e-10 and e2 < 1e-10


def test_lw4_tetra_needs_fc4():
    from spring_amd import model
    from spring_amd.lib import GpuContext, SprError
    m = model.rocksalt2()
    qm, ph = problem(m, (3, 3, 3))
    ctx = GpuContext(m, qm, ph)
    with pytest.raises(SprError, match="no FC4"):
        ctx.lw4_tetra(0)


def test_lw4_tetra_full_size_8(rocksalt):
    """The four-phonon tetrahedron workload at its full size (2-atom with FC4, 8^3 mesh): one seeded
    irreducible q0 of the all-q0 call against the oracle over all 512 q1 rows (16 threads)."""
    from spring_amd.lib import GpuContext
    from oracle import oracle as O
    qm, ph = problem(rocksalt, (8, 8, 8))
    ctx = GpuContext(rocksalt, qm, ph)
    print("lw4_tetra 8^3: all q0 on the GPU", flush=True)
    lw = ctx.lw4_tetra_all()
    assert np.all(np.isfinite(lw)) and np.all(lw.sum(axis=1) > 0.0)
    qi = int(np.random.default_rng(8).integers(1, qm.nq0))
    print(f"lw4_tetra 8^3: oracle q0 #{qi}", flush=True)
    ref = O.lw4_tetra(O.Inputs(rocksalt, qm, ph, which=4), qm, int(qm.irr_idx[qi]), threads=16, chunk=8)
    err = rel_err(lw[qi], ref, 1e-9)
    print(f"lw4_tetra 8^3 q0 #{qi}: {err:.2e}", flush=True)
    assert err < 1e-10
