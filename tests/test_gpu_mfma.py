"""Three-phonon MFMA kernel (FP64, FP32 and split BF16) against the FP64 oracle and against
the other kernels, for cells of 2, 4, 6 and 10 atoms."""
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


def _err(lw, ref):
    s = _sig(ref)
    return np.abs(lw - ref)[s] / np.abs(ref)[s]


def test_mfma_fp64_10atom(ten):
    qm, ph, ctx, ref = ten
    lw = ctx.lw3_all(0.1, "fp64")
    assert rel_err(lw, ref, 1e-9) < 1e-10


def test_mfma_fp32_10atom(ten):
    qm, ph, ctx, ref = ten
    err = _err(ctx.lw3_all(0.1, "fp32"), ref)
    print("fp32 mfma max rel", err.max(), flush=True)
    assert err.max() < 1e-3


def test_mfma_bf16_10atom(ten):
    """Split BF16 operands (hi + lo, three products) keep about 16 mantissa bits."""
    qm, ph, ctx, ref = ten
    lw = ctx.lw3_all(0.1, "bf16")
    err = _err(lw, ref)
    print("bf16 mfma max/mean rel", err.max(), err.mean(), flush=True)
    assert err.max() < 1e-3 and err.mean() < 1e-4


def test_mfma_fp64_10atom_odd_mesh_vs_generic(rocksalt10, monkeypatch):
    """3x3x3 mesh with time reversal (14 irreducible q0, q1 == q2 pairs present): the MFMA
    FP64 kernel and the generic FP64 kernel agree to rounding."""
    from spring_amd.lib import GpuContext
    from spring_amd import mesh, phonons
    qm = mesh.make_mesh(rocksalt10.crystal, (3, 3, 3), time_rev=1)
    ph = phonons.solve(rocksalt10.crystal, rocksalt10.fc2, qm.q_cart, qm.q_int, 300.0)
    ctx = GpuContext(rocksalt10, qm, ph)
    a = ctx.lw3_all(0.1, "fp64")
    monkeypatch.setenv("SPR_KERNEL", "generic")
    b = ctx.lw3_all(0.1, "fp64")
    monkeypatch.delenv("SPR_KERNEL")
    assert rel_err(a, b, 1e-9) < 1e-12


@pytest.mark.parametrize("n", [2, 3])
def test_mfma_intermediate_cells(n):
    """4- and 6-atom cells (Ndof 12, 18: other tile shapes and s2 splits) against the oracle."""
    from spring_amd import model
    from spring_amd.lib import GpuContext
    from oracle import oracle as O
    m = model.rocksalt_chain(n)
    qm, ph = problem(m, (3, 3, 3))
    ctx = GpuContext(m, qm, ph)
    ref = O.lw3_gauss_all(O.Inputs(m, qm, ph), qm.irr_idx, 0.1, threads=8)
    assert rel_err(ctx.lw3_all(0.1, "fp64"), ref, 1e-9) < 1e-10
    assert _err(ctx.lw3_all(0.1, "fp32"), ref).max() < 1e-3
    assert _err(ctx.lw3_all(0.1, "bf16"), ref).max() < 1e-3


def test_mfma_2atom_matches_small_kernel(rocksalt, monkeypatch):
    from spring_amd.lib import GpuContext
    qm, ph = problem(rocksalt, (4, 4, 4))
    ctx = GpuContext(rocksalt, qm, ph)
    ref = ctx.lw3_all(0.1, "fp64")
    monkeypatch.setenv("SPR_KERNEL", "mfma")
    lw64 = ctx.lw3_all(0.1, "fp64")
    lw = ctx.lw3_all(0.1, "fp32")
    bf = ctx.lw3_all(0.1, "bf16")
    monkeypatch.delenv("SPR_KERNEL")
    assert rel_err(lw64, ref, 1e-9) < 1e-12
    assert _err(lw, ref).max() < 1e-3
    assert _err(bf, ref).max() < 1e-3


def test_mfma_pair_terms_fp64_10atom(ten):
    """spr_lw3_pairs on the MFMA kernel: per-pair terms of every q1 sum to the linewidth."""
    qm, ph, ctx, ref = ten
    from spring_amd import mesh
    i = 3
    q1 = np.arange(qm.nq)
    t = ctx.lw3_pairs(np.full(qm.nq, i), q1, 0.1, "fp64")
    assert rel_err(t.sum(axis=0), ref[i], 1e-9) < 1e-10


def test_launch_modes_and_tile_shapes_at_a_tuned_size(rocksalt10, monkeypatch, tmp_path):
    """On a mesh large enough that the launcher tunes (10-atom cell, 8^3 with time reversal: 1040
    items of 64 pairs), every forced launch mode (one workgroup per item, persistent grids of 1, 2
    and 4 workgroups per CU) and the direct launch without a captured graph give bit-identical
    linewidths: each item writes its own partial row and the rows are summed in item order.  The
    two FP64 tile shapes (s2 blocks of 8 and 16) cut the same sums differently and agree to
    rounding."""
    from spring_amd import lib, mesh
    monkeypatch.setenv("SPR_TUNE_CACHE", str(tmp_path / "tune.txt"))
    m = rocksalt10
    qm = mesh.make_mesh(m.crystal, (8, 8, 8), time_rev=1)
    ph = lib.get_phonon(m.crystal, m.fc2, qm.q_cart, qm.q_int, 300.0)
    ctx = lib.GpuContext(m, qm, ph)
    for prec in ("fp64", "fp32"):
        base = ctx.lw3_all(0.1, prec)          # tunes (shape and mode) on first use
        assert np.all(np.isfinite(base)) and np.abs(base).max() > 0
        for mode in ("0", "1", "2", "4"):
            monkeypatch.setenv("SPR_LAUNCH_MODE", mode)
            assert np.array_equal(ctx.lw3_all(0.1, prec), base), (prec, mode)
        monkeypatch.delenv("SPR_LAUNCH_MODE")
        monkeypatch.setenv("SPR_NO_GRAPH", "1")
        assert np.array_equal(ctx.lw3_all(0.1, prec), base), (prec, "no graph")
        monkeypatch.delenv("SPR_NO_GRAPH")
    shapes = {}
    for sh in ("0", "1"):
        monkeypatch.setenv("SPR_TILE_SHAPE", sh)
        shapes[sh] = ctx.lw3_all(0.1, "fp64")
        monkeypatch.setenv("SPR_LAUNCH_MODE", "0")
        assert np.array_equal(ctx.lw3_all(0.1, "fp64"), shapes[sh]), sh
        monkeypatch.delenv("SPR_LAUNCH_MODE")
    assert rel_err(shapes["1"], shapes["0"], 1e-9) < 1e-13
    tune = (tmp_path / "tune.txt").read_text()
    assert "mm_f64|shape" in tune and "mm_f32" in tune
    ctx.close()


@pytest.mark.parametrize("prec", ["fp32", "bf16"])
def test_s2_per_wave_kernel_10atom(ten, rocksalt10, monkeypatch, prec):
    """The s2-per-wave FP32 / BF16 kernel (k_lw3_mx, tile-shape candidate 1 of the 10-atom launch,
    one unit per item, the phase sum once per pair) against the oracle at 2^3; against the per-atom
    kernel (candidate 0) at 3^3 with time reversal (q1 == q2 pairs, items of every size) with and
    without a branch selection rule: the two differ only by FP32 rounding."""
    from spring_amd import mesh
    from spring_amd.lib import GpuContext, get_phonon
    qm, ph, ctx, ref = ten
    monkeypatch.setenv("SPR_TILE_SHAPE", "1")
    lw = ctx.lw3_all(0.1, prec)
    name, _, upi = ctx.last_launch()
    assert name.startswith("mx_") and upi == 1, name
    err = _err(lw, ref)
    print(f"{prec} s2-per-wave max/mean rel", err.max(), err.mean(), flush=True)
    assert err.max() < 1e-3
    if prec == "bf16":
        assert err.mean() < 1e-4
    m = rocksalt10
    q3 = mesh.make_mesh(m.crystal, (3, 3, 3), time_rev=1)
    p3 = get_phonon(m.crystal, m.fc2, q3.q_cart, q3.q_int, 300.0)
    c3 = GpuContext(m, q3, p3)
    tol = 2e-5 if prec == "fp32" else 2e-3
    for rules in ("DDD", "ADD"):
        c3.set_branch_rules(3, rules)
        monkeypatch.setenv("SPR_TILE_SHAPE", "1")
        a = c3.lw3_all(0.1, prec)
        assert c3.last_launch()[0].startswith("mx_")
        monkeypatch.setenv("SPR_TILE_SHAPE", "0")
        b = c3.lw3_all(0.1, prec)
        assert c3.last_launch()[0].startswith("mm_")
        assert np.abs(b).max() > 0
        if rules == "ADD":
            assert np.all(a[:, 3:] == 0.0)
        e = _err(a, b)
        print(f"{prec} {rules}: s2-per-wave vs per-atom max rel {e.max():.2e}", flush=True)
        assert e.max() < tol, rules
    c3.close()


def test_fp64_phase_sum_task_lists_match_single_loop(rocksalt10, monkeypatch):
    """FP64 per-atom kernel, 10-atom cell, 3^3 mesh with time reversal (q1 == q2 pairs): the phase
    sum over the task lists (outputs sorted by cluster count, loads issued four clusters at a time)
    against the loop over every block of the atom (SPR_FT_SCHEDULE=0), for both tile shapes, with
    and without a branch selection rule, and in the W-output mode (tetrahedron path): both sum each
    output in cluster order with the same FMAs, so the linewidths are bit-identical."""
    from spring_amd import mesh
    from spring_amd.lib import GpuContext, get_phonon
    m = rocksalt10
    q3 = mesh.make_mesh(m.crystal, (3, 3, 3), time_rev=1)
    p3 = get_phonon(m.crystal, m.fc2, q3.q_cart, q3.q_int, 300.0)
    c3 = GpuContext(m, q3, p3)
    for rules in ("DDD", "ADD"):
        c3.set_branch_rules(3, rules)
        for sh in ("0", "1"):
            monkeypatch.setenv("SPR_TILE_SHAPE", sh)
            monkeypatch.delenv("SPR_FT_SCHEDULE", raising=False)
            a = c3.lw3_all(0.1, "fp64")
            assert c3.last_launch()[0].startswith("mm_f64")
            monkeypatch.setenv("SPR_FT_SCHEDULE", "0")
            b = c3.lw3_all(0.1, "fp64")
            assert np.abs(b).max() > 0
            if rules == "ADD":
                assert np.all(a[:, 3:] == 0.0)
            e = rel_err(a, b, 1e-9)
            print(f"{rules} shape {sh}: task lists vs single loop max rel {e:.2e}", flush=True)
            assert np.array_equal(a, b), (rules, sh)
    monkeypatch.delenv("SPR_TILE_SHAPE")
    c3.set_branch_rules(3, "DDD")
    # the W-output mode (tetrahedron path of a 10-atom cell)
    monkeypatch.delenv("SPR_FT_SCHEDULE", raising=False)
    ta = c3.lw3_tetra_all()
    monkeypatch.setenv("SPR_FT_SCHEDULE", "0")
    tb = c3.lw3_tetra_all()
    for x, y in zip(ta, tb):
        assert np.array_equal(x, y)
    c3.close()
