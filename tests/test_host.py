"""Host-side pieces (no GPU): model generator, mesh/symmetry, phonons, library exports."""
import re
import os

import numpy as np
import pytest

from spring_amd import constants as C, mesh, model, phonons
from tests.conftest import ROOT, problem


def test_fc3_permutation_symmetry_and_asr(rocksalt):
    f3 = rocksalt.fc3.fc
    assert np.abs(f3 - np.transpose(f3, (0, 2, 1, 3, 5, 4))).max() < 1e-10 * np.abs(f3).max()
    assert np.abs(f3.sum(axis=2)).max() < 1e-12 * np.abs(f3).max()
    f4 = rocksalt.fc4.fc
    assert np.abs(f4 - np.transpose(f4, (0, 1, 3, 2, 4, 5, 7, 6))).max() < 1e-10 * np.abs(f4).max()
    assert np.abs(f4 - np.transpose(f4, (0, 2, 1, 3, 4, 6, 5, 7))).max() < 1e-10 * np.abs(f4).max()
    assert np.abs(f4.sum(axis=3)).max() < 1e-12 * np.abs(f4).max()


def test_fc3_site_permutation_full(rocksalt):
    """Phi(mu a, N1 b, N2 c) = Phi(N1 b, mu a, N2 c) after translating N1 home."""
    f = rocksalt.fc3
    nb = rocksalt.crystal.nb
    lookup = {}
    for mu in range(nb):
        for n in range(f.atm_num[mu]):
            lookup[(mu, tuple(f.atm_indx[mu, n, :3]), int(f.atm_indx[mu, n, 3]))] = n
    worst = 0.0
    for mu in range(nb):
        for n1 in range(f.atm_num[mu]):
            c1, nu = tuple(f.atm_indx[mu, n1, :3]), int(f.atm_indx[mu, n1, 3])
            for n2 in range(f.atm_num[mu]):
                blk = f.fc[mu, n1, n2]
                if not blk.any():
                    continue
                c2, eta = tuple(f.atm_indx[mu, n2, :3]), int(f.atm_indx[mu, n2, 3])
                a = lookup[(nu, tuple(-np.array(c1)), mu)]
                b = lookup[(nu, tuple(np.array(c2) - np.array(c1)), eta)]
                worst = max(worst, np.abs(f.fc[nu, a, b] - np.transpose(blk, (1, 0, 2))).max())
    assert worst < 1e-10 * np.abs(f.fc).max()


def test_irreducible_counts():
    cr = model.rocksalt_primitive()
    for n, expect in ((8, 29), (16, 145), (10, 47)):
        qm = mesh.make_mesh(cr, (n, n, n))
        assert qm.nq0 == expect
        assert qm.multiplicity.sum() == n ** 3
        assert len(qm.rot_frac) == 48
    c10 = model.rocksalt_supercell10()
    assert mesh.make_mesh(c10, (20, 20, 20)).nq0 == 8000
    assert mesh.make_mesh(c10, (20, 20, 20), time_rev=1).nq0 == 4004


def test_pairs_brute_force_and_even_mesh_count():
    cr = model.rocksalt_primitive()
    for n in (5, 8):
        qm = mesh.make_mesh(cr, (n, n, n))
        for i0 in qm.irr_idx[:6]:
            q2 = mesh.pair_partner(qm, int(i0))
            s = qm.q_int[i0] + qm.q_int + qm.q_int[q2]
            assert np.all(np.mod(s, n) == 0)
            q1s, q2s, w = mesh.pair_list(qm, int(i0))
            assert w.sum() == qm.nq
            nfix = int(np.sum(q2 == np.arange(qm.nq)))
            assert len(q1s) == (qm.nq + nfix) // 2
    # Gamma on the 8^3 mesh: 260 pairs, not the Nq/2+1 = 257 the reference asserts
    qm = mesh.make_mesh(cr, (8, 8, 8))
    assert len(mesh.pair_list(qm, 0)[0]) == 260


def test_tetrahedra_cover_mesh():
    qm = mesh.make_mesh(model.rocksalt_primitive(), (4, 4, 4))
    assert qm.tetra.shape == (6 * 64, 4)
    assert np.array_equal(np.bincount(qm.tetra.ravel(), minlength=64), np.full(64, 24))


def test_phonons_basic(rocksalt):
    qm, ph = problem(rocksalt, (4, 4, 4))
    assert np.all(np.abs(ph.omega[0, :3]) < 1e-5)
    assert np.all(ph.inv_omega[0, :3] == 0) and np.all(ph.nbe[0, :3] == 0)
    D, _ = phonons.dynamical_matrix(rocksalt.crystal, rocksalt.fc2, qm.q_cart)
    assert np.abs(D - np.conj(np.transpose(D, (0, 2, 1)))).max() < 1e-12 * np.abs(D).max()
    V = np.transpose(ph.evec, (0, 2, 1))
    lam = (ph.omega / C.W_THZ) ** 2
    res = np.einsum('qij,qjs->qis', D, V) - V * lam[:, None, :]
    assert np.abs(res).max() < 1e-10 * np.abs(D).max()
    eye = np.einsum('qis,qit->qst', V.conj(), V)
    assert np.abs(eye - np.eye(6)).max() < 1e-12


def test_gamma_optical_closed_form():
    """Nearest-neighbour springs only: w_TO(Gamma)^2 = (2 kL + 4 kT)(1/m1 + 1/m2)."""
    cr = model.rocksalt_primitive()
    fc2 = model.spring_fc2(cr, [(5.64 / 2, 1.5, 0.2)], seed=3)
    kl, kt = 1.5 * fc2.meta["factors"][0, 0], 0.2 * fc2.meta["factors"][0, 1]
    ph = phonons.solve(cr, fc2, np.zeros((1, 3)), np.zeros((1, 3), int), 300.0)
    w = np.sqrt((2 * kl + 4 * kt) * (1 / cr.mass[0] + 1 / cr.mass[1])) * C.W_THZ
    assert np.allclose(ph.omega[0, 3:], w, rtol=1e-12)


def test_group_velocity_finite_difference(rocksalt):
    q = np.array([[0.11, 0.07, -0.05]])
    ph = phonons.solve(rocksalt.crystal, rocksalt.fc2, q, np.ones((1, 3), int), 300.0)
    h = 1e-6
    for k in range(3):
        dq = np.zeros((1, 3))
        dq[0, k] = h
        p = phonons.solve(rocksalt.crystal, rocksalt.fc2, q + dq, np.ones((1, 3), int), 300.0)
        m = phonons.solve(rocksalt.crystal, rocksalt.fc2, q - dq, np.ones((1, 3), int), 300.0)
        fd = (p.omega - m.omega) / (2 * h) / C.W_THZ * C.VG_MS   # d(w)/dq in m/s
        assert np.allclose(ph.grp_vel[0, :, k], fd[0], rtol=1e-5, atol=1e-3)


def test_library_exports_every_declared_symbol():
    hdr = open(os.path.join(ROOT, "include", "spring_gpu.h")).read()
    declared = set(re.findall(r"\b(spr_[a-z0-9_]+)\s*\(", hdr))
    from spring_amd import lib
    L = lib.load()
    for name in sorted(declared):
        assert hasattr(L, name), name
    assert declared <= set(lib.SYMBOLS)


def test_get_phonon_rejects_bad_arguments_on_the_host():
    """Argument checks of spr_get_phonon run before any device call (no GPU needed)."""
    from spring_amd.lib import SprError, get_phonon
    m = model.rocksalt_chain(8)                                 # Ndof = 48 > 32
    with pytest.raises(SprError, match="Ndof"):
        get_phonon(m.crystal, m.fc2, np.zeros((1, 3)), np.zeros((1, 3), int), 300.0)
    m = model.rocksalt2()
    with pytest.raises(SprError, match="bad argument"):
        get_phonon(m.crystal, m.fc2, np.zeros((1, 3)), np.zeros((1, 3), int), -1.0)


def test_lw4_tetra_rejects_a_null_context_on_the_host():
    import ctypes as ct
    from spring_amd import lib
    L = lib.load()
    out = np.zeros(6)
    assert L.spr_lw4_tetra(None, ct.c_int(0), out.ctypes.data_as(ct.c_void_p)) < 0
    assert b"bad argument" in L.spr_last_error()
    assert L.spr_lw4_tetra_all(None, out.ctypes.data_as(ct.c_void_p)) < 0


def test_head_tail_draw_is_unbiased():
    """Deterministic head + sampled tail: the weighted sum is an unbiased estimate of the full sum
    (seeded Monte Carlo mean within 4 standard errors), and a budget covering every q1 is exact."""
    from spring_amd import learned as L
    rng = np.random.default_rng(0)
    S = rng.exponential(size=400)
    t = S * rng.uniform(0.5, 1.5, size=400)
    est = []
    for _ in range(3000):
        q, w = L.head_tail_draw(S, 40, 0.5, rng)
        est.append((t[q] * w).sum())
    est = np.array(est)
    assert abs(est.mean() - t.sum()) < 4 * est.std() / np.sqrt(est.size)
    q, w = L.head_tail_draw(S, 400, 1.0, rng)
    assert np.isclose((t[q] * w).sum(), t.sum(), rtol=1e-12)

