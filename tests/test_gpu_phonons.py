"""GPU phonons (spr_get_phonon) against the oracle's get_phonon restatement.

Eigenvectors are unique only up to a phase (and, for degenerate modes, up to a unitary
rotation inside the degenerate set), for the reference's LAPACK call as for the device
Jacobi solver, so the comparison is on the quantities that do not depend on that choice:
  eigenvalues lambda = sign(omega) (omega/wTHz)^2       to 1e-12 of max |lambda|
  omega, 1/omega, n of every mode off Gamma              to 1e-9 relative
  |D e - lambda e| (D of the oracle) and e^H e = 1        to 1e-12
  the projector onto each degenerate set                 to 1e-9
  sum over each degenerate set of omega v                to 1e-10 of max |omega v|
  Gamma: the three acoustic modes have v = n = 1/omega = 0 exactly.
"""
import numpy as np
import pytest

from oracle import oracle as O
from tests.test_phonon_oracle import degenerate_sets, fcc_monatomic

pytestmark = pytest.mark.gpu


def _check(cr, fc2, q, q_int, T=300.0):
    from spring_amd.lib import get_phonon
    g = get_phonon(cr, fc2, q, q_int, T)
    o = O.phonons(cr, fc2, q, q_int, T)
    nq, nd = g.omega.shape
    lam_g = np.sign(g.omega) * (g.omega / O.PH_WTHZ) ** 2
    lam_o = np.sign(o["omega"]) * (o["omega"] / O.PH_WTHZ) ** 2
    lmax = np.abs(lam_o).max()
    e_lam = np.abs(lam_g - lam_o).max() / lmax
    V = np.transpose(g.evec, (0, 2, 1))                      # V[q][row][s]
    res = np.einsum('qij,qjs->qis', o["D"], V) - V * lam_g[:, None, :]
    e_res = np.abs(res).max() / np.abs(o["D"]).max()
    e_orth = np.abs(np.einsum('qis,qit->qst', V.conj(), V) - np.eye(nd)).max()
    gamma = np.all(np.asarray(q_int) == 0, axis=1)
    sel = np.ones((nq, nd), bool)
    sel[gamma, :3] = False
    e_w = max(np.max(np.abs(g.omega[sel] / o["omega"][sel] - 1.0)),
              np.max(np.abs(g.inv_omega[sel] / o["inv_omega"][sel] - 1.0)),
              np.max(np.abs(g.nbe[sel] / o["nbe"][sel] - 1.0)))
    e_p = e_v = 0.0
    wv_max = np.abs(o["omega"][:, :, None] * o["grp_vel"]).max()
    Vo = np.transpose(o["evec"], (0, 2, 1))
    for i in range(nq):
        for grp in degenerate_sets(o["omega"][i]):
            if gamma[i] and grp[0] < 3:
                continue
            Pg = V[i][:, grp] @ V[i][:, grp].conj().T
            Po = Vo[i][:, grp] @ Vo[i][:, grp].conj().T
            e_p = max(e_p, np.abs(Pg - Po).max())
            a = (g.omega[i, grp, None] * g.grp_vel[i, grp]).sum(0)
            b = (o["omega"][i, grp, None] * o["grp_vel"][i, grp]).sum(0)
            e_v = max(e_v, np.abs(a - b).max() / wv_max)
    print(f"nq={nq} nd={nd}: lambda {e_lam:.1e} residual {e_res:.1e} orth {e_orth:.1e} "
          f"omega/n/1/omega {e_w:.1e} projector {e_p:.1e} omega*v {e_v:.1e}", flush=True)
    assert e_lam < 1e-12 and e_res < 1e-12 and e_orth < 1e-12
    assert e_w < 1e-9 and e_p < 1e-9 and e_v < 1e-10
    for k in ("grp_vel", "nbe", "inv_omega"):
        assert np.all(getattr(g, k)[gamma][:, :3] == 0)
    return g


def _qset(qm, n_random, seed):
    rng = np.random.default_rng(seed)
    q = np.concatenate([qm.q_cart, rng.uniform(-1.0, 1.0, size=(n_random, 3))])
    qi = np.concatenate([qm.q_int, np.ones((n_random, 3), int)])
    return q, qi


def test_phonons_rocksalt2(rocksalt):
    from spring_amd import mesh
    qm = mesh.make_mesh(rocksalt.crystal, (4, 4, 4))
    q, qi = _qset(qm, 50, 1)
    _check(rocksalt.crystal, rocksalt.fc2, q, qi)


def test_phonons_rocksalt10(rocksalt10):
    from spring_amd import mesh
    qm = mesh.make_mesh(rocksalt10.crystal, (3, 3, 3))
    q, qi = _qset(qm, 20, 2)
    _check(rocksalt10.crystal, rocksalt10.fc2, q, qi)


def test_phonons_monatomic_odd_ndof():
    # Ndof = 3: the round-robin pairing carries one idle slot
    cr, fc2 = fcc_monatomic()
    q = np.concatenate([np.zeros((1, 3)), np.random.default_rng(3).uniform(-1.5, 1.5, size=(30, 3))])
    qi = np.concatenate([np.zeros((1, 3), int), np.ones((30, 3), int)])
    _check(cr, fc2, q, qi)


def test_phonons_feed_the_linewidth_path(rocksalt):
    """Linewidths computed from GPU phonons equal those from oracle phonons, summed over each
    degenerate set of q0 modes (how the sum splits between the modes of a set depends on the
    basis chosen inside it; sums over all q1 and q2 modes do not)."""
    from spring_amd import mesh
    from spring_amd.lib import GpuContext, get_phonon
    from spring_amd.phonons import Phonons
    qm = mesh.make_mesh(rocksalt.crystal, (5, 5, 5))
    g = get_phonon(rocksalt.crystal, rocksalt.fc2, qm.q_cart, qm.q_int, 300.0)
    o = O.phonons(rocksalt.crystal, rocksalt.fc2, qm.q_cart, qm.q_int, 300.0)
    ph_o = Phonons(o["evec"], o["omega"], o["inv_omega"], o["nbe"], o["grp_vel"], 300.0)
    sigma = 1.0
    lw_g = GpuContext(rocksalt, qm, g).lw3_all(sigma)
    lw_o = GpuContext(rocksalt, qm, ph_o).lw3_all(sigma)
    e = 0.0
    for k, gi in enumerate(qm.irr_idx):
        for grp in degenerate_sets(o["omega"][gi]):
            if gi == 0 and grp[0] < 3:
                continue
            e = max(e, abs(lw_g[k, grp].sum() - lw_o[k, grp].sum()) / np.abs(lw_o).max())
    print(f"linewidth from GPU phonons vs oracle phonons, degenerate-set sums: {e:.1e}", flush=True)
    assert e < 1e-9

