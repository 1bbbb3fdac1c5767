"""The Fortran host from the reference's own inputs only: `kappa_gpu -inp input.nml -T 300` in a
directory holding input.nml (&FCInfo, &CrystalInfo, &AtmInfo, &AnhInfo) and the FC*_F.h5 files;
it builds the mesh, irreducible points and tetrahedra itself (spr_mesh_build), solves the phonons
on the device (get_phonon_gpu) and runs the tetrahedron linewidths and kappa.  Checked against the
oracle fed with the same device phonons at the same q points (so that degenerate modes carry the
same eigenvectors): linewidth_IBZ_*.txt, the kappa file's tensor, LineWidth, PhaseSpace3rd,
ModeResolvedKappa and SpectralResolvedKappa, with and without a selection rule and with the
four-phonon term."""
import os
import re
import subprocess

import numpy as np
import pytest

from tests.conftest import ROOT, rel_err

pytestmark = pytest.mark.gpu


def _run(tmp_path, args):
    exe = os.path.join(ROOT, "fortran", "kappa_gpu")
    assert os.path.exists(exe), "fortran/kappa_gpu is not built"
    r = subprocess.run([exe, *args], cwd=tmp_path, capture_output=True, text=True, timeout=300)
    print(r.stdout[-3000:], r.stderr[-2000:], flush=True)
    assert r.returncode == 0
    return r.stdout


def _oracle(m, dims, rules3="DDD", four=False):
    """Oracle linewidths, P3 and kappa on the mesh the host builds, with device phonons."""
    from spring_amd import mesh
    from spring_amd.lib import get_phonon, mesh_build
    from oracle import oracle as O
    qm = mesh.make_mesh(m.crystal, dims)
    mb = mesh_build(m.crystal, dims, symprec=1e-5)
    assert np.array_equal(mb["irr_idx"], qm.irr_idx) and np.array_equal(mb["tetra"], qm.tetra)
    ph = get_phonon(m.crystal, m.fc2, mb["q_cart"], mb["q_int"], 300.0)
    nd = m.crystal.ndof
    pol3 = O.polarization(3, nd, rules3)
    inp = O.Inputs(m, qm, ph)
    inp4 = O.Inputs(m, qm, ph, which=4) if four else None
    lw = np.zeros((qm.nq0, nd))
    p3 = np.zeros((qm.nq0, nd))
    for k, g in enumerate(qm.irr_idx):
        _, W = O.lw3_gauss(inp, int(g), 1.0, want_w=True)
        lw[k], p3[k] = O.lw3_tetra(qm, ph, int(g), W, pol=pol3)
        if four:
            lw[k] += 0.5 * O.lw4_tetra(inp4, qm, int(g), threads=8, chunk=4)
    kap = O.kappa_rta_symm(qm, ph, lw, 300.0, m.crystal.vol, srange=[pol3[0], pol3[1]])
    return qm, ph, lw, p3, kap


@pytest.mark.parametrize("rules3", ["DDD", "ADD"])
def test_namelist_host_three_phonon(rocksalt, tmp_path, rules3):
    from spring_amd import bundle, h5io
    bundle.write_input_nml(str(tmp_path), rocksalt, (4, 4, 4), 300.0, prefix="NaClnml", rules3=rules3)
    out = _run(tmp_path, ["-inp", "input.nml", "-T", "300", "-Spectral_k", "T", "-Mode_resolved_k", "T"])
    assert "Number of irreducible q-points for (    4 X   4 X   4 ) mesh is:      8" in out
    qm, ph, ref, p3, kref = _oracle(rocksalt, (4, 4, 4), rules3)
    nd, nq0 = 6, qm.nq0
    w, lw = bundle.read_linewidth_file(str(tmp_path / "linewidth_IBZ_NaClnml300.0K.txt"), nd)
    assert np.allclose(w, ph.omega[qm.irr_idx], rtol=1e-11, atol=1e-9)
    assert rel_err(lw, ref, 1e-8) < 1e-10
    k = h5io.read_kappa(str(tmp_path / "kappa_NaClnml300.0K.h5"), nd, nq0, p3=True, mode_k=True)
    assert rel_err(k["lw"], ref, 1e-9) < 1e-10
    assert np.array_equal(k["omega"], ph.omega[qm.irr_idx])
    p3_fs = p3 * 1000.0 / (9.0 * nd ** 2 * qm.nq)
    assert rel_err(k["p3"], p3_fs, 1e-9) < 1e-10
    assert np.allclose(k["kappa"], kref, rtol=1e-10, atol=1e-12 * np.abs(kref).max())
    # ModeResolvedKappa(9, Ndof) sums over the modes to kappa; SpectralResolvedKappa summed / Nq too
    assert np.allclose(k["mode_k"].sum(0), k["kappa"].ravel(), rtol=1e-10, atol=1e-12 * np.abs(kref).max())
    spec = h5io.read_spectral(str(tmp_path / "kappa_NaClnml300.0K.h5"), nd, nq0)
    assert np.array_equal(spec[:, 0], np.repeat(np.arange(1, nd + 1), nq0))
    assert np.allclose(spec[:, 2:].sum(0) / qm.nq, k["kappa"].ravel(), rtol=1e-10,
                       atol=1e-12 * np.abs(kref).max())
    if rules3 == "ADD":
        assert "non-defult selection rule" in out
        assert np.all(lw[:, 3:] == 0.0) and np.all(spec[3 * nq0:, 2:] == 0.0)
    # the printed tensor (three rows of three G20.12 values)
    nums = [float(x) for x in re.findall(r"k[xyz][xyz] =\s+([-+0-9.E]+)", out)]
    assert np.allclose(np.array(nums), k["kappa"].ravel(), rtol=1e-10, atol=1e-9 * np.abs(kref).max())
    assert "3-phonon scattering phase-space:" in out


def test_namelist_host_four_phonon(rocksalt, tmp_path):
    from spring_amd import bundle
    bundle.write_input_nml(str(tmp_path), rocksalt, (3, 3, 3), 300.0, prefix="NaCl4", four_phonon=True)
    _run(tmp_path, ["-inp", "input.nml", "-T", "300"])
    qm, ph, ref, _, _ = _oracle(rocksalt, (3, 3, 3), four=True)
    _, lw = bundle.read_linewidth_file(str(tmp_path / "linewidth_IBZ_NaCl4300.0K.txt"), 6)
    assert rel_err(lw, ref, 1e-8) < 1e-10
