"""The LineWidth.x host from the reference's own inputs only: `linewidth_gpu -inp input.nml -T 300`
in a directory holding input.nml (&FCInfo, &CrystalInfo, &AtmInfo, &LWInfo, &HighSymPathLW) and
the FC2/FC3 files.  It builds the path, the q1 mesh and the phonons itself and writes
linewidth_<prefix><T>K.txt.  Checked against the oracle: the path's distances and points
(q_points_highsymm), and per point the off-mesh linewidth (lw3_offmesh, pinned in
test_offmesh_oracle.py) with the q0, q1 and q2 phonons from the same device solver.  The nodes are
generic points, so no q2 has degenerate branches and the tetrahedron sums agree branch by branch;
the path has zero-centred and non-zero-centred segments."""
import os
import subprocess

import numpy as np
import pytest

from tests.conftest import ROOT, rel_err

pytestmark = pytest.mark.gpu

NODES = [[0.02, 0.13, 0.05, 0],
         [0.31, -0.12, 0.07, 2],     # a negative coordinate: zero-centred
         [0.61, 0.72, 0.83, 1],      # starts at a negative coordinate: zero-centred
         [0.66, 0.93, 0.71, 2]]      # both ends outside [-1/2, 1/2]: not zero-centred


def _run(tmp_path, args):
    exe = os.path.join(ROOT, "fortran", "linewidth_gpu")
    assert os.path.exists(exe), "fortran/linewidth_gpu is not built"
    r = subprocess.run([exe, *args], cwd=tmp_path, capture_output=True, text=True, timeout=300)
    print(r.stdout[-3000:], r.stderr[-2000:], flush=True)
    assert r.returncode == 0
    return r.stdout


@pytest.mark.parametrize("delta", ["tetra", "gauss"])
def test_linewidth_host_path(rocksalt, tmp_path, delta):
    from spring_amd import bundle
    from spring_amd.lib import get_phonon
    from oracle import oracle as O
    dims, sigma = (4, 4, 4), 0.3
    bundle.write_linewidth_nml(str(tmp_path), rocksalt, dims, 300.0, NODES, prefix="NaCllw", sigma=sigma)
    args = ["-inp", "input.nml", "-T", "300"] + (["-delta", "gauss"] if delta == "gauss" else [])
    out = _run(tmp_path, args)
    assert "Linewidths written in file: linewidth_NaCllw300.0K.txt" in out
    nd = rocksalt.crystal.ndof
    dist, w, lw = bundle.read_linewidth_path_file(str(tmp_path / "linewidth_NaCllw300.0K.txt"), nd)
    q0, d_ref, zc = O.q_points_highsymm(NODES, rocksalt.crystal.G)
    assert zc.tolist() == [True, True, True, False, False]
    assert np.allclose(dist, d_ref, rtol=1e-11)
    # the mesh's phonons as the host solves them (degenerate modes of grid points carry the same
    # eigenvectors on both sides)
    from spring_amd import mesh
    from spring_amd.lib import mesh_build
    qm = mesh.make_mesh(rocksalt.crystal, dims)
    mb = mesh_build(rocksalt.crystal, dims, symprec=1e-5)
    assert np.array_equal(mb["tetra"], qm.tetra)
    cr = rocksalt.crystal
    ph = get_phonon(cr, rocksalt.fc2, mb["q_cart"], mb["q_int"], 300.0)
    p0 = get_phonon(cr, rocksalt.fc2, q0, np.ones((q0.shape[0], 3), np.int32), 300.0)
    assert np.allclose(w, p0.omega, rtol=1e-11)
    for i in range(q0.shape[0]):
        _, q2c, exists = O.offmesh_q2(rocksalt, qm, q0[i], bool(zc[i]))
        g2 = np.all(np.abs(q2c) < 1e-10, axis=1)
        p2 = get_phonon(cr, rocksalt.fc2, q2c, np.repeat(np.where(g2, 0, 1)[:, None], 3, axis=1), 300.0)
        ph0 = dict(evec=p0.evec[i], omega=p0.omega[i], inv_omega=p0.inv_omega[i], nbe=p0.nbe[i])
        ph2 = dict(evec=p2.evec, omega=p2.omega, inv_omega=p2.inv_omega, nbe=p2.nbe)
        ref, _, _ = O.lw3_offmesh(rocksalt, qm, ph, q0[i], ph0, 300.0, delta, sigma, bool(zc[i]), ph2)
        assert ref.max() > 0
        e = rel_err(lw[i], ref, 1e-8)
        print(f"path point {i} (zero-centred {zc[i]}): lw rel err {e:.2e}", flush=True)
        assert e < 1e-10, i
