"""Fortran host driver on the GPU: linewidth_IBZ file and kappa against the oracle."""
import os
import re
import subprocess

import numpy as np
import pytest

from tests.conftest import ROOT, problem, rel_err

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("mode", ["loop", "all"])
def test_fortran_driver_parity(rocksalt, tmp_path, mode):
    from spring_amd import bundle
    from oracle import oracle as O
    exe = os.path.join(ROOT, "fortran", "kappa_gpu")
    assert os.path.exists(exe), "fortran/kappa_gpu is not built"
    qm, ph = problem(rocksalt, (4, 4, 4))
    # sigma = 1 rad/ps: every mode of the coarse mesh has a finite linewidth, so kappa is finite
    bundle.write_bundle(str(tmp_path / "b"), rocksalt, qm, ph, 300.0, 1.0, prefix="NaClsynth")
    r = subprocess.run([exe, str(tmp_path / "b"), "fp64", mode], cwd=tmp_path, capture_output=True,
                       text=True, timeout=300)
    print(r.stdout[-2000:], r.stderr[-2000:], flush=True)
    assert r.returncode == 0
    w, lw = bundle.read_linewidth_file(str(tmp_path / "linewidth_IBZ_NaClsynth300.0K.txt"), 6)
    ref = O.lw3_gauss_all(O.Inputs(rocksalt, qm, ph), qm.irr_idx, 1.0)
    assert np.allclose(w, ph.omega[qm.irr_idx], rtol=1e-11, atol=1e-9)
    assert rel_err(lw, ref, 1e-8) < 1e-10
    nums = [float(x) for x in re.findall(r"[-+]?\d\.\d+E[-+]\d+", r.stdout)]
    kap = np.array(nums[-9:]).reshape(3, 3)
    kref = O.kappa_rta_symm(qm, ph, ref, 300.0, rocksalt.crystal.vol)
    assert np.allclose(kap, kref, rtol=1e-10)
