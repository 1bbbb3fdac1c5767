"""The Fortran host driver builds, reads its input bundle and reports boundary errors the
way the reference does (message + ERROR STOP)."""
import os
import subprocess

import pytest

from tests.conftest import ROOT, problem


def _build():
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "fortran")], check=True)
    return os.path.join(ROOT, "fortran", "kappa_gpu")


def test_fortran_driver_error_path(rocksalt, tmp_path):
    from spring_amd import bundle
    exe = _build()
    qm, ph = problem(rocksalt, (3, 3, 3))
    bundle.write_bundle(str(tmp_path / "b"), rocksalt, qm, ph, 300.0, 0.1)
    env = dict(os.environ, HIP_VISIBLE_DEVICES="-1", ROCR_VISIBLE_DEVICES="-1")
    r = subprocess.run([exe, str(tmp_path / "b"), "fp64", "loop"], cwd=tmp_path, env=env,
                       capture_output=True, text=True, timeout=120)
    assert r.returncode != 0
    assert "spr_ctx_create failed" in r.stdout
