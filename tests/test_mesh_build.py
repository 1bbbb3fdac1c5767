"""spr_mesh_build (host code in the library, the mesh tables of make_tetrahedron + FindIrrQ,
src/Kappa/Src/TetrahedronQ_m.f90:56-331) against spring_amd.mesh.make_mesh, whose tables are
pinned in test_host.py (irreducible counts, tetrahedron coverage, orbit invariants): identical
integer tables and rotations, for cubic, low-symmetry and monatomic cells, odd, even and mixed
meshes, with and without time reversal.  Runs without a GPU (no device call)."""
import numpy as np
import pytest

from spring_amd import mesh, model
from tests.test_phonon_oracle import fcc_monatomic


_CACHE = {}


def _crystal(name):
    if name not in _CACHE:
        _CACHE[name] = {"rocksalt2": model.rocksalt_primitive, "rocksalt10": lambda: model.rocksalt10().crystal,
                        "rocksalt4": lambda: model.rocksalt_chain(2).crystal,
                        "fcc1": lambda: fcc_monatomic()[0]}[name]()
    return _CACHE[name]


CASES = [(n, d, t) for n in ("rocksalt2", "fcc1") for d in ((4, 4, 4), (5, 5, 5), (3, 4, 6)) for t in (0, 1)]
CASES += [("rocksalt4", (4, 4, 4), 0), ("rocksalt4", (3, 4, 6), 1), ("rocksalt10", (4, 4, 4), 1)]


@pytest.mark.parametrize("name,dims,time_rev", CASES)
def test_mesh_build_matches_python(name, dims, time_rev):
    from spring_amd.lib import mesh_build
    cr = _crystal(name)
    ref = mesh.make_mesh(cr, dims, time_rev=time_rev)
    got = mesh_build(cr, dims, time_rev=time_rev)
    for key, want in (("q_int", ref.q_int), ("irr_idx", ref.irr_idx), ("pos", ref.pos),
                      ("multiplicity", ref.multiplicity), ("tetra", ref.tetra)):
        assert np.array_equal(got[key], want), (name, dims, time_rev, key)
    assert np.allclose(got["q_cart"], ref.q_cart, rtol=0, atol=1e-13)
    rot = ref.rot_cart if not time_rev else np.concatenate([ref.rot_cart, -ref.rot_cart])
    assert got["rot"].shape == rot.shape and np.allclose(got["rot"], rot, rtol=0, atol=1e-12)


def test_mesh_build_counts_and_identity():
    from spring_amd.lib import mesh_build
    cr = model.rocksalt_primitive()
    for dims, nq0 in (((8, 8, 8), 29), ((16, 16, 16), 145)):
        got = mesh_build(cr, dims)
        assert got["irr_idx"].size == nq0 and got["rot"].shape[0] == 48
        assert got["multiplicity"].sum() == np.prod(dims)
    got = mesh_build(cr, (4, 4, 4), symmetry=False)
    assert got["irr_idx"].size == 64 and got["rot"].shape == (1, 3, 3)
    assert np.allclose(got["rot"][0], np.eye(3), rtol=0, atol=1e-14)


def test_mesh_build_rejects_bad_input():
    from spring_amd.lib import SprError, mesh_build
    cr = model.rocksalt_primitive()
    with pytest.raises(SprError):
        mesh_build(cr, (0, 4, 4))
    with pytest.raises(SprError):
        mesh_build(cr, (4, 4, 4), symprec=0.0)
