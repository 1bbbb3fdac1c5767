"""HDF5 files of the path (libspring_h5.so): the FC files kappa.x reads and the kappa file it
writes (src/Kappa/Src/hdf5_wrap_mod.f90:40-337, 650-765; main.f90:192-235, 773-798).

The reference ships no .h5 file (no fixture exists), so the layout is pinned against the
reference's reader and writer code through h5dump, the HDF5 tool, which prints what is in a
file independently of this library: dataset names and groups, stored types (32-bit integers,
IEEE doubles) and shapes (h5dump lists C-order dimensions, the reverse of the Fortran ones the
reference declares), and the element order (a tensor whose every element encodes its Fortran
indices must show fC(gamma, beta, alpha, N2, N1) at the position HDF5 reports).
"""
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

from spring_amd import h5io, model

H5DUMP = shutil.which("h5dump") or "/opt/conda/bin/h5dump"
needs_dump = pytest.mark.skipif(not os.path.exists(H5DUMP), reason="h5dump not installed")


def _dump(path, *args):
    return subprocess.run([H5DUMP, *args, path], capture_output=True, text=True, check=True).stdout


def _same(a, b):
    assert a.order == b.order
    assert np.array_equal(a.atm_num, b.atm_num)
    assert np.array_equal(a.atm_indx, b.atm_indx)
    assert a.fc.shape == b.fc.shape and np.array_equal(a.fc, b.fc)


def test_fc_round_trip(rocksalt, rocksalt10, tmp_path):
    for m in (rocksalt, rocksalt10):
        for f in (m.fc2, m.fc3, m.fc4):
            if f is None:
                continue
            p = str(tmp_path / h5io.fc_file_name(f.order, m.name, 300.0))
            h5io.write_fc(p, f)
            _same(h5io.read_fc(p, f.order), f)


@needs_dump
def test_fc_layout_matches_the_reference(rocksalt, tmp_path):
    m = rocksalt
    nb, natm3 = m.fc3.atm_indx.shape[:2]
    p3 = str(tmp_path / "FC3rd_t300.0K_F.h5")
    h5io.write_fc(p3, m.fc3)
    head = _dump(p3, "-H")
    # r3rdFC: NumAtoms (Nb), atm_Indx (4, Natm, Nb) integers, group FC3rd with FC3_mu1..Nb of
    # shape (3, 3, 3, Natm, Natm) doubles
    assert re.search(r'DATASET "NumAtoms" \{\s*DATATYPE\s+H5T_STD_I32LE\s*DATASPACE\s+SIMPLE \{ \( %d \)' % nb, head)
    assert re.search(r'DATASET "atm_Indx" \{\s*DATATYPE\s+H5T_STD_I32LE\s*DATASPACE\s+SIMPLE \{ \( %d, %d, 4 \)'
                     % (nb, natm3), head)
    assert 'GROUP "FC3rd"' in head
    for mu in range(1, nb + 1):
        assert re.search(r'DATASET "FC3_mu%d" \{\s*DATATYPE\s+H5T_IEEE_F64LE\s*DATASPACE\s+SIMPLE \{ \( %d, %d, 3, 3, 3 \)'
                         % (mu, natm3, natm3), head)
    p2 = str(tmp_path / "FC2nd_t300.0K_F.h5")
    h5io.write_fc(p2, m.fc2)
    natm2 = m.fc2.atm_indx.shape[1]
    assert re.search(r'DATASET "FC2" \{\s*DATATYPE\s+H5T_IEEE_F64LE\s*DATASPACE\s+SIMPLE \{ \( %d, %d, 3, 3, 1 \)'
                     % (nb, natm2), _dump(p2, "-H"))
    p4 = str(tmp_path / "FC4th_t300.0K_F.h5")
    h5io.write_fc(p4, m.fc4)
    natm4 = m.fc4.atm_indx.shape[1]
    assert re.search(r'DATASET "FC4_mu2" \{\s*DATATYPE\s+H5T_IEEE_F64LE\s*DATASPACE\s+SIMPLE \{ \( %d, %d, %d, 3, 3, 3, 3 \)'
                     % (natm4, natm4, natm4), _dump(p4, "-H", "-g", "FC4th"))


@needs_dump
def test_fc3_element_order(tmp_path):
    """Element (N1, N2, alpha, beta, gamma) of basis atom mu, encoded as a number, must sit at the
    HDF5 index (N1, N2, alpha, beta, gamma) of FC3_mu<mu+1>: Fortran fC(gamma, beta, alpha, N2, N1)."""
    nb, natm = 2, 3
    fc = np.zeros((nb, natm, natm, 3, 3, 3))
    for idx in np.ndindex(fc.shape):
        fc[idx] = float("".join(str(i + 1) for i in idx))
    indx = np.zeros((nb, natm, 4), np.int32)
    indx[..., 3] = np.arange(natm)[None, :] % nb
    f = model.IFC(3, np.array([natm, natm - 1], np.int32), indx, fc, {})
    p = str(tmp_path / "FC3rd_code300.0K_F.h5")
    h5io.write_fc(p, f)
    out = _dump(p, "-d", "/FC3rd/FC3_mu2", "-m", "%.0f")
    vals = [float(x) for x in re.findall(r"\b\d{6}\b", out.split("DATA {")[1])]
    assert len(vals) == natm * natm * 27
    for k, v in enumerate(vals):                   # h5dump prints in C (row-major) order
        n1, rest = divmod(k, natm * 27)
        n2, rest = divmod(rest, 27)
        a, rest = divmod(rest, 9)
        b, g = divmod(rest, 3)
        assert v == float(f"2{n1 + 1}{n2 + 1}{a + 1}{b + 1}{g + 1}")
    # atm_Indx holds the basis index 1-based; NumAtoms as given
    txt = _dump(p, "-d", "/atm_Indx")
    assert re.findall(r"\(0,1,0\): ([^\n]+)", txt)[0].split(",")[3].strip() == "2"


@needs_dump
def test_kappa_file(tmp_path):
    rng = np.random.default_rng(0)
    nq0, nd = 5, 6
    kappa = rng.normal(size=(3, 3))
    w = rng.uniform(1, 50, size=(nq0, nd))
    lw = rng.uniform(0, 1, size=(nq0, nd))
    p3 = rng.uniform(0, 1, size=(nq0, nd))
    mk = rng.normal(size=(nd, 9))
    p = str(tmp_path / h5io.kappa_file_name("NaCl", 300.0))
    assert os.path.basename(p) == "kappa_NaCl300.0K.h5"
    h5io.write_kappa(p, kappa, w, lw, p3, mk)
    back = h5io.read_kappa(p, nd, nq0, p3=True, mode_k=True)
    assert np.array_equal(back["kappa"], kappa) and np.array_equal(back["omega"], w)
    assert np.array_equal(back["lw"], lw) and np.array_equal(back["p3"], p3) and np.array_equal(back["mode_k"], mk)
    head = _dump(p, "-H")
    # w1dh5_no_grp / w2dh5_no_grp_type2 / w2dh5_no_grp: (9), (2 Ndof, Nq0), (9, Ndof) in Fortran order
    assert re.search(r'DATASET "ThermalConductivityTensor" \{\s*DATATYPE\s+H5T_IEEE_F64LE\s*DATASPACE\s+SIMPLE \{ \( 9 \)', head)
    assert re.search(r'DATASET "LineWidth" \{\s*DATATYPE\s+H5T_IEEE_F64LE\s*DATASPACE\s+SIMPLE \{ \( %d, %d \)' % (nq0, 2 * nd), head)
    assert re.search(r'DATASET "PhaseSpace3rd" \{\s*DATATYPE\s+H5T_IEEE_F64LE\s*DATASPACE\s+SIMPLE \{ \( %d, %d \)' % (nq0, 2 * nd), head)
    assert re.search(r'DATASET "ModeResolvedKappa" \{\s*DATATYPE\s+H5T_IEEE_F64LE\s*DATASPACE\s+SIMPLE \{ \( %d, 9 \)' % nd, head)
    # LineWidth(1:Ndof, q0) = omega, LineWidth(Ndof+1:, q0) = linewidth (w2dh5_no_grp_type2)
    txt = re.sub(r"\(\d+,\d+\):", " ", _dump(p, "-d", "/LineWidth", "-m", "%.17g").split("DATA {")[1])
    vals = np.array([float(x) for x in re.findall(r"[-+]?\d+\.\d*(?:[eE][-+]?\d+)?", txt)])
    assert vals.size == nq0 * 2 * nd
    rows = vals.reshape(nq0, 2 * nd)                  # h5dump order: q0 outer, then 2 Ndof
    assert np.array_equal(rows[:, :nd], w) and np.array_equal(rows[:, nd:], lw)
    # -Spectral_k: SpectralResolvedKappa (11, Ndof Nq0) added to the same file (main.f90:780-794)
    spec = rng.normal(size=(nd * nq0, 11))
    h5io.write_spectral(p, spec)
    assert np.array_equal(h5io.read_spectral(p, nd, nq0), spec)
    head = _dump(p, "-H")
    assert re.search(r'DATASET "SpectralResolvedKappa" \{\s*DATATYPE\s+H5T_IEEE_F64LE\s*DATASPACE\s+SIMPLE \{ \( %d, 11 \)' % (nd * nq0), head)
    assert "ThermalConductivityTensor" in head          # the earlier datasets stay


def test_errors_are_reported(tmp_path):
    p = str(tmp_path / "missing.h5")
    with pytest.raises(h5io.H5Error, match="cannot open"):
        h5io.read_fc(p, 3)
    f = model.rocksalt2().fc3
    q = str(tmp_path / "fc3.h5")
    h5io.write_fc(q, f)
    with pytest.raises(h5io.H5Error, match="no group FC4th"):
        h5io.read_fc(q, 4)


def test_library_exports_every_declared_symbol():
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    hdr = open(os.path.join(root, "include", "spring_h5.h")).read()
    declared = set(re.findall(r"\b(spr_h5_[a-z0-9_]+)\s*\(", hdr))
    L = h5io.load()
    for name in sorted(declared):
        assert hasattr(L, name), name
    assert declared == set(h5io.SYMBOLS)
