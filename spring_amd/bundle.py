"""Input bundle for the Fortran host driver (fortran/kappa_gpu).

A directory holding `bundle.nml` (sizes, T, sigma) and one raw little-endian file per
array, in the reference's column-major layout with its 1-based index conventions
(atmIndx(4,:,:) basis index, pos, irreducible grid index), which is what the Fortran side
of the boundary receives in the reference program.
"""
from __future__ import annotations

import os

import numpy as np


def write_bundle(path, model, qm, ph, T: float, sigma: float, prefix: str = "synth",
                 with_fc4: bool = False, h5: bool = False):
    """h5: also write the force constants as the reference's own files
    FC2nd_/FC3rd_/FC4th_<prefix><T>K_F.h5 (the driver reads those when present)."""
    os.makedirs(path, exist_ok=True)
    cr = model.crystal

    def put(name, arr, dt):
        np.ascontiguousarray(arr, dtype=dt).tofile(os.path.join(path, name))

    f3 = model.fc3
    idx3 = f3.atm_indx.copy()
    idx3[..., 3] += 1
    put("mass.bin", cr.mass, "<f8")
    put("atm_num3.bin", f3.atm_num, "<i4")
    put("atm_indx3.bin", idx3, "<i4")
    put("fc3.bin", f3.fc, "<f8")
    put("q_int.bin", qm.q_int, "<i4")
    put("irr_grid.bin", qm.irr_idx + 1, "<i4")
    put("pos.bin", qm.pos + 1, "<i4")
    put("multiplicity.bin", qm.multiplicity, "<i4")
    rot = qm.rot_cart if not qm.time_rev else np.concatenate([qm.rot_cart, -qm.rot_cart])
    put("rot.bin", np.transpose(rot, (0, 2, 1)), "<f8")          # Rxyz(3,3,nsym)
    put("evec.bin", ph.evec, "<c16")
    put("omega.bin", ph.omega, "<f8")
    put("inv_omega.bin", ph.inv_omega, "<f8")
    put("nbe.bin", ph.nbe, "<f8")
    put("grp_vel.bin", ph.grp_vel, "<f8")
    # harmonic inputs of get_phonon: FC2type (atmIndx2 1-based, fC2(1,3,3,Natm2,Nb)), the lattice
    # vectors as the columns of latvec(3,3), Cartesian q points (3, Nq)
    f2 = model.fc2
    idx2 = f2.atm_indx.copy()
    idx2[..., 3] += 1
    put("atm_num2.bin", f2.atm_num, "<i4")
    put("atm_indx2.bin", idx2, "<i4")
    put("fc2.bin", f2.fc, "<f8")
    put("latvec.bin", np.asarray(cr.latvec).T, "<f8")
    put("q_cart.bin", qm.q_cart, "<f8")
    if getattr(qm, "tetra", None) is not None:
        put("tetra.bin", np.asarray(qm.tetra) + 1, "<i4")          # tetrahedrons(4, 6 Nq), 1-based
    natm4 = 1
    if with_fc4:
        f4 = model.fc4
        idx4 = f4.atm_indx.copy()
        idx4[..., 3] += 1
        put("atm_num4.bin", f4.atm_num, "<i4")
        put("atm_indx4.bin", idx4, "<i4")
        put("fc4.bin", f4.fc, "<f8")
        natm4 = f4.natm_max
    if h5:
        from . import h5io
        for f in (model.fc2, model.fc3, model.fc4 if with_fc4 else None):
            if f is not None:
                h5io.write_fc(os.path.join(path, h5io.fc_file_name(f.order, prefix, T)), f)
    with open(os.path.join(path, "bundle.nml"), "w") as fh:
        fh.write("&bundle\n")
        fh.write(f"  prefix = '{prefix}'\n  Nbasis = {cr.nb}\n  Natm2 = {f2.natm_max}\n  Natm3 = {f3.natm_max}\n")
        fh.write(f"  Natm4 = {natm4}\n  mesh = {int(qm.mesh[0])}, {int(qm.mesh[1])}, {int(qm.mesh[2])}\n")
        fh.write(f"  Nq = {qm.nq}\n  Nq0 = {qm.nq0}\n  Nsym = {rot.shape[0]}\n")
        fh.write(f"  T = {T!r}\n  Sigma = {sigma!r}\n  vol = {cr.vol!r}\n  FourPhonon = {int(with_fc4)}\n/\n")
    return path


def _cell_nml(model, prefix, four_phonon=False):
    """&FCInfo, &CrystalInfo, &AtmInfo lines (cell%init_data, src/Kappa/Src/unit_mod.f90:64-214).
    Species become types (typ_indx 1-based, one mass per type); reals in shortest round-trip
    form, so the reader gets the same doubles."""
    cr = model.crystal
    species, typ = np.unique(np.asarray(cr.species), return_inverse=True)
    mass_typ = [float(np.asarray(cr.mass)[typ == t][0]) for t in range(species.size)]
    return ["&FCInfo", f"  prefix = '{prefix}'", "  SecondFC = .true.", "  ThirdFC = .true.",
            f"  FourthFC = {'.true.' if four_phonon else '.false.'}", "  Force_dd = .false.", "/",
            "&CrystalInfo", f"  Nbasis = {cr.nb}", f"  Ntyp = {species.size}",
            f"  Latvec = {_reals(np.asarray(cr.latvec).T.ravel())}", "  Sup_Cell = 5, 5, 5", "/",
            "&AtmInfo", f"  atm_pos = {_reals(np.asarray(cr.basis_frac).ravel())}",
            f"  typ_indx = {', '.join(str(int(t) + 1) for t in typ)}",
            f"  typ_lbl = {', '.join(repr('X%d' % t) for t in range(species.size))}",
            f"  mass = {_reals(mass_typ)}", "/"]


def _reals(xs):
    return ", ".join(repr(float(x)) for x in xs)


def _write_fc_files(path, model, prefix, T, orders):
    from . import h5io
    for ifc in (model.fc2, model.fc3, model.fc4 if 4 in orders else None):
        if ifc is not None:
            h5io.write_fc(os.path.join(path, h5io.fc_file_name(ifc.order, prefix, T)), ifc)


def write_input_nml(path, model, qmesh, T: float, prefix: str = "synth", sigma: float = 5.1,
                    time_rev: int = 0, symprec: float = 1e-5, rules3: str = "DDD", rules4: str = "DDDD",
                    four_phonon: bool = False):
    """A kappa.x input: `input.nml` with &FCInfo, &CrystalInfo, &AtmInfo and &AnhInfo (ReadNmlKappa,
    src/Kappa/Src/Helper_m.f90:37-132), plus the force constants as
    FC2nd_/FC3rd_[/FC4th_]<prefix><T>K_F.h5 beside it."""
    os.makedirs(path, exist_ok=True)
    lines = _cell_nml(model, prefix, four_phonon) + [
        "&AnhInfo", f"  SelectionRule3 = {', '.join(repr(c) for c in rules3)}",
        f"  SelectionRule4 = {', '.join(repr(c) for c in rules4)}", f"  Sigma = {float(sigma)!r}",
        f"  symprec = {float(symprec)!r}", f"  qmesh = {', '.join(str(int(x)) for x in qmesh)}",
        f"  time_rev = {int(time_rev)}", f"  FourPhonon = {'.true.' if four_phonon else '.false.'}", "/", ""]
    with open(os.path.join(path, "input.nml"), "w") as fh:
        fh.write("\n".join(lines))
    _write_fc_files(path, model, prefix, T, (2, 3, 4) if four_phonon else (2, 3))
    return path


def write_linewidth_nml(path, model, qmesh, T: float, q_high_sym, prefix: str = "synth", sigma: float = 0.1):
    """A LineWidth.x input: `input.nml` with the cell namelists, &LWInfo (Sigma, qmesh, num_points)
    and &HighSymPathLW (q_high_sym(:, n) = reduced coordinates and the point count of the segment
    ending at node n; src/LineWidth/Src/main.f90:70-76), plus FC2nd_/FC3rd_<prefix><T>K_F.h5."""
    os.makedirs(path, exist_ok=True)
    nodes = np.asarray(q_high_sym, float)
    lines = _cell_nml(model, prefix) + [
        "&LWInfo", f"  Sigma = {float(sigma)!r}", f"  qmesh = {', '.join(str(int(x)) for x in qmesh)}",
        f"  num_points = {nodes.shape[0]}", "/", "&HighSymPathLW"]
    lines += [f"  q_high_sym(:, {n + 1}) = {_reals(row[:3])}, {int(row[3])}" for n, row in enumerate(nodes)]
    lines += ["/", ""]
    with open(os.path.join(path, "input.nml"), "w") as fh:
        fh.write("\n".join(lines))
    _write_fc_files(path, model, prefix, T, (2, 3))
    return path


def read_linewidth_path_file(fname, ndof):
    """Rows of LineWidth.x's linewidth_*.txt: (distance, omega(ndof), lw(ndof)) per path point."""
    data = _read_rows(fname)
    return data[:, 0], data[:, 1:ndof + 1], data[:, ndof + 1:2 * ndof + 1]


def read_linewidth_file(fname, ndof):
    """Rows of linewidth_IBZ_*.txt: (omega(ndof), lw(ndof)) per irreducible q0."""
    data = _read_rows(fname)
    return data[:, :ndof], data[:, ndof:2 * ndof]


def _read_rows(fname):
    """Fortran's G20.12 drops the exponent letter when the exponent has three digits
    ('0.4920108584-171'), so such fields are repaired before parsing."""
    import re
    rows = []
    for line in open(fname):
        if line.lstrip().startswith("#") or not line.strip():
            continue
        fields = [re.sub(r"(\d)([-+]\d{3})$", r"\1E\2", f) for f in line.split()]
        rows.append([float(f) for f in fields])
    return np.atleast_2d(np.array(rows))
