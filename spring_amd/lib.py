"""ctypes binding of libspring_gpu.so (include/spring_gpu.h).

The product path has no CPU fallback: if the library is missing or no HIP device is
present, constructing a context raises.
"""
from __future__ import annotations

import ctypes as ct
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libspring_gpu.so")

PREC = {"fp64": 0, "fp32": 1, "bf16": 2}

SYMBOLS = [
    "spr_ctx_create", "spr_ctx_destroy", "spr_last_error", "spr_v3_w", "spr_lw3",
    "spr_lw3_all", "spr_lw4", "spr_lw4_all", "spr_kappa_rta", "spr_kernel_time",
    "spr_kernel_time_reset", "spr_last_launch", "spr_lw3_pairs", "spr_v3_w_pairs", "spr_mlp_create", "spr_mlp_destroy",
    "spr_mlp_num_params", "spr_mlp_get_params", "spr_mlp_set_params", "spr_mlp_grad",
    "spr_mlp_train", "spr_mlp_predict", "spr_lw3_tetra", "spr_lw3_tetra_all",
    "spr_lw4_tetra", "spr_lw4_tetra_all", "spr_lw3_kappa_all", "spr_mlp_score",
    "spr_bte_irr", "spr_bte_tau_inv", "spr_bte_delta_f", "spr_bte_solve", "spr_get_phonon",
    "spr_lw3_slice", "spr_lw3_pairs_noalloc", "spr_set_branch_rules", "spr_kappa_rta_spectral",
    "spr_lw3_offmesh", "spr_mesh_build", "spr_pair_scores", "spr_lw3_sampled",
]


class SprError(RuntimeError):
    pass


class _Crystal(ct.Structure):
    _fields_ = [("nb", ct.c_int), ("mass", ct.c_void_p), ("vol", ct.c_double)]


class _IFC(ct.Structure):
    _fields_ = [("order", ct.c_int), ("natm", ct.c_int), ("atm_num", ct.c_void_p),
                ("atm_indx", ct.c_void_p), ("fc", ct.c_void_p)]


class _Mesh(ct.Structure):
    _fields_ = [("mesh", ct.c_int * 3), ("nq", ct.c_int), ("q_int", ct.c_void_p),
                ("nq0", ct.c_int), ("irr_grid", ct.c_void_p), ("pos", ct.c_void_p),
                ("multiplicity", ct.c_void_p), ("nsym", ct.c_int), ("rot", ct.c_void_p),
                ("tetra", ct.c_void_p)]


class _Phonons(ct.Structure):
    _fields_ = [("nq", ct.c_int), ("ndof", ct.c_int), ("evec", ct.c_void_p),
                ("omega", ct.c_void_p), ("inv_omega", ct.c_void_p), ("nbe", ct.c_void_p),
                ("grp_vel", ct.c_void_p)]


_lib = None


def load():
    """Load libspring_gpu.so; raise if it is not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise SprError(f"{LIB_PATH} is not built: run __graft_entry__.build() "
                           "(or python -m spring_amd.build)")
        lib = ct.CDLL(LIB_PATH)
        lib.spr_last_error.restype = ct.c_char_p
        _lib = lib
    return _lib


def _ptr(a):
    return a.ctypes.data_as(ct.c_void_p)


class GpuContext:
    """One spr_ctx: model + mesh + phonons resident on the GPU."""

    def __init__(self, model, qm, ph, device: int = 0, with_fc4: bool | None = None):
        self.lib = load()
        cr = model.crystal
        self.nd = cr.ndof
        self.nq0 = qm.nq0
        self.nq = qm.nq
        keep = []

        def arr(x, dt):
            a = np.ascontiguousarray(x, dtype=dt)
            keep.append(a)
            return a

        mass = arr(cr.mass, np.float64)
        self._cr = _Crystal(cr.nb, _ptr(mass), cr.vol)

        def ifc(f):
            if f is None:
                return None
            s = _IFC(f.order, f.natm_max, _ptr(arr(f.atm_num, np.int32)),
                     _ptr(arr(f.atm_indx, np.int32)), _ptr(arr(f.fc, np.float64)))
            return s

        self._fc3 = ifc(model.fc3)
        use4 = model.fc4 is not None if with_fc4 is None else with_fc4
        self._fc4 = ifc(model.fc4) if use4 else None
        rot = qm.rot_cart
        if qm.time_rev:
            rot = np.concatenate([rot, -rot])
        # Fortran Rxyz(3,3,nsym): R(a,b) at a + 3 b  -> C array [r][b][a] = R^T rows
        rotf = arr(np.transpose(rot, (0, 2, 1)), np.float64)
        self._mesh = _Mesh((ct.c_int * 3)(*[int(x) for x in qm.mesh]), qm.nq,
                           _ptr(arr(qm.q_int, np.int32)), qm.nq0, _ptr(arr(qm.irr_idx, np.int32)),
                           _ptr(arr(qm.pos, np.int32)), _ptr(arr(qm.multiplicity, np.int32)),
                           rot.shape[0], _ptr(rotf),
                           _ptr(arr(qm.tetra, np.int32)) if getattr(qm, "tetra", None) is not None else None)
        self._ph = _Phonons(ph.nq, self.nd, _ptr(arr(ph.evec, np.complex128)),
                            _ptr(arr(ph.omega, np.float64)), _ptr(arr(ph.inv_omega, np.float64)),
                            _ptr(arr(ph.nbe, np.float64)), _ptr(arr(ph.grp_vel, np.float64)))
        self._keep = keep
        h = ct.c_void_p()
        rc = self.lib.spr_ctx_create(ct.byref(self._cr), ct.byref(self._fc3),
                                     ct.byref(self._fc4) if self._fc4 is not None else None,
                                     ct.byref(self._mesh), ct.byref(self._ph), ct.c_int(device),
                                     ct.byref(h))
        self._check(rc, "spr_ctx_create")
        self.h = h
        self._keep = None   # the library holds device copies now

    def _check(self, rc, what):
        if rc != 0:
            raise SprError(f"{what} failed ({rc}): {self.lib.spr_last_error().decode()}")

    def close(self):
        if getattr(self, "h", None):
            self.lib.spr_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- entry points ------------------------------------------------------------------
    def v3_w(self, i0: int) -> np.ndarray:
        """Sctrq1(s2,s1,s0,q1) as numpy [q1][s0][s1][s2]."""
        W = np.zeros((self.nq, self.nd, self.nd, self.nd))
        self._check(self.lib.spr_v3_w(self.h, ct.c_int(i0), _ptr(W)), "spr_v3_w")
        return W

    def v3_w_pairs(self, q0_index, q1_grid) -> np.ndarray:
        """W for (irreducible q0, q1) pairs as numpy [pair][s0][s1][s2]."""
        q0 = np.ascontiguousarray(q0_index, np.int32)
        q1 = np.ascontiguousarray(q1_grid, np.int32)
        W = np.zeros((q0.size, self.nd, self.nd, self.nd))
        self._check(self.lib.spr_v3_w_pairs(self.h, ct.c_int(q0.size), _ptr(q0), _ptr(q1), _ptr(W)),
                    "spr_v3_w_pairs")
        return W

    def set_branch_rules(self, order: int, rules):
        """Branch selection of every later `order`-phonon linewidth and of kappa: `rules` is a
        string of order letters ('A' acoustic, 'O' optical, anything else every branch), as
        SelectionRule3 / SelectionRule4 of &AnhInfo, or an explicit Polarization array
        (order x 2, 1-based inclusive)."""
        if isinstance(rules, str):
            pol = []
            for r in rules[:order]:
                pol += [1, 3] if r in "Aa" else [4, self.nd] if r in "Oo" else [1, self.nd]
        else:
            pol = list(np.asarray(rules).ravel())
        p = np.ascontiguousarray(pol, np.int32)
        self._check(self.lib.spr_set_branch_rules(self.h, ct.c_int(order), _ptr(p)), "spr_set_branch_rules")

    def lw3(self, i0: int, sigma: float, prec: str = "fp64") -> np.ndarray:
        lw = np.zeros(self.nd)
        self._check(self.lib.spr_lw3(self.h, ct.c_int(i0), ct.c_double(sigma),
                                     ct.c_int(PREC[prec]), _ptr(lw)), "spr_lw3")
        return lw

    def lw3_all(self, sigma: float, prec: str = "fp64") -> np.ndarray:
        lw = np.zeros((self.nq0, self.nd))
        self._check(self.lib.spr_lw3_all(self.h, ct.c_double(sigma), ct.c_int(PREC[prec]),
                                         _ptr(lw)), "spr_lw3_all")
        return lw

    def lw3_slice(self, q0_off: int, q0_cnt: int, sigma: float, prec: str = "fp64") -> np.ndarray:
        """Linewidths of the irreducible q0 [q0_off, q0_off + q0_cnt) in one launch."""
        lw = np.zeros((q0_cnt, self.nd))
        self._check(self.lib.spr_lw3_slice(self.h, ct.c_int(q0_off), ct.c_int(q0_cnt), ct.c_double(sigma),
                                           ct.c_int(PREC[prec]), _ptr(lw)), "spr_lw3_slice")
        return lw

    def lw4(self, i0: int, sigma: float, prec: str = "fp64") -> np.ndarray:
        lw = np.zeros(self.nd)
        self._check(self.lib.spr_lw4(self.h, ct.c_int(i0), ct.c_double(sigma),
                                     ct.c_int(PREC[prec]), _ptr(lw)), "spr_lw4")
        return lw

    def lw4_all(self, sigma: float, prec: str = "fp64") -> np.ndarray:
        lw = np.zeros((self.nq0, self.nd))
        self._check(self.lib.spr_lw4_all(self.h, ct.c_double(sigma), ct.c_int(PREC[prec]),
                                         _ptr(lw)), "spr_lw4_all")
        return lw

    def kappa(self, lw_all: np.ndarray, T: float, symm: bool = True):
        k = np.zeros(9)
        modes = np.zeros((self.nd, 9))
        lw = np.ascontiguousarray(lw_all, np.float64)
        self._check(self.lib.spr_kappa_rta(self.h, _ptr(lw), ct.c_double(T), ct.c_int(int(symm)),
                                           _ptr(k), _ptr(modes)), "spr_kappa_rta")
        return k.reshape(3, 3), modes

    def kappa_spectral(self, lw_all: np.ndarray, T: float, symm: bool = True):
        """(kappa (3, 3), mode_k (Ndof, 9), Spectral_k as (Ndof * nq0, 11) rows (s + 1, omega,
        nine kappa_ab terms x kW_m_k / (T^2 V)), row qi0 + s nq0)."""
        k = np.zeros(9)
        modes = np.zeros((self.nd, 9))
        spec = np.zeros((self.nd * self.nq0, 11))
        lw = np.ascontiguousarray(lw_all, np.float64)
        self._check(self.lib.spr_kappa_rta_spectral(self.h, _ptr(lw), ct.c_double(T), ct.c_int(int(symm)),
                                                    _ptr(k), _ptr(modes), _ptr(spec)), "spr_kappa_rta_spectral")
        return k.reshape(3, 3), modes, spec

    def lw3_kappa_all(self, sigma: float, prec: str, T: float, symm: bool = True):
        """Linewidths of every irreducible q0 and the RTA kappa in one call, LW_Allq0 kept on the
        device in between: (lw (nq0, Ndof), kappa (3, 3), mode_k (Ndof, 9))."""
        lw = np.zeros((self.nq0, self.nd))
        k = np.zeros(9)
        modes = np.zeros((self.nd, 9))
        self._check(self.lib.spr_lw3_kappa_all(self.h, ct.c_double(sigma), ct.c_int(PREC[prec]), ct.c_double(T),
                                               ct.c_int(int(symm)), _ptr(lw), _ptr(k), _ptr(modes)),
                    "spr_lw3_kappa_all")
        return lw, k.reshape(3, 3), modes

    def lw3_pairs(self, q0_index, q1_grid, sigma: float, prec: str = "fp32") -> np.ndarray:
        """Per-pair linewidth terms (n, Ndof) for the given (irreducible q0, q1) pairs."""
        q0 = np.ascontiguousarray(q0_index, np.int32)
        q1 = np.ascontiguousarray(q1_grid, np.int32)
        out = np.zeros((q0.size, self.nd))
        self._check(self.lib.spr_lw3_pairs(self.h, ct.c_int(q0.size), _ptr(q0), _ptr(q1), ct.c_double(sigma),
                                           ct.c_int(PREC[prec]), _ptr(out)), "spr_lw3_pairs")
        return out

    def lw3_sampled(self, q0_index, q1_grid, weight, sigma: float, prec: str = "fp32") -> np.ndarray:
        """Weighted sums of the q1 terms per irreducible q0, (nq0, Ndof): row qi is the sum of
        weight[k] * lw3_pairs term k over the samples k with q0_index[k] == qi (spr_lw3_sampled)."""
        q0 = np.ascontiguousarray(q0_index, np.int32)
        q1 = np.ascontiguousarray(q1_grid, np.int32)
        w = np.ascontiguousarray(weight, np.float64)
        if not (q0.size == q1.size == w.size):
            raise ValueError("lw3_sampled: q0_index, q1_grid and weight differ in length")
        est = np.zeros((self.nq0, self.nd))
        self._check(self.lib.spr_lw3_sampled(self.h, ct.c_int(q0.size), _ptr(q0), _ptr(q1), _ptr(w),
                                             ct.c_double(sigma), ct.c_int(PREC[prec]), _ptr(est)), "spr_lw3_sampled")
        return est

    def lw3_pairs_noalloc(self, q0_index, q1_grid, sigma: float, prec: str = "fp32") -> np.ndarray:
        """lw3_pairs through the entry point that allocates no device memory (at most
        SPR_NOALLOC_MAX_PAIRS = 4096 pairs per call)."""
        q0 = np.ascontiguousarray(q0_index, np.int32)
        q1 = np.ascontiguousarray(q1_grid, np.int32)
        out = np.zeros((q0.size, self.nd))
        self._check(self.lib.spr_lw3_pairs_noalloc(self.h, ct.c_int(q0.size), _ptr(q0), _ptr(q1), ct.c_double(sigma),
                                                   ct.c_int(PREC[prec]), _ptr(out)), "spr_lw3_pairs_noalloc")
        return out

    def lw3_offmesh(self, model, q0_cart, ph0, T: float, delta: str = "tetra", sigma: float = 0.1,
                    zero_centered: bool = True, want_w: bool = False):
        """LineWidth.x's linewidth of an off-mesh q0 (spr_lw3_offmesh).  ph0: dict or object with
        evec (Ndof, Ndof) [s][row], omega, inv_omega of q0; model gives the crystal, lattice and
        FC2 for the q2 phonons.  Returns lw (Ndof) [and W[q1][s0][s1][s2]]."""
        cr, fc2 = model.crystal, model.fc2
        get = (lambda k: ph0[k]) if isinstance(ph0, dict) else (lambda k: getattr(ph0, k))
        e0 = np.ascontiguousarray(get("evec"), np.complex128)
        w0 = np.ascontiguousarray(get("omega"), np.float64)
        iw0 = np.ascontiguousarray(get("inv_omega"), np.float64)
        q0 = np.ascontiguousarray(q0_cart, np.float64)
        mass = np.ascontiguousarray(cr.mass, np.float64)
        crs = _Crystal(cr.nb, _ptr(mass), cr.vol)
        an = np.ascontiguousarray(fc2.atm_num, np.int32)
        ai = np.ascontiguousarray(fc2.atm_indx, np.int32)
        fc = np.ascontiguousarray(fc2.fc, np.float64)
        f2 = _IFC(2, fc2.natm_max, _ptr(an), _ptr(ai), _ptr(fc))
        lat = np.ascontiguousarray(cr.latvec, np.float64)
        lw = np.zeros(self.nd)
        W = np.zeros((self.nq, self.nd, self.nd, self.nd)) if want_w else None
        self._check(self.lib.spr_lw3_offmesh(self.h, _ptr(q0), _ptr(e0), _ptr(w0), _ptr(iw0), ct.byref(crs), _ptr(lat),
                                             ct.byref(f2), ct.c_double(T), ct.c_int(int(zero_centered)),
                                             ct.c_int(1 if delta == "tetra" else 0), ct.c_double(sigma), _ptr(lw),
                                             _ptr(W) if want_w else None), "spr_lw3_offmesh")
        return (lw, W) if want_w else lw

    def lw3_tetra(self, i0: int):
        """Tetrahedron-method linewidth and P3_q0 of irreducible q0 number i0 (FP64)."""
        lw = np.zeros(self.nd)
        p3 = np.zeros(self.nd)
        self._check(self.lib.spr_lw3_tetra(self.h, ct.c_int(i0), _ptr(lw), _ptr(p3)), "spr_lw3_tetra")
        return lw, p3

    def lw3_tetra_all(self):
        lw = np.zeros((self.nq0, self.nd))
        p3 = np.zeros((self.nq0, self.nd))
        self._check(self.lib.spr_lw3_tetra_all(self.h, _ptr(lw), _ptr(p3)), "spr_lw3_tetra_all")
        return lw, p3

    def lw4_tetra(self, i0: int):
        """Four-phonon tetrahedron-method linewidth sum of irreducible q0 number i0 (FP64)."""
        lw = np.zeros(self.nd)
        self._check(self.lib.spr_lw4_tetra(self.h, ct.c_int(i0), _ptr(lw)), "spr_lw4_tetra")
        return lw

    def lw4_tetra_all(self):
        lw = np.zeros((self.nq0, self.nd))
        self._check(self.lib.spr_lw4_tetra_all(self.h, _ptr(lw)), "spr_lw4_tetra_all")
        return lw

    # ---- iterative Boltzmann transport (time-reversal irreducible points) ----
    def bte_irr(self):
        """(grid index of every irreducible point, number of grid points it stands for)."""
        n = ct.c_int()
        self._check(self.lib.spr_bte_irr(self.h, ct.byref(n), None, None), "spr_bte_irr")
        irr = np.zeros(n.value, np.int32)
        deg = np.zeros(n.value, np.int32)
        self._check(self.lib.spr_bte_irr(self.h, ct.byref(n), irr.ctypes.data_as(ct.c_void_p),
                                         deg.ctypes.data_as(ct.c_void_p)), "spr_bte_irr")
        return irr, deg

    def bte_tau_inv(self):
        irr, _ = self.bte_irr()
        out = np.zeros((len(irr), self.nd))
        self._check(self.lib.spr_bte_tau_inv(self.h, _ptr(out)), "spr_bte_tau_inv")
        return out

    def bte_delta_f(self, F):
        F = np.ascontiguousarray(F, np.float64)
        dF = np.zeros_like(F)
        self._check(self.lib.spr_bte_delta_f(self.h, _ptr(F), _ptr(dF)), "spr_bte_delta_f")
        return dF

    def bte_solve(self, T: float, eps_k: float = 1e-4, max_iter: int = 100):
        irr, _ = self.bte_irr()
        k = np.zeros(9)
        k0 = np.zeros(9)
        it = ct.c_int()
        F = np.zeros((len(irr), self.nd, 3))
        self._check(self.lib.spr_bte_solve(self.h, ct.c_double(T), ct.c_double(eps_k), ct.c_int(max_iter),
                                           _ptr(k), _ptr(k0), ct.byref(it), _ptr(F)), "spr_bte_solve")
        return dict(kappa=k.reshape(3, 3), kappa0=k0.reshape(3, 3), iterations=it.value, F=F)

    def kernel_time(self):
        n = ct.c_int()
        ms = ct.c_double()
        self._check(self.lib.spr_kernel_time(self.h, ct.byref(n), ct.byref(ms)), "spr_kernel_time")
        return n.value, ms.value

    def last_launch(self):
        """(variant:shape, launch mode, units per item) of the last work-item launch."""
        buf = ct.create_string_buffer(128)
        mode, upi = ct.c_int(), ct.c_int()
        self._check(self.lib.spr_last_launch(self.h, buf, 128, ct.byref(mode), ct.byref(upi)), "spr_last_launch")
        return buf.value.decode(), mode.value, upi.value

    def kernel_time_reset(self):
        self._check(self.lib.spr_kernel_time_reset(self.h), "spr_kernel_time_reset")


def get_phonon(cr, fc2, q_cart, q_int, T: float, device: int = 0):
    """Phonons of every q on the GPU (spr_get_phonon); the same fields and layout as
    phonons.solve, which is the host path it replaces."""
    from .phonons import Phonons
    lib = load()
    q_cart = np.ascontiguousarray(q_cart, dtype=np.float64).reshape(-1, 3)
    nq, nd = q_cart.shape[0], cr.ndof
    gamma = np.ascontiguousarray(np.all(np.asarray(q_int) == 0, axis=1), dtype=np.int32)
    mass = np.ascontiguousarray(cr.mass, dtype=np.float64)
    lat = np.ascontiguousarray(cr.latvec, dtype=np.float64)
    num = np.ascontiguousarray(fc2.atm_num, dtype=np.int32)
    indx = np.ascontiguousarray(fc2.atm_indx, dtype=np.int32)
    fc = np.ascontiguousarray(fc2.fc, dtype=np.float64)
    c = _Crystal(cr.nb, _ptr(mass), cr.vol)
    f = _IFC(fc2.order, fc2.natm_max, _ptr(num), _ptr(indx), _ptr(fc))
    evec = np.zeros((nq, nd, nd), np.complex128)
    omega, inv_omega, nbe = (np.zeros((nq, nd)) for _ in range(3))
    vel = np.zeros((nq, nd, 3))
    rc = lib.spr_get_phonon(ct.byref(c), _ptr(lat), ct.byref(f), ct.c_int(nq), _ptr(q_cart), _ptr(gamma),
                            ct.c_double(T), ct.c_int(device), _ptr(evec), _ptr(omega), _ptr(inv_omega),
                            _ptr(nbe), _ptr(vel))
    if rc != 0:
        raise SprError(f"spr_get_phonon failed ({rc}): {lib.spr_last_error().decode()}")
    return Phonons(evec, omega, inv_omega, nbe, vel, float(T))


def mesh_build(cr, mesh, time_rev: int = 0, symmetry: bool = True, symprec: float = 1e-2):
    """The q mesh from the crystal on the host side of the library (spr_mesh_build): a dict with
    q_int (nq, 3), q_cart (nq, 3), irr_idx (nq0), pos (nq), multiplicity (nq0), rot (nsym, 3, 3)
    Cartesian, tetra (6 nq, 4), all 0-based; the same tables as mesh.make_mesh."""
    lib = load()
    lat = np.ascontiguousarray(cr.latvec, np.float64)
    frac = np.ascontiguousarray(cr.basis_frac, np.float64)
    typ = np.ascontiguousarray(cr.species, np.int32)
    m = np.ascontiguousarray(mesh, np.int32)
    nq0, nsym = ct.c_int(), ct.c_int()
    args = (_ptr(lat), ct.c_int(cr.nb), _ptr(frac), _ptr(typ), _ptr(m), ct.c_int(int(time_rev)),
            ct.c_int(int(symmetry)), ct.c_double(symprec), ct.byref(nq0), ct.byref(nsym))
    if lib.spr_mesh_build(*args, *([None] * 7)) != 0:
        raise SprError(f"spr_mesh_build failed: {lib.spr_last_error().decode()}")
    nq = int(np.prod(m))
    out = dict(q_int=np.zeros((nq, 3), np.int32), q_cart=np.zeros((nq, 3)), irr_idx=np.zeros(nq0.value, np.int32),
               pos=np.zeros(nq, np.int32), multiplicity=np.zeros(nq0.value, np.int32),
               rotf=np.zeros((nsym.value, 3, 3)), tetra=np.zeros((6 * nq, 4), np.int32))
    keys = ("q_int", "q_cart", "irr_idx", "pos", "multiplicity", "rotf", "tetra")
    if lib.spr_mesh_build(*args, *[_ptr(out[k]) for k in keys]) != 0:
        raise SprError(f"spr_mesh_build failed: {lib.spr_last_error().decode()}")
    out["rot"] = np.transpose(out.pop("rotf"), (0, 2, 1)).copy()   # Rxyz(3,3,k) -> [k][i][j]
    return out


class Mlp:
    """Device MLP (spr_mlp_*): F -> 128 -> 128 -> 1, tanh."""

    def __init__(self, n_features: int, seed: int = 0, device: int = 0):
        self.lib = load()
        self.F = n_features
        h = ct.c_void_p()
        self._check(self.lib.spr_mlp_create(ct.c_int(n_features), ct.c_ulonglong(seed), ct.c_int(device),
                                            ct.byref(h)), "spr_mlp_create")
        self.h = h
        n = ct.c_long()
        self._check(self.lib.spr_mlp_num_params(self.h, ct.byref(n)), "spr_mlp_num_params")
        self.n_params = n.value

    def _check(self, rc, what):
        if rc != 0:
            raise SprError(f"{what} failed ({rc}): {self.lib.spr_last_error().decode()}")

    def close(self):
        if getattr(self, "h", None):
            self.lib.spr_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def params(self) -> np.ndarray:
        p = np.zeros(self.n_params, np.float32)
        self._check(self.lib.spr_mlp_get_params(self.h, _ptr(p)), "spr_mlp_get_params")
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        assert p.size == self.n_params
        self._check(self.lib.spr_mlp_set_params(self.h, _ptr(p)), "spr_mlp_set_params")

    def grad(self, X, y):
        X = np.ascontiguousarray(X, np.float32)
        y = np.ascontiguousarray(y, np.float32)
        g = np.zeros(self.n_params, np.float32)
        loss = ct.c_double()
        self._check(self.lib.spr_mlp_grad(self.h, _ptr(X), _ptr(y), ct.c_long(y.size), _ptr(g), ct.byref(loss)),
                    "spr_mlp_grad")
        return g, loss.value

    def train(self, X, y, epochs: int, batch: int = 4096, lr: float = 1e-3):
        X = np.ascontiguousarray(X, np.float32)
        y = np.ascontiguousarray(y, np.float32)
        losses = np.zeros(epochs)
        self._check(self.lib.spr_mlp_train(self.h, _ptr(X), _ptr(y), ct.c_long(y.size), ct.c_int(epochs),
                                           ct.c_int(batch), ct.c_float(lr), _ptr(losses)), "spr_mlp_train")
        return losses

    def score(self, ctx: "GpuContext", q0_index, sigma: float, window: float, y_std: float, wmax: float):
        """Device proposal scores S(q1) (n, nq) for irreducible q0 indices and the evaluation count."""
        return score(ctx, self, q0_index, sigma, window, y_std, wmax)

    def predict(self, X, prec: str = "fp32") -> np.ndarray:
        X = np.ascontiguousarray(X, np.float32)
        out = np.zeros(X.shape[0], np.float32)
        self._check(self.lib.spr_mlp_predict(self.h, _ptr(X), ct.c_long(X.shape[0]), ct.c_int(PREC[prec]),
                                             _ptr(out)), "spr_mlp_predict")
        return out


def pair_scores(ctx: GpuContext, mlp: "Mlp | None", q0_index, sigma: float, window: float, wmax: float,
                want_x: bool = False):
    """spr_pair_scores: per (q0, q1) the resonance weight S (mlp=None) or the network's
    normalised output (n, nq), and optionally the pair feature rows (n, nq, 12)."""
    lib = load()
    q0 = np.ascontiguousarray(q0_index, np.int32)
    S = np.zeros((q0.size, ctx.nq))
    X = np.zeros((q0.size, ctx.nq, 12), np.float32) if want_x else None
    rc = lib.spr_pair_scores(ctx.h, mlp.h if mlp is not None else None, ct.c_int(q0.size), _ptr(q0),
                             ct.c_double(sigma), ct.c_double(window), ct.c_float(wmax), _ptr(S),
                             _ptr(X) if want_x else None)
    if rc != 0:
        raise SprError(f"spr_pair_scores failed ({rc}): {lib.spr_last_error().decode()}")
    return (S, X) if want_x else S


def score(ctx: GpuContext, mlp: "Mlp | None", q0_index, sigma: float, window: float, y_std: float = 1.0,
          wmax: float = 1.0):
    """spr_mlp_score: S(q1) (n, nq) for irreducible q0 indices and the number of network
    evaluations; mlp=None scores |G| alone."""
    lib = load()
    q0 = np.ascontiguousarray(q0_index, np.int32)
    S = np.zeros((q0.size, ctx.nq))
    ne = ct.c_longlong()
    rc = lib.spr_mlp_score(ctx.h, mlp.h if mlp is not None else None, ct.c_int(q0.size), _ptr(q0),
                           ct.c_double(sigma), ct.c_double(window), ct.c_float(y_std), ct.c_float(wmax), _ptr(S),
                           ct.byref(ne))
    if rc != 0:
        raise SprError(f"spr_mlp_score failed ({rc}): {lib.spr_last_error().decode()}")
    return S, ne.value
