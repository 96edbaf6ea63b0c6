"""ctypes bindings of the CPU oracle (TEST INFRASTRUCTURE ONLY).

``Oracle`` wraps ``oracle/libwmoracle.so`` (the FP64 C restatement of the
reference's hot path, built by oracle/Makefile).  ``RefBuild`` wraps
``oracle/_ref/libwmref.so``, the reference's own common/particle.f90 and
common/sort.f90 compiled unchanged by oracle/build_ref.sh; it exists only
where the reference tree was present at build time.

Only tests/, the smoke check and bench.py's CPU-baseline leg import this
module; the product never does.
"""
from __future__ import annotations

import ctypes
import os
import shutil
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_LIB = os.path.join(HERE, "libwmoracle.so")
REF_LIB = os.path.join(HERE, "_ref", "libwmref.so")
REF_LIB_OMP = os.path.join(HERE, "_ref", "libwmref_omp.so")   # OpenMP loops on

OR_OK, OR_ERR_CG_ITEMAX, OR_ERR_NP_OVER, OR_ERR_RANGE, OR_ERR_ARG = 0, 1, 2, 3, 4
MAXSP = 8


class OrParams(ctypes.Structure):
    _fields_ = [("ndim", ctypes.c_int), ("np", ctypes.c_int), ("nsp", ctypes.c_int),
                ("nxgs", ctypes.c_int), ("nxge", ctypes.c_int), ("nygs", ctypes.c_int),
                ("nyge", ctypes.c_int), ("nys", ctypes.c_int), ("nye", ctypes.c_int),
                ("delx", ctypes.c_double), ("delt", ctypes.c_double), ("c", ctypes.c_double),
                ("gfac", ctypes.c_double), ("q", ctypes.c_double * MAXSP),
                ("r", ctypes.c_double * MAXSP), ("nxs", ctypes.c_int), ("nxe", ctypes.c_int),
                ("bc", ctypes.c_int), ("u0", ctypes.c_double)]


BC_PERIODIC, BC_SHOCK, BC_RECONNECTION = 0, 1, 2


def _p(a):
    if a is None:
        return None
    if not a.flags.f_contiguous:
        raise OracleError("arrays must be Fortran-contiguous (the reference layout)")
    return a.ctypes.data


class OracleError(RuntimeError):
    pass


class Oracle:
    """FP64 restatement; one instance = one grid and its SAVE state."""

    def __init__(self, p: dict, threads: int = 1, lib: str | None = None):
        path = lib or ORACLE_LIB
        if not os.path.exists(path):
            raise OracleError(f"{path} not built (run make -C oracle)")
        self.lib = ctypes.CDLL(path)
        L = self.lib
        vp = ctypes.c_void_p
        L.or_state_new.restype = vp
        L.or_state_new.argtypes = [vp]
        L.or_state_free.argtypes = [vp]
        for n in ("or_state_df", "or_state_uj", "or_state_gkl"):
            getattr(L, n).restype = vp
            getattr(L, n).argtypes = [vp]
        L.or_state_cg_iters.argtypes = [vp, vp]
        L.or_state_set_cg_err.argtypes = [vp, ctypes.c_double]
        L.or_add_rd.restype = ctypes.c_double
        L.or_add_rd.argtypes = [ctypes.c_double, ctypes.c_double]
        L.or_cgm.argtypes = [vp, vp, vp, ctypes.c_double, vp]
        L.or_set_threads.argtypes = [ctypes.c_int]
        L.or_bc_phi_l.argtypes = [vp, vp, ctypes.c_int]
        L.or_bc_injection.argtypes = [vp, vp, vp, ctypes.c_double]
        for n, k in (("or_particle_solv", 5), ("or_ele_cur", 5), ("or_bc_curre", 2),
                     ("or_bc_dfield", 2), ("or_bc_phi", 2), ("or_field_rhs", 4),
                     ("or_field_fdtd_i", 6), ("or_field_advance", 4), ("or_bc_particle_x", 3),
                     ("or_bc_particle_y", 4),
                     ("or_sort_bucket", 5), ("or_energy", 5), ("or_step", 7),
                     ("or_mom_calc_accl", 5), ("or_mom_calc_nvt", 4), ("or_bc_mom", 2)):
            getattr(L, n).argtypes = [vp] * k
        self.p = p
        self.par = OrParams()
        for k in ("ndim", "np", "nsp", "nxgs", "nxge", "nygs", "nyge", "nys", "nye"):
            setattr(self.par, k, int(p[k]))
        for k in ("delx", "delt", "c", "gfac"):
            setattr(self.par, k, float(p[k]))
        for s in range(p["nsp"]):
            self.par.q[s] = float(p["q"][s])
            self.par.r[s] = float(p["r"][s])
        # boundary family and active x range (the whole grid unless given)
        self.par.bc = int(p.get("bc", BC_PERIODIC))
        self.par.nxs = int(p.get("nxs", p["nxgs"]))
        self.par.nxe = int(p.get("nxe", p["nxge"]))
        self.par.u0 = float(p.get("u0", 0.0))
        self.pp = ctypes.addressof(self.par)
        self.state = L.or_state_new(self.pp)
        if not self.state:
            raise OracleError("oracle state allocation failed")
        L.or_set_threads(int(threads))

    def __del__(self):
        try:
            self.lib.or_state_free(self.state)
        except Exception:
            pass

    def _ck(self, st, what):
        if st != OR_OK:
            raise OracleError(f"{what}: oracle status {st}")

    def set_threads(self, n: int):
        self.lib.or_set_threads(int(n))

    # reference routines (same argument order as the Fortran)
    def particle__solv(self, gp, up, uf, cumcnt):
        self._ck(self.lib.or_particle_solv(self.pp, _p(gp), _p(up), _p(uf), _p(cumcnt)), "particle__solv")

    def ele_cur(self, uj, up, gp, cumcnt):
        self._ck(self.lib.or_ele_cur(self.pp, _p(uj), _p(up), _p(gp), _p(cumcnt)), "ele_cur")

    def bc_curre(self, uj):
        self.lib.or_bc_curre(self.pp, _p(uj))

    def bc_dfield(self, df):
        self.lib.or_bc_dfield(self.pp, _p(df))

    def bc_phi(self, phi, l: int = 1):
        self.lib.or_bc_phi_l(self.pp, _p(phi), int(l))

    def bc_injection(self, up, np2, u0: float):
        self._ck(self.lib.or_bc_injection(self.pp, _p(up), _p(np2), ctypes.c_double(u0)),
                 "injection")

    def set_active_x(self, nxs: int, nxe: int):
        """The active x range [nxs, nxe] (the shock's box grows by relocate)."""
        self.par.nxs, self.par.nxe = int(nxs), int(nxe)

    def field_rhs(self, gkl, uf, uj):
        self.lib.or_field_rhs(self.pp, _p(gkl), _p(uf), _p(uj))

    def cgm(self, db, gkl, err=1e-6):
        it = np.zeros(3, dtype=np.int32)
        st = self.lib.or_cgm(self.pp, _p(db), _p(gkl), ctypes.c_double(err), _p(it))
        return st, it

    def field__fdtd_i(self, uf, up, gp, cumcnt):
        self._ck(self.lib.or_field_fdtd_i(self.pp, ctypes.c_void_p(self.state), _p(uf), _p(up),
                                          _p(gp), _p(cumcnt)), "field__fdtd_i")

    def field_advance(self, uf, uj):
        """field.f90:125-184 from a folded current uj (ghosts filled)."""
        self._ck(self.lib.or_field_advance(self.pp, ctypes.c_void_p(self.state), _p(uf), _p(uj)),
                 "field_advance")

    def particle_x(self, up, np2):
        self._ck(self.lib.or_bc_particle_x(self.pp, _p(up), _p(np2)), "particle_x")

    def particle_y(self, up, np2):
        return self.lib.or_bc_particle_y(self.pp, ctypes.c_void_p(self.state), _p(up), _p(np2))

    def sort__bucket(self, gp, up, cumcnt, np2):
        self._ck(self.lib.or_sort_bucket(self.pp, _p(gp), _p(up), _p(cumcnt), _p(np2)), "sort__bucket")

    def mom_calc_accl(self, gp, up, uf, cumcnt):
        self._ck(self.lib.or_mom_calc_accl(self.pp, _p(gp), _p(up), _p(uf), _p(cumcnt)),
                 "mom_calc__accl")

    def mom_calc_nvt(self, mom, up, np2):
        self._ck(self.lib.or_mom_calc_nvt(self.pp, _p(mom), _p(up), _p(np2)), "mom_calc__nvt")

    def bc_mom(self, mom):
        self.lib.or_bc_mom(self.pp, _p(mom))

    def energy(self, up, uf, np2):
        out = np.zeros(self.p["nsp"] + 2)
        self.lib.or_energy(self.pp, _p(up), _p(uf), _p(np2), _p(out))
        return out

    def step(self, up, gp, uf, cumcnt, np2):
        self._ck(self.lib.or_step(self.pp, ctypes.c_void_p(self.state), _p(up), _p(gp), _p(uf),
                                  _p(cumcnt), _p(np2)), "step")

    def cg_iters(self):
        it = np.zeros(3, dtype=np.int32)
        self.lib.or_state_cg_iters(ctypes.c_void_p(self.state), _p(it))
        return it

    def set_cg_err(self, err):
        self.lib.or_state_set_cg_err(ctypes.c_void_p(self.state), ctypes.c_double(err))

    def _view(self, ptr, shape):
        n = int(np.prod(shape))
        buf = (ctypes.c_double * n).from_address(ptr)
        return np.ndarray(shape, dtype=np.float64, buffer=buf, order="F")

    def df(self):
        p = self.p
        return self._view(self.lib.or_state_df(ctypes.c_void_p(self.state)),
                          (6, p["nx"] + 4, p["ny"] + 4))

    def uj(self):
        p = self.p
        return self._view(self.lib.or_state_uj(ctypes.c_void_p(self.state)),
                          (3, p["nx"] + 4, p["ny"] + 4))

    def add_rd(self, a, b):
        return self.lib.or_add_rd(a, b)


def ref_available(omp: bool = False) -> bool:
    return os.path.exists(REF_LIB_OMP if omp else REF_LIB)


class RefBuild:
    """The reference's particle.f90 / sort.f90 / mom_calc.f90, compiled unchanged.

    Module procedures are called through their compiler symbols with every
    argument by reference.  The reference's *_init routines allocate module
    arrays once per process, so every instance loads its own private copy of
    the shared object."""

    def __init__(self, p: dict, omp: bool = False):
        if not ref_available(omp):
            raise OracleError("oracle/_ref/libwmref%s.so not built" % ("_omp" if omp else ""))
        self._tmp = tempfile.NamedTemporaryFile(suffix=".so", delete=False)
        self._tmp.close()
        shutil.copyfile(REF_LIB_OMP if omp else REF_LIB, self._tmp.name)
        self.lib = ctypes.CDLL(self._tmp.name, mode=os.RTLD_LOCAL)
        vp = ctypes.c_void_p
        self.lib._QMparticlePparticle__init.argtypes = [vp] * 14
        self.lib._QMparticlePparticle__solv.argtypes = [vp] * 6
        self.lib._QMsortPsort__init.argtypes = [vp] * 9
        self.lib._QMsortPsort__bucket.argtypes = [vp] * 6
        self.lib._QMmom_calcPmom_calc__init.argtypes = [vp] * 14
        self.lib._QMmom_calcPmom_calc__accl.argtypes = [vp] * 6
        self.lib._QMmom_calcPmom_calc__nvt.argtypes = [vp] * 3
        self.p = p
        self._keep = []
        self._q = np.ascontiguousarray(p["q"], dtype=np.float64)
        self._r = np.ascontiguousarray(p["r"], dtype=np.float64)
        g = [self.I(p[k]) for k in ("ndim", "np", "nsp", "nxgs", "nxge", "nygs", "nyge", "nys", "nye")]
        self.lib._QMparticlePparticle__init(*g, self.D(p["delx"]), self.D(p["delt"]),
                                            self.D(p["c"]), _p(self._q), _p(self._r))
        self.lib._QMsortPsort__init(*g)
        self.lib._QMmom_calcPmom_calc__init(*g, self.D(p["delx"]), self.D(p["delt"]),
                                            self.D(p["c"]), _p(self._q), _p(self._r))

    def I(self, v):
        o = ctypes.c_int(int(v))
        self._keep.append(o)
        return ctypes.addressof(o)

    def D(self, v):
        o = ctypes.c_double(float(v))
        self._keep.append(o)
        return ctypes.addressof(o)

    def __del__(self):
        try:
            os.unlink(self._tmp.name)
        except Exception:
            pass

    def particle__solv(self, gp, up, uf, cumcnt):
        I = self.I
        self.lib._QMparticlePparticle__solv(_p(gp), _p(up), _p(uf), _p(cumcnt),
                                            I(self.p["nxgs"]), I(self.p["nxge"]))

    def mom_calc__accl(self, gp, up, uf, cumcnt):
        I = self.I
        self.lib._QMmom_calcPmom_calc__accl(_p(gp), _p(up), _p(uf), _p(cumcnt),
                                            I(self.p["nxgs"]), I(self.p["nxge"]))

    def mom_calc__nvt(self, mom, up, np2):
        self.lib._QMmom_calcPmom_calc__nvt(_p(mom), _p(up), _p(np2))

    def sort__bucket(self, gp, up, cumcnt, np2):
        I = self.I
        self.lib._QMsortPsort__bucket(_p(gp), _p(up), _p(cumcnt), _p(np2), I(self.p["nxgs"]),
                                      I(self.p["nxge"]))
