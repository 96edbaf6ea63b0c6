"""wumingpic2d_amd -- MI355X-native hot path of WumingPIC2D.

The package is a thin host layer over ``libwmpic.so`` (hand-written HIP
kernels for gfx950 behind the C ABI of ``include/wmpic.h``).  ``Context``
mirrors the reference's module procedures one for one (same names, argument
order and meaning, same fatal-error messages raised as :class:`WmError`):

=====================================  ======================================
reference (WumingCode/WumingPIC2D)     Context method
=====================================  ======================================
common/particle.f90:18   particle__init          ``particle__init``
common/particle.f90:48   particle__solv          ``particle__solv``
common/field.f90:22      field__init             ``field__init``
common/field.f90:66      field__fdtd_i           ``field__fdtd_i``
common/sort.f90:16       sort__init              ``sort__init``
common/sort.f90:36       sort__bucket            ``sort__bucket``
common/boundary_periodic.f90:26  __init          ``boundary_periodic__init``
common/boundary_periodic.f90:61  __particle_x    ``boundary_periodic__particle_x``
common/boundary_periodic.f90:99  __particle_y    ``boundary_periodic__particle_y``
common/boundary_periodic.f90:251 __dfield        ``boundary_periodic__dfield``
common/boundary_periodic.f90:357 __curre         ``boundary_periodic__curre``
common/boundary_periodic.f90:511 __phi           ``boundary_periodic__phi``
common/mpi_set.f90:19    mpi_set__init           :func:`mpi_set__init`
=====================================  ======================================

Arrays are numpy arrays in Fortran order with the reference's shapes, e.g.
``up`` is ``(ndim, np, ny, nsp)`` float64 and ``cumcnt`` is
``(nx+1, ny, nsp)`` int32.  The device-resident methods (``state_upload``,
``init_weibel``, ``step``, ``energy``) keep the whole run in HBM.

There is no CPU fallback: importing works without a GPU, but every compute
call needs the built library and a visible device and raises otherwise.
"""
from __future__ import annotations

import ctypes
import os
from typing import Sequence

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libwmpic.so")

WM_OK = 0
WM_ERR_NOT_INIT = 1
WM_ERR_CG_ITEMAX = 2
WM_ERR_NP_OVER = 3
WM_ERR_RANGE = 4
WM_ERR_ARG = 5
WM_ERR_HIP = 6
WM_ERR_NOMEM = 7
WM_ERR_NODEV = 8
BC_PERIODIC = 0
BC_SHOCK = 1
BC_RECONNECTION = 2
FORM_FP64 = 0
FORM_FP32 = 1
SOLVER_CG = 0
SOLVER_FCG = 1

# Symbols declared in include/wmpic.h (checked by the CPU test suite).
EXPORTS = (
    "wm_create", "wm_destroy", "wm_last_error", "wm_set_capacity", "wm_set_cg_tolerance",
    "wm_set_form",
    "wm_mpi_set_init", "wm_particle_init", "wm_field_init", "wm_sort_init", "wm_bcp_init",
    "wm_particle_solv", "wm_field_fdtd_i", "wm_bcp_particle_x", "wm_bcp_particle_y",
    "wm_sort_bucket", "wm_bcp_dfield", "wm_bcp_curre", "wm_bcp_phi", "wm_state_upload",
    "wm_state_download", "wm_state_init_weibel", "wm_step", "wm_energy", "wm_cg_iters",
    "wm_phase_times", "wm_particle_count", "wm_capacity_stats", "wm_id_checksum",
    "wm_set_solver", "wm_pc_collect", "wm_pc_pair_count", "wm_pc_init", "wm_pc_train",
    "wm_pc_num_params", "wm_pc_params", "wm_pc_apply", "wm_pc_apply_noalloc", "wm_pc_grad",
    "wm_solve_stats",
    "wm_mom_calc_init", "wm_mom_calc_accl", "wm_mom_calc_nvt", "wm_bcp_mom", "wm_moments",
    "wm_paraio_init", "wm_paraio_output", "wm_paraio_input", "wm_paraio_output_state",
    "wm_paraio_input_state", "wm_paraio_param", "wm_paraio_mom", "wm_paraio_ptcl", "wm_paraio_orb",
    "wm_bcs_init", "wm_bcs_particle_x", "wm_bcs_injection", "wm_bcs_particle_y", "wm_bcs_dfield",
    "wm_bcs_curre", "wm_bcs_phi", "wm_bcs_mom", "wm_bcr_init", "wm_bcr_particle_x",
    "wm_bcr_particle_y", "wm_bcr_dfield", "wm_bcr_curre", "wm_bcr_phi", "wm_bcr_mom",
    "wm_set_walls",
    # include/wmpic_json.h: the JSON model behind the Fortran json_module
    "wm_json_new_object", "wm_json_load", "wm_json_parse", "wm_json_free", "wm_json_child",
    "wm_json_get_int", "wm_json_get_real", "wm_json_get_bool", "wm_json_get_str",
    "wm_json_array_size", "wm_json_add_child", "wm_json_add_int", "wm_json_add_real",
    "wm_json_add_bool", "wm_json_add_str", "wm_json_add_ints", "wm_json_add_reals",
    "wm_json_update_str", "wm_json_update_int", "wm_json_update_real", "wm_json_print",
    "wm_json_serialize",
)


class WmError(RuntimeError):
    """A non-zero status from libwmpic; ``code`` is the WM_ERR_* value."""

    def __init__(self, code: int, message: str):
        super().__init__(message)
        self.code = code


_lib = None


def load_library(path: str | None = None) -> ctypes.CDLL:
    """Load libwmpic.so (raises if it has not been built)."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise WmError(WM_ERR_NODEV, f"libwmpic is not built ({p} missing); run __graft_entry__.build()")
    lib = ctypes.CDLL(p)
    _declare(lib)
    if path is None:
        _lib = lib
    return lib


_P = ctypes.c_void_p
_I = ctypes.c_int
_D = ctypes.c_double


def _declare(lib: ctypes.CDLL) -> None:
    sig = {
        "wm_create": [ctypes.POINTER(_P), _I],
        "wm_destroy": [_P],
        "wm_set_capacity": [_P, _I, _I],
        "wm_set_cg_tolerance": [_P, _D],
        "wm_set_form": [_P, _I],
        "wm_mpi_set_init": [_I, _I, _I] + [ctypes.POINTER(_I)] * 6,
        "wm_particle_init": [_P] + [_I] * 9 + [_D, _D, _D, _P, _P],
        "wm_field_init": [_P] + [_I] * 13 + [_D, _D, _D, _P, _P, _D],
        "wm_sort_init": [_P] + [_I] * 9,
        "wm_bcp_init": [_P] + [_I] * 15 + [_P, _I, _D, _D, _D],
        "wm_particle_solv": [_P, _P, _P, _P, _P, _I, _I],
        "wm_field_fdtd_i": [_P, _P, _P, _P, _P, _I, _I, _I],
        "wm_bcp_particle_x": [_P, _P, _P],
        "wm_bcp_particle_y": [_P, _P, _P],
        "wm_sort_bucket": [_P, _P, _P, _P, _P, _I, _I],
        "wm_bcp_dfield": [_P, _P] + [_I] * 6,
        "wm_bcp_curre": [_P, _P] + [_I] * 6,
        "wm_bcp_phi": [_P, _P] + [_I] * 5,
        "wm_state_upload": [_P, _P, _P, _P, _P],
        "wm_state_download": [_P, _P, _P, _P, _P],
        "wm_state_init_weibel": [_P, ctypes.c_uint64, _I, _D, _D],
        "wm_step": [_P, _I],
        "wm_energy": [_P, _P],
        "wm_cg_iters": [_P, _P],
        "wm_phase_times": [_P, _P, _I],
        "wm_particle_count": [_P, ctypes.POINTER(ctypes.c_longlong)],
        "wm_id_checksum": [_P, _P],
        "wm_capacity_stats": [_P, _P, _I],
        "wm_set_solver": [_P, _I],
        "wm_pc_collect": [_P, _I, _I],
        "wm_pc_pair_count": [_P],
        "wm_pc_init": [_P, ctypes.c_uint64],
        "wm_pc_train": [_P, _I, _D, ctypes.c_uint64, _P],
        "wm_pc_num_params": [],
        "wm_pc_params": [_P, _P, _I, _I],
        "wm_pc_apply": [_P, _P, _P],
        "wm_pc_apply_noalloc": [_P, _P, _P, _I, _I],
        "wm_pc_grad": [_P, _P, _P, _P, _P],
        "wm_solve_stats": [_P, _P, _P],
        "wm_mom_calc_init": [_P] + [_I] * 9 + [_D, _D, _D, _P, _P],
        "wm_mom_calc_accl": [_P, _P, _P, _P, _P, _I, _I],
        "wm_mom_calc_nvt": [_P, _P, _P, _P],
        "wm_bcp_mom": [_P, _P],
        "wm_moments": [_P, _P],
        "wm_paraio_init": [_P] + [_I] * 11 + [_D, _D, _D, _P, _P, ctypes.c_char_p],
        "wm_paraio_output": [_P, _P, _P, _P, _I, _I, _I, ctypes.c_char_p],
        "wm_paraio_input": [_P, _P, _P, _P, _P, _P, _P, ctypes.c_char_p],
        "wm_paraio_output_state": [_P, _I, ctypes.c_char_p],
        "wm_paraio_input_state": [_P, _P, ctypes.c_char_p],
        "wm_paraio_param": [_P, _I] + [_D] * 6 + [ctypes.c_char_p],
        "wm_paraio_mom": [_P, _P, _P, _I],
        "wm_paraio_ptcl": [_P, _P, _P, _P, _I],
        "wm_paraio_orb": [_P, _P, _P, _P, _I],
        "wm_set_walls": [_P, _I, _I, _D],
        "wm_bcs_injection": [_P, _P, _P, _I, _I, _D],
    }
    for fam in ("bcs", "bcr"):
        sig[f"wm_{fam}_init"] = sig["wm_bcp_init"]
        sig[f"wm_{fam}_particle_x"] = [_P, _P, _P, _I, _I]
        sig[f"wm_{fam}_particle_y"] = [_P, _P, _P]
        sig[f"wm_{fam}_dfield"] = sig["wm_bcp_dfield"]
        sig[f"wm_{fam}_curre"] = sig["wm_bcp_curre"]
        sig[f"wm_{fam}_phi"] = sig["wm_bcp_phi"]
        sig[f"wm_{fam}_mom"] = [_P, _P]
    for name, args in sig.items():
        f = getattr(lib, name)
        f.argtypes = args
        f.restype = _I
    lib.wm_last_error.argtypes = [_P]
    lib.wm_last_error.restype = ctypes.c_char_p


def _ptr(a: np.ndarray | None, dtype, name: str, n: int | None = None):
    """Address of a Fortran-ordered array of at least n elements (the C side
    reads and writes the reference's full shapes)."""
    if a is None:
        return None
    if not isinstance(a, np.ndarray) or a.dtype != np.dtype(dtype):
        raise TypeError(f"{name} must be a numpy array of {np.dtype(dtype)}")
    if not a.flags.f_contiguous:
        raise TypeError(f"{name} must be Fortran-contiguous (the reference's column-major layout)")
    if n is not None and a.size < n:
        raise ValueError(f"{name} holds {a.size} elements; the grid needs {n}")
    return a.ctypes.data


def mpi_set__init(nygs: int, nyge: int, nproc: int):
    """Single-rank ``mpi_set__init`` (common/mpi_set.f90:19-49).

    Returns ``dict(nys, nye, nup, ndown, nrank, nsize)``; ``nproc`` must be 1
    (the reference aborts on a rank-count mismatch, mpi_set.f90:30-34)."""
    lib = load_library()
    out = [_I() for _ in range(6)]
    st = lib.wm_mpi_set_init(nygs, nyge, nproc, *[ctypes.byref(o) for o in out])
    if st != WM_OK:
        raise WmError(st, "error in proc no.")
    keys = ("nys", "nye", "nup", "ndown", "nrank", "nsize")
    return {k: o.value for k, o in zip(keys, out)}


class Context:
    """One device, one grid: the reference's module SAVE state, in HBM."""

    def __init__(self, device: int = 0):
        self._lib = load_library()
        h = _P()
        st = self._lib.wm_create(ctypes.byref(h), int(device))
        if st != WM_OK:
            raise WmError(st, f"no usable GPU {device} (status {st})")
        self._h = h
        self.device = device
        self.grid: dict | None = None

    # ------------------------------------------------------------ plumbing
    def close(self) -> None:
        if getattr(self, "_h", None):
            self._lib.wm_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _check(self, st: int) -> None:
        if st != WM_OK:
            msg = self._lib.wm_last_error(self._h)
            raise WmError(st, msg.decode() if msg else f"status {st}")

    def set_capacity(self, cap_per_row: int, mover_cap: int = 0) -> None:
        """Device slots per (species, row) and migration buffer per row."""
        self._check(self._lib.wm_set_capacity(self._h, int(cap_per_row), int(mover_cap)))

    def set_form(self, form: int) -> None:
        """FORM_FP64 (reference arithmetic) or FORM_FP32 (int32 cell + FP32
        offsets and momenta; FP64 current accumulation, solve and sums)."""
        self._check(self._lib.wm_set_form(self._h, int(form)))

    def set_cg_tolerance(self, err: float) -> None:
        self._check(self._lib.wm_set_cg_tolerance(self._h, float(err)))

    def _n(self, kind: str) -> int | None:
        """Elements of the reference array `kind` on this context's grid."""
        g = self.grid
        if g is None:
            return None
        nx, ny, nsp = g["nx"], g["ny"], g["nsp"]
        return {"up": g["ndim"] * g["np"] * ny * nsp, "uf": 6 * (nx + 4) * (ny + 4),
                "uj": 3 * (nx + 4) * (ny + 4), "cumcnt": (nx + 1) * ny * nsp, "np2": ny * nsp,
                "phi": (nx + 2) * (ny + 2), "mom": 7 * (nx + 2) * (ny + 2) * nsp}[kind]

    def _a(self, a, kind: str, name: str | None = None):
        dtype = np.int32 if kind in ("cumcnt", "np2") else np.float64
        return _ptr(a, dtype, name or kind, self._n(kind))

    def _remember(self, ndim, np_, nsp, nxgs, nxge, nygs, nyge, nys, nye):
        self.grid = dict(ndim=ndim, np=np_, nsp=nsp, nxgs=nxgs, nxge=nxge, nygs=nygs,
                         nyge=nyge, nys=nys, nye=nye, nx=nxge - nxgs + 1, ny=nye - nys + 1)

    # ------------------------------------------------------ initialisers
    def particle__init(self, ndim, np_, nsp, nxgs, nxge, nygs, nyge, nys, nye, delx, delt, c,
                       q: Sequence[float], r: Sequence[float]) -> None:
        qa = np.ascontiguousarray(q, dtype=np.float64)
        ra = np.ascontiguousarray(r, dtype=np.float64)
        self._check(self._lib.wm_particle_init(self._h, ndim, np_, nsp, nxgs, nxge, nygs, nyge,
                                               nys, nye, delx, delt, c, qa.ctypes.data,
                                               ra.ctypes.data))
        self._remember(ndim, np_, nsp, nxgs, nxge, nygs, nyge, nys, nye)

    def field__init(self, ndim, np_, nsp, nxgs, nxge, nygs, nyge, nys, nye, mnpr, ncomw, opsum,
                    nerr, delx, delt, c, q, r, gfac) -> None:
        qa = np.ascontiguousarray(q, dtype=np.float64)
        ra = np.ascontiguousarray(r, dtype=np.float64)
        self._check(self._lib.wm_field_init(self._h, ndim, np_, nsp, nxgs, nxge, nygs, nyge, nys,
                                            nye, mnpr, ncomw, opsum, nerr, delx, delt, c,
                                            qa.ctypes.data, ra.ctypes.data, gfac))
        self._remember(ndim, np_, nsp, nxgs, nxge, nygs, nyge, nys, nye)

    def sort__init(self, ndim, np_, nsp, nxgs, nxge, nygs, nyge, nys, nye) -> None:
        self._check(self._lib.wm_sort_init(self._h, ndim, np_, nsp, nxgs, nxge, nygs, nyge, nys,
                                           nye))
        self._remember(ndim, np_, nsp, nxgs, nxge, nygs, nyge, nys, nye)

    def boundary_periodic__init(self, ndim, np_, nsp, nxgs, nxge, nygs, nyge, nys, nye, nup,
                                ndown, mnpi, mnpr, ncomw, nerr, nstat, delx, delt, c) -> None:
        ns = np.ascontiguousarray(nstat, dtype=np.int32)
        self._check(self._lib.wm_bcp_init(self._h, ndim, np_, nsp, nxgs, nxge, nygs, nyge, nys,
                                          nye, nup, ndown, mnpi, mnpr, ncomw, nerr,
                                          ns.ctypes.data, ns.size, delx, delt, c))
        self._remember(ndim, np_, nsp, nxgs, nxge, nygs, nyge, nys, nye)

    def mom_calc__init(self, ndim, np_, nsp, nxgs, nxge, nygs, nyge, nys, nye, delx, delt, c,
                       q: Sequence[float], r: Sequence[float]) -> None:
        qa = np.ascontiguousarray(q, dtype=np.float64)
        ra = np.ascontiguousarray(r, dtype=np.float64)
        self._check(self._lib.wm_mom_calc_init(self._h, ndim, np_, nsp, nxgs, nxge, nygs, nyge,
                                               nys, nye, delx, delt, c, qa.ctypes.data,
                                               ra.ctypes.data))
        self._remember(ndim, np_, nsp, nxgs, nxge, nygs, nyge, nys, nye)

    # ------------------------------------------- host-array entry points
    def particle__solv(self, gp, up, uf, cumcnt, nxs, nxe) -> None:
        self._check(self._lib.wm_particle_solv(self._h, self._a(gp, "up", "gp"),
                                               self._a(up, "up"),
                                               self._a(uf, "uf"),
                                               self._a(cumcnt, "cumcnt"), nxs, nxe))

    def field__fdtd_i(self, uf, up, gp, cumcnt, nxs, nxe, bc_family: int = BC_PERIODIC) -> None:
        self._check(self._lib.wm_field_fdtd_i(self._h, self._a(uf, "uf"),
                                              self._a(up, "up"),
                                              self._a(gp, "up", "gp"),
                                              self._a(cumcnt, "cumcnt"), nxs, nxe,
                                              bc_family))

    def boundary_periodic__particle_x(self, up, np2) -> None:
        self._check(self._lib.wm_bcp_particle_x(self._h, self._a(up, "up"),
                                                self._a(np2, "np2")))

    def boundary_periodic__particle_y(self, up, np2) -> None:
        self._check(self._lib.wm_bcp_particle_y(self._h, self._a(up, "up"),
                                                self._a(np2, "np2")))

    def sort__bucket(self, gp, up, cumcnt, np2, nxs, nxe) -> None:
        self._check(self._lib.wm_sort_bucket(self._h, self._a(gp, "up", "gp"),
                                             self._a(up, "up"),
                                             self._a(cumcnt, "cumcnt"),
                                             self._a(np2, "np2"), nxs, nxe))

    def boundary_periodic__dfield(self, df, nxs, nxe, nys, nye, nxgs, nxge) -> None:
        self._check(self._lib.wm_bcp_dfield(self._h, self._a(df, "uf", "df"), nxs, nxe, nys,
                                            nye, nxgs, nxge))

    def boundary_periodic__curre(self, uj, nxs, nxe, nys, nye, nxgs, nxge) -> None:
        self._check(self._lib.wm_bcp_curre(self._h, self._a(uj, "uj"), nxs, nxe, nys,
                                           nye, nxgs, nxge))

    def boundary_periodic__phi(self, phi, nxs, nxe, nys, nye, l) -> None:
        self._check(self._lib.wm_bcp_phi(self._h, self._a(phi, "phi"), nxs, nxe, nys,
                                         nye, l))

    def mom_calc__accl(self, gp, up, uf, cumcnt, nxs, nxe) -> None:
        self._check(self._lib.wm_mom_calc_accl(self._h, self._a(gp, "up", "gp"),
                                               self._a(up, "up"),
                                               self._a(uf, "uf"),
                                               self._a(cumcnt, "cumcnt"), nxs, nxe))

    def mom_calc__nvt(self, mom, up, np2) -> None:
        self._check(self._lib.wm_mom_calc_nvt(self._h, self._a(mom, "mom"),
                                              self._a(up, "up"),
                                              self._a(np2, "np2")))

    def boundary_periodic__mom(self, mom) -> None:
        self._check(self._lib.wm_bcp_mom(self._h, self._a(mom, "mom")))

    # ------------------- the wall families (boundary_shock, boundary_reconnection)
    def _bc_init(self, fam, ndim, np_, nsp, nxgs, nxge, nygs, nyge, nys, nye, nup, ndown, mnpi,
                 mnpr, ncomw, nerr, nstat, delx, delt, c) -> None:
        ns = np.ascontiguousarray(nstat, dtype=np.int32)
        self._check(getattr(self._lib, f"wm_{fam}_init")(
            self._h, ndim, np_, nsp, nxgs, nxge, nygs, nyge, nys, nye, nup, ndown, mnpi, mnpr,
            ncomw, nerr, ns.ctypes.data, ns.size, delx, delt, c))
        self._remember(ndim, np_, nsp, nxgs, nxge, nygs, nyge, nys, nye)

    def boundary_shock__init(self, *args) -> None:
        self._bc_init("bcs", *args)

    def boundary_reconnection__init(self, *args) -> None:
        self._bc_init("bcr", *args)

    def _phi_n(self, nxs, nxe):
        return (nxe - nxs + 3) * (self.grid["ny"] + 2) if self.grid else None

    def _walls(self, fam):
        lib, a = self._lib, self._a

        class W:
            def particle_x(_, up, np2, nxs, nxe):
                self._check(getattr(lib, f"wm_{fam}_particle_x")(self._h, a(up, "up"),
                                                                   a(np2, "np2"), nxs, nxe))

            def particle_y(_, up, np2):
                self._check(getattr(lib, f"wm_{fam}_particle_y")(self._h, a(up, "up"),
                                                                   a(np2, "np2")))

            def dfield(_, df, nxs, nxe, nys, nye, nxgs, nxge):
                self._check(getattr(lib, f"wm_{fam}_dfield")(self._h, a(df, "uf", "df"), nxs,
                                                               nxe, nys, nye, nxgs, nxge))

            def curre(_, uj, nxs, nxe, nys, nye, nxgs, nxge):
                self._check(getattr(lib, f"wm_{fam}_curre")(self._h, a(uj, "uj"), nxs, nxe,
                                                              nys, nye, nxgs, nxge))

            def phi(_, phi, nxs, nxe, nys, nye, l):
                p = _ptr(phi, np.float64, "phi", self._phi_n(nxs, nxe))
                self._check(getattr(lib, f"wm_{fam}_phi")(self._h, p, nxs, nxe, nys, nye, l))

            def mom(_, mom):
                self._check(getattr(lib, f"wm_{fam}_mom")(self._h, a(mom, "mom")))

        return W()

    @property
    def boundary_shock(self):
        """boundary_shock__* (proj/shock/boundary_shock.f90) as methods:
        particle_x, particle_y, dfield, curre, phi, mom (and injection on
        the context)."""
        return self._walls("bcs")

    @property
    def boundary_reconnection(self):
        """boundary_reconnection__* (proj/reconnection/boundary_reconnection.f90)."""
        return self._walls("bcr")

    def boundary_shock__injection(self, up, np2, nxs, nxe, u0: float) -> None:
        self._check(self._lib.wm_bcs_injection(self._h, self._a(up, "up"), self._a(np2, "np2"),
                                               nxs, nxe, float(u0)))

    def set_walls(self, nxs: int, nxe: int, u0: float = 0.0) -> None:
        """Active columns and the shock's u0 for device-resident steps."""
        self._check(self._lib.wm_set_walls(self._h, nxs, nxe, float(u0)))

    # ------------------------------------------------- resident run
    def state_upload(self, up, np2, cumcnt, uf) -> None:
        self._check(self._lib.wm_state_upload(self._h, self._a(up, "up"),
                                              self._a(np2, "np2"),
                                              self._a(cumcnt, "cumcnt"),
                                              self._a(uf, "uf")))

    def state_download(self, up=None, np2=None, cumcnt=None, uf=None) -> None:
        self._check(self._lib.wm_state_download(self._h, self._a(up, "up"),
                                                self._a(np2, "np2"),
                                                self._a(cumcnt, "cumcnt"),
                                                self._a(uf, "uf")))

    def init_weibel(self, seed: int, ppc: int, vth: float, udrift: float) -> None:
        self._check(self._lib.wm_state_init_weibel(self._h, int(seed), int(ppc), float(vth),
                                                   float(udrift)))

    def step(self, nsteps: int = 1) -> None:
        self._check(self._lib.wm_step(self._h, int(nsteps)))

    def energy(self) -> np.ndarray:
        nsp = self.grid["nsp"]
        out = np.zeros(nsp + 2)
        self._check(self._lib.wm_energy(self._h, out.ctypes.data))
        return out

    def moments(self) -> np.ndarray:
        """mom (7, nx+2, ny+2, nsp) of the resident state: mom_calc__accl,
        mom_calc__nvt and boundary_periodic__mom on the device."""
        g = self.grid
        mom = np.zeros((7, g["nx"] + 2, g["ny"] + 2, g["nsp"]), order="F")
        self._check(self._lib.wm_moments(self._h, mom.ctypes.data))
        return mom

    # ------------------------------------------ restart files (paraio)
    def paraio__init(self, ndim, np_, nsp, nxgs, nxge, nygs, nyge, nys, nye, nproc, nrank, delx,
                     delt, c, q, r, dir_: str) -> None:
        q = np.ascontiguousarray(q, dtype=np.float64)
        r = np.ascontiguousarray(r, dtype=np.float64)
        self._check(self._lib.wm_paraio_init(self._h, ndim, np_, nsp, nxgs, nxge, nygs, nyge, nys,
                                             nye, nproc, nrank, delx, delt, c, q.ctypes.data,
                                             r.ctypes.data, dir_.encode()))

    def paraio__output(self, up, uf, np2, nxs, nxe, it, filename: str) -> None:
        self._check(self._lib.wm_paraio_output(self._h, self._a(up, "up"), self._a(uf, "uf"),
                                               self._a(np2, "np2"), nxs, nxe, it,
                                               filename.encode()))

    def paraio__input(self, up, uf, np2, filename: str):
        """Reads into up, uf, np2; returns (nxs, nxe, it)."""
        a, b, c = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        self._check(self._lib.wm_paraio_input(self._h, self._a(up, "up"), self._a(uf, "uf"),
                                              self._a(np2, "np2"), ctypes.byref(a),
                                              ctypes.byref(b), ctypes.byref(c),
                                              filename.encode()))
        return a.value, b.value, c.value

    def paraio__param(self, n0, wpe, wpi, wge, wgi, vti, vte, filename: str) -> None:
        self._check(self._lib.wm_paraio_param(self._h, n0, wpe, wpi, wge, wgi, vti, vte,
                                              filename.encode()))

    def paraio__mom(self, mom, uf, it) -> None:
        self._check(self._lib.wm_paraio_mom(self._h, self._a(mom, "mom"), self._a(uf, "uf"), it))

    def paraio__ptcl(self, up, uf, np2, it) -> None:
        self._check(self._lib.wm_paraio_ptcl(self._h, self._a(up, "up"), self._a(uf, "uf"),
                                             self._a(np2, "np2"), it))

    def paraio__orb(self, up, uf, np2, it) -> None:
        self._check(self._lib.wm_paraio_orb(self._h, self._a(up, "up"), self._a(uf, "uf"),
                                            self._a(np2, "np2"), it))

    def save_restart(self, it: int, filename: str) -> None:
        """paraio__output of the device-resident state, row by row."""
        self._check(self._lib.wm_paraio_output_state(self._h, int(it), filename.encode()))

    def load_restart(self, filename: str) -> int:
        """paraio__input into the device-resident state (rows sorted by cell, CG
        warm start cleared); returns it."""
        it = ctypes.c_int()
        self._check(self._lib.wm_paraio_input_state(self._h, ctypes.byref(it), filename.encode()))
        return it.value

    def cg_iters(self) -> np.ndarray:
        out = np.zeros(3, dtype=np.int32)
        self._check(self._lib.wm_cg_iters(self._h, out.ctypes.data))
        return out

    def phase_times(self) -> np.ndarray:
        out = np.zeros(8)
        n = self._lib.wm_phase_times(self._h, out.ctypes.data, 8)
        return out[:max(n, 0)]

    def particle_count(self) -> int:
        c = ctypes.c_longlong()
        self._check(self._lib.wm_particle_count(self._h, ctypes.byref(c)))
        return c.value

    def id_checksum(self) -> tuple:
        """(sum, xor) of the active particles' 64-bit IDs, on the device."""
        out = np.zeros(2, dtype=np.uint64)
        self._check(self._lib.wm_id_checksum(self._h, out.ctypes.data))
        return int(out[0]), int(out[1])

    def capacity_stats(self) -> dict:
        """FP32 form: the particle store's capacities and their use since the
        state was made (maxima over the pushes)."""
        out = np.zeros(12, dtype=np.int64)
        n = self._lib.wm_capacity_stats(self._h, out.ctypes.data, 12)
        if n != 12:                     # an error status (12 entries on success)
            self._check(n)
        keys = ("max_bin_fill", "max_bin_movers", "max_workgroup_movers", "max_row_records",
                "max_row_spill", "relayouts", "slots_per_row", "records_per_row",
                "workgroup_lds_list", "spill_per_row", "full_bins", "segment_relayouts")
        return {k: int(v) for k, v in zip(keys, out[:n])}

    # ------------------------------------------ learned preconditioner
    def set_solver(self, solver: int) -> None:
        """SOLVER_CG (the reference's plain CG) or SOLVER_FCG (flexible CG
        with the learned preconditioner; plain CG judges and falls back)."""
        self._check(self._lib.wm_set_solver(self._h, int(solver)))

    def pc_collect(self, nslot: int, per_solve: int) -> None:
        """Record (residual, correction) pairs of the next plain solves."""
        self._check(self._lib.wm_pc_collect(self._h, int(nslot), int(per_solve)))

    def pc_pair_count(self) -> int:
        return int(self._lib.wm_pc_pair_count(self._h))

    def pc_init(self, seed: int = 0) -> None:
        self._check(self._lib.wm_pc_init(self._h, int(seed)))

    def pc_train(self, steps: int, lr: float = 3e-3, seed: int = 0) -> np.ndarray:
        """Adam on the recorded pairs; returns [first loss, mean of last ten]."""
        out = np.zeros(2)
        self._check(self._lib.wm_pc_train(self._h, int(steps), float(lr), int(seed),
                                          out.ctypes.data))
        return out

    def pc_params(self) -> np.ndarray:
        n = self._lib.wm_pc_num_params()
        th = np.zeros(n, dtype=np.float32)
        self._check(self._lib.wm_pc_params(self._h, th.ctypes.data, n, 0))
        return th

    def pc_set_params(self, theta: np.ndarray) -> None:
        th = np.ascontiguousarray(theta, dtype=np.float32)
        self._check(self._lib.wm_pc_params(self._h, th.ctypes.data, th.size, 1))

    def pc_apply(self, r: np.ndarray) -> np.ndarray:
        """z = M(r) for three images r[3, ny, nx]."""
        r = np.ascontiguousarray(r, dtype=np.float64)
        z = np.zeros_like(r)
        self._check(self._lib.wm_pc_apply(self._h, r.ctypes.data, z.ctypes.data))
        return z

    def pc_apply_noalloc(self, r_dev: int, z_dev: int, nx: int, ny: int) -> None:
        """z = M(r) on device images (addresses r_dev, z_dev of [3, ny, nx]
        doubles), allocating nothing: the activations live in the library's
        static scratch array (nx * ny <= 262144)."""
        self._check(self._lib.wm_pc_apply_noalloc(self._h, ctypes.c_void_p(r_dev),
                                                  ctypes.c_void_p(z_dev), int(nx), int(ny)))

    def pc_grad(self, r: np.ndarray, e: np.ndarray):
        """(loss, dL/dtheta) for images r, e [3, ny, nx]."""
        r = np.ascontiguousarray(r, dtype=np.float64)
        e = np.ascontiguousarray(e, dtype=np.float64)
        g = np.zeros(self._lib.wm_pc_num_params(), dtype=np.float32)
        loss = ctypes.c_double()
        self._check(self._lib.wm_pc_grad(self._h, r.ctypes.data, e.ctypes.data, g.ctypes.data,
                                         ctypes.byref(loss)))
        return loss.value, g

    def solve_stats(self):
        """(flexible-CG iterations, plain-CG iterations) of the last solve."""
        a = np.zeros(3, dtype=np.int32)
        b = np.zeros(3, dtype=np.int32)
        self._check(self._lib.wm_solve_stats(self._h, a.ctypes.data, b.ctypes.data))
        return a, b
