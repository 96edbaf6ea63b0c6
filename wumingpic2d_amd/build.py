"""Build recipe of libwmpic.so (HIP, gfx950) -- run by __graft_entry__.build().

Each ``csrc/*.hip`` file (and host-only ``csrc/*.cpp``) is compiled to an
object in parallel, then linked
into ``wumingpic2d_amd/libwmpic.so`` in-tree.  Packed FP32 instructions are
disabled for every device compile.
"""
from __future__ import annotations

import concurrent.futures as cf
import glob
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libwmpic.so")

ARCH = os.environ.get("WM_OFFLOAD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", "hipcc")
FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17", "-Wall",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    f"-I{os.path.join(ROOT, 'include')}",
]


def _run(cmd: list[str]) -> None:
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"command failed: {' '.join(cmd)}")


def build_hip(verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    srcs = sorted(glob.glob(os.path.join(CSRC, "*.hip")) + glob.glob(os.path.join(CSRC, "*.cpp")))
    objs = [os.path.join(BUILD, os.path.splitext(os.path.basename(s))[0] + ".o") for s in srcs]
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        futs = [ex.submit(_run, [HIPCC, *FLAGS, "-c", s, "-o", o]) for s, o in zip(srcs, objs)]
        for f in futs:
            f.result()
    tmp = LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs])
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


FORTRAN = os.path.join(HERE, "fortran")
FMOD = os.path.join(FORTRAN, "mod")
FLIB = os.path.join(HERE, "libwmpic_f.so")
FSRC = ["apps/mpi.f90", "wmpic_capi.f90", "wm_mpi_set.f90", "wm_particle.f90", "wm_field.f90",
        "wm_sort.f90", "wm_boundary_periodic.f90", "wm_boundary_shock.f90",
        "wm_boundary_reconnection.f90", "wm_mom_calc.f90",
        # what the reference's apps use besides the hot path (single process)
        "apps/json_module.f90", "apps/mpiio.f90", "apps/jsonio.f90", "apps/wm_paraio.f90",
        "apps/wuming_utils.f90", "apps/wuming2d.f90"]
FCSRC = ["apps/wm_mpi_stub.c"]
DRIVER_SRC = os.path.join(ROOT, "tests", "fortran", "weibel_loop.f90")
DRIVER = os.path.join(ROOT, "tests", "fortran", "weibel_loop.x")
JSON_SRC = os.path.join(ROOT, "tests", "fortran", "json_check.f90")
JSON_X = os.path.join(ROOT, "tests", "fortran", "json_check.x")


def build_fortran(verbose: bool = False) -> str:
    """Drop-in Fortran modules (particle, field, sort, boundary_periodic,
    mpi_set) over the C ABI -> libwmpic_f.so + .mod files in fortran/mod,
    and the Fortran test driver of the Weibel time loop."""
    fc = os.environ.get("FC", "amdflang")
    os.makedirs(FMOD, exist_ok=True)
    os.makedirs(BUILD, exist_ok=True)
    srcs = [os.path.join(FORTRAN, f) for f in FSRC]
    cobjs = []
    for c in FCSRC:
        o = os.path.join(BUILD, os.path.basename(c)[:-2] + ".o")
        _run([os.environ.get("CC", "gcc"), "-O2", "-fPIC", "-c", os.path.join(FORTRAN, c), "-o", o])
        cobjs.append(o)
    tmp = FLIB + ".tmp"
    _run([fc, "-O2", "-fPIC", "-shared", "-module-dir", FMOD, *srcs, *cobjs, "-o", tmp,
          f"-L{HERE}", "-lwmpic", "-Wl,-rpath,$ORIGIN"])
    os.replace(tmp, FLIB)
    _run([fc, "-O2", "-I", FMOD, DRIVER_SRC, "-o", DRIVER + ".tmp", f"-L{HERE}", "-lwmpic_f",
          "-lwmpic", "-Wl,-rpath,$ORIGIN/../../wumingpic2d_amd"])
    os.replace(DRIVER + ".tmp", DRIVER)
    _run([fc, "-O2", "-I", FMOD, JSON_SRC, "-o", JSON_X + ".tmp", f"-L{HERE}", "-lwmpic_f",
          "-lwmpic", "-Wl,-rpath,$ORIGIN/../../wumingpic2d_amd"])
    os.replace(JSON_X + ".tmp", JSON_X)
    if verbose:
        print(f"built {FLIB} and {DRIVER}")
    return FLIB


if __name__ == "__main__":
    build_hip(verbose=True)
    build_fortran(verbose=True)
