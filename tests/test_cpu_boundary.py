"""CPU checks of the drop-in boundary and the host layer (no GPU compute)."""
import ctypes
import os
import re
import subprocess
import sys

import numpy as np
import pytest

import wumingpic2d_amd as wm
from wumingpic2d_amd import weibel

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "wmpic.h")


def _declared():
    txt = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char \*)\s*(wm_\w+)\s*\(", txt, re.M)))


def test_library_exports_every_declared_symbol():
    names = _declared()
    assert len(names) >= 25
    lib = ctypes.CDLL(wm.LIB_PATH)
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(wm.EXPORTS)


def test_fortran_dropin_modules_export_the_reference_procedures():
    lib = os.path.join(ROOT, "wumingpic2d_amd", "libwmpic_f.so")
    out = subprocess.run(["nm", "-D", lib], capture_output=True, text=True).stdout
    for mod, procs in {
        "particle": ["particle__init", "particle__solv"],
        "field": ["field__init", "field__fdtd_i"],
        "sort": ["sort__init", "sort__bucket"],
        "boundary_periodic": ["boundary_periodic__init", "boundary_periodic__particle_x",
                              "boundary_periodic__particle_y", "boundary_periodic__dfield",
                              "boundary_periodic__curre", "boundary_periodic__phi"],
        "mpi_set": ["mpi_set__init"],
    }.items():
        for p in procs:
            assert f"_QM{mod}P{p}" in out, (mod, p)


def test_mpi_set_single_rank_decomposition():
    d = wm.mpi_set__init(2, 129, 1)
    assert d == dict(nys=2, nye=129, nup=0, ndown=0, nrank=0, nsize=1)
    with pytest.raises(wm.WmError):
        wm.mpi_set__init(2, 129, 4)


def test_no_device_fails_loudly():
    """Without a visible GPU the product raises; there is no CPU fallback."""
    try:
        import torch  # noqa: F401
        has_gpu = torch.cuda.is_available()
    except Exception:
        has_gpu = False
    if has_gpu:
        pytest.skip("a GPU is visible")
    with pytest.raises(wm.WmError) as e:
        wm.Context(0)
    assert e.value.code == wm.WM_ERR_NODEV


def test_array_layout_is_checked():
    p = weibel.grid_params(8, 4, 2)
    up, np2, cc, uf = weibel.init_state(p, 1)
    with pytest.raises(TypeError):
        wm._ptr(np.ascontiguousarray(up), np.float64, "up")
    with pytest.raises(TypeError):
        wm._ptr(cc.astype(np.int64, order="F"), np.int32, "cumcnt")


@pytest.mark.parametrize("nx,ny,ppc,seed", [(16, 8, 4, 1), (33, 5, 7, 2)])
def test_initial_state_invariants(nx, ny, ppc, seed):
    p = weibel.grid_params(nx, ny, ppc)
    up, np2, cc, uf = weibel.init_state(p, seed, vth=0.05, udrift=0.5)
    n = ppc * nx
    assert np.all(np2 == n)
    ids = up[5, :n].view(np.int64).ravel()
    assert len(np.unique(ids)) == ids.size and ids.max() < 0
    for s in range(2):
        for jl in range(ny):
            x, y = up[0, :n, jl, s], up[1, :n, jl, s]
            key = x.astype(int) - p["nxgs"]
            assert np.all(np.diff(key) >= 0)                  # sorted by cell
            assert np.array_equal(cc[:, jl, s], np.concatenate([[0], np.cumsum(np.bincount(key, minlength=nx))]))
            assert np.all((y >= jl + p["nys"]) & (y < jl + p["nys"] + 1))
            assert np.all((x >= p["nxgs"]) & (x < p["nxge"] + 1))
    # species share positions (zero net charge) and counter-stream in z
    a = np.sort(up[0, :n, :, 0].ravel())
    b = np.sort(up[0, :n, :, 1].ravel())
    assert np.array_equal(a, b)
    uz = up[4, :n, :, 0]
    assert abs(np.mean(np.abs(uz)) - 0.5) < 0.05
    q, r = weibel.pair_charges(ppc, 0.125)
    assert q[0] == -q[1] and r[0] == r[1] == 1.0
    # plasma frequency: 4 pi n q^2 / m = wpe^2 with n = ppc
    assert abs(4 * np.pi * ppc * q[0] ** 2 / r[0] - 0.125 ** 2) < 1e-15


def test_bench_refuses_several_gpus():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=60)
    assert r.returncode == 2
    assert "one GPU" in r.stderr


def test_undersized_arrays_are_refused_before_any_device_call():
    """The C side reads and writes the reference's full shapes: an array
    smaller than its grid shape is a ValueError on the host."""
    from wumingpic2d_amd import _ptr
    a = np.zeros((6, 10, 4, 2), order="F")
    assert _ptr(a, np.float64, "up", a.size) == a.ctypes.data
    with pytest.raises(ValueError):
        _ptr(a, np.float64, "up", a.size + 1)
    with pytest.raises(TypeError):
        _ptr(np.zeros((6, 10), order="C"), np.float64, "up", 60)
