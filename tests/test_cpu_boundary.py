"""CPU checks of the drop-in boundary and the host layer (no GPU compute)."""
import ctypes
import json
import os
import re
import subprocess
import sys

import numpy as np
import pytest

import wumingpic2d_amd as wm
from wumingpic2d_amd import weibel

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADERS = [os.path.join(ROOT, "include", h) for h in sorted(os.listdir(os.path.join(ROOT, "include")))
           if h.endswith(".h")]


def _declared():
    txt = "".join(open(h).read() for h in HEADERS)
    return sorted(set(re.findall(r"^\s*(?:int|void|void \*|const char \*)\s*\*?\s*(wm_\w+)\s*\(",
                                 txt, re.M)))


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


def test_json_c_abi_parses_queries_updates_and_prints(tmp_path):
    """The JSON model behind json_module (include/wmpic_json.h), on text
    formatted as json-fortran prints it (reals like 0.1E+1, arrays on lines)."""
    lib = ctypes.CDLL(wm.LIB_PATH)
    vp, cp = ctypes.c_void_p, ctypes.c_char_p
    lib.wm_json_parse.restype = vp
    lib.wm_json_parse.argtypes = [cp, ctypes.c_int]
    lib.wm_json_child.restype = vp
    lib.wm_json_child.argtypes = [vp, cp]
    for n in ("wm_json_get_int", "wm_json_get_real", "wm_json_get_str", "wm_json_update_str",
              "wm_json_print", "wm_json_serialize", "wm_json_array_size"):
        getattr(lib, n).restype = ctypes.c_int
    lib.wm_json_get_int.argtypes = [vp, cp, ctypes.POINTER(ctypes.c_longlong)]
    lib.wm_json_get_real.argtypes = [vp, cp, ctypes.POINTER(ctypes.c_double)]
    lib.wm_json_get_str.argtypes = [vp, cp, ctypes.c_char_p, ctypes.c_int]
    lib.wm_json_update_str.argtypes = [vp, cp, cp]
    lib.wm_json_print.argtypes = [vp, cp]
    lib.wm_json_serialize.argtypes = [vp, ctypes.c_char_p, ctypes.c_int]
    lib.wm_json_array_size.argtypes = [vp, cp]
    lib.wm_json_free.argtypes = [vp]
    text = b"""{
  "attribute": {
    "delt": { "datatype": "f8", "offset": 60, "size": 8, "ndim": 1,
              "shape": [
                1
              ], "description": "", "data": 0.1E+1 },
    "name": "a \\"quoted\\" string"
  },
  "flag": true
}"""
    root = lib.wm_json_parse(text, len(text))
    assert root
    i, r = ctypes.c_longlong(), ctypes.c_double()
    assert lib.wm_json_get_int(root, b"attribute.delt.offset", ctypes.byref(i)) == 1 and i.value == 60
    assert lib.wm_json_get_real(root, b"attribute.delt.data", ctypes.byref(r)) == 1 and r.value == 1.0
    assert lib.wm_json_get_int(root, b"attribute.delt.shape( 1)", ctypes.byref(i)) == 1 and i.value == 1
    assert lib.wm_json_array_size(root, b"attribute.delt.shape") == 1
    assert lib.wm_json_get_int(root, b"attribute.nothing", ctypes.byref(i)) == 0
    buf = ctypes.create_string_buffer(64)
    n = lib.wm_json_get_str(root, b"attribute.name", buf, 64)
    assert buf.raw[:n] == b'a "quoted" string'
    assert lib.wm_json_update_str(root, b"attribute.name", b"new") == 1
    assert lib.wm_json_update_str(root, b"attribute.added", b"x") == 0
    n = lib.wm_json_serialize(root, None, 0)
    out = ctypes.create_string_buffer(n)
    lib.wm_json_serialize(root, out, n)
    back = json.loads(out.raw[:n])
    assert back["attribute"]["name"] == "new" and back["attribute"]["added"] == "x"
    assert back["attribute"]["delt"]["data"] == 1.0 and back["flag"] is True
    path = str(tmp_path / "p.json")
    assert lib.wm_json_print(root, path.encode()) == 1
    assert json.load(open(path)) == back
    lib.wm_json_free(root)
    assert not lib.wm_json_parse(b"{ bad", 5)


def test_fortran_json_module_mpi_and_mpiio_as_the_app_uses_them(tmp_path):
    exe = os.path.join(ROOT, "tests", "fortran", "json_check.x")
    cfg = tmp_path / "config.json"
    cfg.write_text(json.dumps({"config": {"verbose": 1, "datadir": "./out", "max_it": 200,
                                          "restart_file": "0000004_restart"},
                               "parameter": {"num_process": 1, "mass_ratio": 0.25E+2}}))
    r = subprocess.run([exe, str(cfg)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    got = dict(ln.split(None, 1) for ln in r.stdout.splitlines())
    assert got["max_it"] == "200" and got["datadir"] == "./out"
    assert got["restart_found"] == "T" and got["num_process"] == "1"
    assert float(got["mass_ratio"]) == 25.0
    assert got["failed_after_missing"] == "T" and got["updated_existing"] == "T"
    assert json.load(open(str(cfg) + ".out"))["config"]["restart_file"] == "0000011_restart"
    assert got["reduced"].split() == ["1.50", "2.50", "3.50"]
    assert got["append_disp"] == "12" and got["read_back"].split() == ["7", "0.25"]
    assert got["endian"] == "1" and got["uniform_in_range"] == "T"
