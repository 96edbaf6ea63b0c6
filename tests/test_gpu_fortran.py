"""The reference's time loop, written in Fortran against the drop-in modules
(particle, field, sort, boundary_periodic, mom_calc, mpi_set with the
reference's module and procedure names and argument lists), run on the device
and compared with the oracle over the same steps and the moments of an output
step."""
import os
import subprocess

import numpy as np
import pytest

from oracle import Oracle
from wmhelpers import active, bits, state

pytestmark = pytest.mark.gpu

DRIVER = os.path.join(os.path.dirname(__file__), "fortran", "weibel_loop.x")


def _write(d, p, up, np2, cc, uf, nsteps):
    with open(os.path.join(d, "params.txt"), "w") as f:
        f.write(f"{p['ndim']} {p['np']} {p['nsp']} {p['nxgs']} {p['nxge']} {p['nygs']} {p['nyge']}\n")
        f.write(f"{p['delx']!r} {p['delt']!r} {p['c']!r} {p['gfac']!r}\n")
        f.write(" ".join(repr(float(v)) for v in p["q"]) + "\n")
        f.write(" ".join(repr(float(v)) for v in p["r"]) + "\n")
        f.write(f"{nsteps}\n")
    with open(os.path.join(d, "state.bin"), "wb") as f:
        for a in (up, np2, cc, uf):
            f.write(np.asfortranarray(a).tobytes(order="F"))


def _read(d, up, np2, cc, uf, mom):
    raw = open(os.path.join(d, "out.bin"), "rb").read()
    out, off = [], 0
    for a in (up, np2, cc, uf, mom):
        n = a.nbytes
        out.append(np.frombuffer(raw[off:off + n], dtype=a.dtype).reshape(a.shape, order="F"))
        off += n
    return out


def test_fortran_weibel_loop_matches_oracle(tmp_path):
    assert os.path.exists(DRIVER), "tests/fortran/weibel_loop.x is not built"
    p, up, np2, cc, uf = state(32, 16, 8, 31, vth=0.1)
    nsteps = 5
    _write(str(tmp_path), p, up, np2, cc, uf, nsteps)
    r = subprocess.run([DRIVER, str(tmp_path)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    mom = np.zeros((7, p["nx"] + 2, p["ny"] + 2, p["nsp"]), order="F")
    fup, fnp2, fcc, fuf, fmom = _read(str(tmp_path), up, np2, cc, uf, mom)
    orc = Oracle(p)
    gp = up.copy(order="F")
    for _ in range(nsteps):
        orc.step(up, gp, uf, cc, np2)
    orc.mom_calc_accl(gp, up, uf, cc)            # proj/weibel/app.f90:119-121
    orc.mom_calc_nvt(mom, gp, np2)
    orc.bc_mom(mom)
    for s in range(p["nsp"]):
        for k in range(7):
            scale = np.abs(mom[k, ..., s]).max()
            assert np.abs(fmom[k, ..., s] - mom[k, ..., s]).max() <= 1e-11 * scale
    assert np.array_equal(fnp2, np2)
    assert np.array_equal(fcc, cc)
    for s in range(p["nsp"]):
        for jl in range(p["ny"]):
            n = np2[jl, s]
            assert np.array_equal(bits(fup[5, :n, jl, s]), bits(up[5, :n, jl, s]))
            assert np.allclose(active(fup, np2, jl, s), active(up, np2, jl, s), rtol=1e-12,
                               atol=1e-13)
    assert np.abs(fuf - uf).max() <= 1e-11 * np.abs(uf).max()


def test_fortran_fatal_message_is_the_references(tmp_path):
    """A row overflowing np stops the program with the reference's message
    (boundary_periodic.f90:231-234)."""
    assert os.path.exists(DRIVER)
    p, up, np2, cc, uf = state(16, 8, 4, 32, vth=0.3, np_cap=64)
    _write(str(tmp_path), p, up, np2, cc, uf, 3)
    r = subprocess.run([DRIVER, str(tmp_path)], capture_output=True, text=True, timeout=120)
    assert "memory over (np2 > np)" in (r.stdout + r.stderr)
