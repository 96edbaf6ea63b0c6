"""The reference's Weibel app, unchanged (proj/weibel/app.f90 and main.f90,
compiled by tests/fortran/build_app.sh against this package's drop-in
modules: particle, field, sort, boundary_periodic, mpi_set, mom_calc, the
single-rank mpi, json_module, jsonio, mpiio, paraio, wuming_utils and the
wuming2d umbrella), restarted from a file the library wrote, runs its own
loop of host-array calls on the device (app.f90:100-146), writes its energy
history, moments and final restart file, and rewrites its configuration.
Its final state is checked against the oracle resumed from the same state
with the app's own physical constants (read back from the app's restart
file) and the fresh solver state of a restarted process."""
import json
import os
import subprocess

import numpy as np
import pytest

from oracle import Oracle
from wmhelpers import grid_args, make_ctx, state

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
APP = os.path.join(ROOT, "tests", "fortran", "_app", "weibel_app.x")


def _attr(d, name, key):
    meta = json.load(open(os.path.join(d, name + ".json")))
    m = meta["attribute"][key]
    dt = {"i4": np.int32, "f8": np.float64}[m["datatype"]]
    raw = open(os.path.join(d, name + ".raw"), "rb").read()
    return np.frombuffer(raw, dtype=dt, count=m["size"] // np.dtype(dt).itemsize, offset=m["offset"])


def _dataset(d, name, key):
    meta = json.load(open(os.path.join(d, name + ".json")))
    m = meta["dataset"][key]
    dt = {"i4": np.int32, "f8": np.float64}[m["datatype"]]
    raw = open(os.path.join(d, name + ".raw"), "rb").read()
    a = np.frombuffer(raw, dtype=dt, count=m["size"] // np.dtype(dt).itemsize, offset=m["offset"])
    return a.reshape(m["shape"], order="F")


@pytest.mark.skipif(not os.path.exists(APP), reason="the reference app was not built here "
                    "(tests/fortran/build_app.sh needs the reference tree)")
def test_unchanged_weibel_app_resumes_on_the_device_and_matches_the_oracle(tmp_path):
    d = str(tmp_path) + "/"
    nx, ny, ppc, it0, it1 = 32, 16, 8, 4, 10
    p, up, np2, cc, uf = state(nx, ny, ppc, 71, vth=0.1)     # np = 5 ppc nx, as app.f90:248
    w = make_ctx(p)
    w.paraio__init(*grid_args(p), 1, 0, p["delx"], p["delt"], p["c"], p["q"], p["r"], d)
    w.paraio__output(up, uf, np2, p["nxgs"], p["nxge"], it0, "%07d_restart" % it0)
    w.close()
    open(d + "energy.dat", "w").close()          # energy_history appends after a restart
    cfg = {"config": {"verbose": 1, "datadir": d, "max_elapsed": 1.0e6, "max_it": it1,
                      "intvl_ptcl": 100000, "intvl_mom": it1, "intvl_orb": 100000,
                      "restart_file": "%07d_restart" % it0},
           "parameter": {"num_process": 1, "n_ppc": ppc, "n_x": nx, "n_y": ny,
                         "mass_ratio": 1.0, "sigma_e": 0.0, "omega_pe": 0.125, "v_the": 0.1,
                         "v_thi": 0.1, "t_ani": 1.0}}
    json.dump(cfg, open(d + "config.json", "w"), indent=2)
    # no argument: the app reads config.json in its directory (app.f90:22,
    # 184-185); its argument form passes an unallocated deferred-length
    # string to get_command_argument (app.f90:24, 188), which reads as blank
    r = subprocess.run([APP], cwd=d, capture_output=True, text=True, timeout=300)
    print(r.stdout[-1500:], r.stderr[-1500:])
    assert r.returncode == 0
    assert r.stdout.count("Time step:") == it1 - it0
    final = "%07d_restart" % (it1 + 1)
    # the app's constants (app.f90:290-301: delt = cfl delx / c, q and r from
    # the mass ratio and omega_pe) as its restart file records them
    pa = dict(p, delt=float(_attr(d, final, "delt")[0]), q=_attr(d, final, "q").copy(),
              r=_attr(d, final, "r").copy())
    assert pa["delt"] == 1.0 and np.array_equal(pa["r"], [1.0, 1.0])
    orc = Oracle(pa)                              # a restarted process: CG warm start at zero
    gp = up.copy(order="F")
    for _ in range(it1 - it0):
        orc.step(up, gp, uf, cc, np2)
    assert int(_attr(d, final, "it")[0]) == it1 + 1
    gup = _dataset(d, final, "up")[..., 0]
    gnp2 = _dataset(d, final, "np2")[..., 0]
    guf = _dataset(d, final, "uf")[..., 0]
    assert np.array_equal(gnp2, np2)
    for s in range(2):
        for jl in range(ny):
            n = np2[jl, s]
            assert np.array_equal(gup[5, :n, jl, s].view(np.int64), up[5, :n, jl, s].view(np.int64))
            assert np.allclose(gup[:5, :n, jl, s], up[:5, :n, jl, s], rtol=1e-12, atol=1e-12)
    assert np.abs(guf - uf).max() <= 1e-11 * np.abs(uf).max()
    # energy.dat: one line at it1 in (f8.2, 4(1x, e12.5)) (app.f90:537-542)
    lines = [ln for ln in open(d + "energy.dat").read().splitlines() if ln.strip()]
    assert len(lines) == 1
    t, kin, ee, be, tot = (float(v) for v in lines[0].split())
    e = orc.energy(up, uf, np2)
    assert t == it1 * pa["delt"]
    for got, want in ((kin, e[0] + e[1]), (ee, e[2]), (be, e[3]), (tot, e.sum())):
        assert abs(got - want) <= 5e-5 * abs(want) + 1e-300
    # the moments of the output step and the rewritten configuration
    mom = json.load(open(d + "%07d_mom.json" % it1))
    assert list(mom["dataset"]) == ["den", "vel", "temp", "uf"]
    assert json.load(open(d + "config.json"))["config"]["restart_file"] == final
