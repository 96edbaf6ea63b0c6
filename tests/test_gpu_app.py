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


RECON_APP = os.path.join(ROOT, "tests", "fortran", "_app", "reconnection_app.x")
SHOCK_APP = os.path.join(ROOT, "tests", "fortran", "_app", "shock_app.x")


def _resume_files(p, up, np2, uf, d, it0):
    """The state as a restart file (the app resumes from it, app.f90's
    io__input), written through the library's paraio__output."""
    w = make_ctx(p)
    w.paraio__init(*grid_args(p), 1, 0, p["delx"], p["delt"], p["c"], p["q"], p["r"], d)
    w.paraio__output(up, uf, np2, p["nxs"], p["nxe"], it0, "%07d_restart" % it0)
    w.close()


def _app_constants(p, d, name):
    """The app's own delt, q, r as its restart file records them."""
    return dict(p, delt=float(_attr(d, name, "delt")[0]), q=_attr(d, name, "q").copy(),
                r=_attr(d, name, "r").copy())


@pytest.mark.skipif(not os.path.exists(RECON_APP), reason="the reference app was not built here "
                    "(tests/fortran/build_app.sh needs the reference tree)")
def test_unchanged_reconnection_app_resumes_on_the_device_and_matches_the_oracle(tmp_path):
    """proj/reconnection/app.f90 and main.f90, unchanged, with this package's
    boundary_reconnection: conducting walls at nxs and nxe, its loop
    (app.f90:101-106) of host-array calls on the device, resumed from a
    restart file and checked against the oracle resumed the same way with
    the app's constants (np = (nbg + ncs) nx, app.f90:248-249)."""
    from oracle import BC_RECONNECTION
    from wmhelpers import wall_state
    d = str(tmp_path) + "/"
    nx, ny, ppc, it0, it1 = 32, 16, 8, 3, 9
    nbg, ncs = 2, 8
    p, up, np2, cc, uf = wall_state(BC_RECONNECTION, nx, ny, ppc, 72, vth=0.1,
                                    np_cap=(nbg + ncs) * nx)
    rng = np.random.default_rng(72)
    uf[2] = rng.normal(0, 0.05, uf.shape[1:])
    _resume_files(p, up, np2, uf, d, it0)
    # the app saves no final restart (app.f90:150-154): its particle output
    # at it1 (io__ptcl, every active particle in row order, and the fields)
    # carries the final state and, as attributes, its constants
    cfg = {"config": {"verbose": 1, "datadir": d, "max_elapsed": 1.0e6, "max_it": it1,
                      "intvl_ptcl": it1, "intvl_mom": 100000, "intvl_orb": 100000,
                      "restart_file": "%07d_restart" % it0},
           "parameter": {"num_process": 1, "n_x": nx, "n_y": ny, "mass_ratio": 4.0, "alpha": 2.0,
                         "rtemp": 1.0, "lcs": 2.0, "nbg": nbg, "ncs": ncs}}
    json.dump(cfg, open(d + "config.json", "w"), indent=2)
    r = subprocess.run([RECON_APP], cwd=d, capture_output=True, text=True, timeout=300)
    print(r.stdout[-1500:], r.stderr[-1500:])
    assert r.returncode == 0
    assert r.stdout.count("Time step:") == it1 - it0
    out = "%07d_ptcl" % it1
    pa = _app_constants(p, d, out)
    assert pa["r"][0] == 4.0 and pa["delt"] == 0.5
    orc = Oracle(pa)
    gp = up.copy(order="F")
    for _ in range(it1 - it0):
        orc.step(up, gp, uf, cc, np2)
    meta = json.load(open(d + out + ".json"))
    raw = open(d + out + ".raw", "rb").read()
    get = lambda m: np.frombuffer(raw, dtype=np.float64, count=m["size"] // 8,
                                  offset=m["offset"]).reshape(m["shape"], order="F")
    for s in range(2):
        want = np.concatenate([up[:, :np2[jl, s], jl, s] for jl in range(ny)], axis=1)
        got = get(meta["dataset"]["up%02d" % (s + 1)])
        assert got.shape == want.shape
        assert np.array_equal(got[5].view(np.int64), want[5].view(np.int64))
        assert np.allclose(got[:5], want[:5], rtol=1e-12, atol=1e-12)
    guf = get(meta["dataset"]["uf"])
    assert np.abs(guf - uf[:, 2:-2, 2:-2]).max() <= 1e-11 * np.abs(uf).max()


@pytest.mark.skipif(not os.path.exists(SHOCK_APP), reason="the reference app was not built here "
                    "(tests/fortran/build_app.sh needs the reference tree)")
def test_unchanged_shock_app_resumes_on_the_device(tmp_path):
    """proj/shock/app.f90 and main.f90, unchanged, with this package's
    boundary_shock: the wall at nxs, the injection edge moving with u0, an
    active range nxs..nxe short of the grid.  One step from a restart file
    against the oracle's (the app's inject() then appends its random
    particles to each row and resets Bx..Ez of columns nxe-1, nxe,
    app.f90:695-850, which the comparison leaves out); then a longer run in
    which relocate() grows the box every 4 steps (app.f90:613-690) ends
    with the box grown, every particle inside it and finite fields."""
    from oracle import BC_SHOCK
    from wmhelpers import wall_state
    nx, ny, ppc, extra, nppc = 24, 12, 8, 9, 4
    ntot = nx + extra + 1
    it0 = 5
    base = dict(num_process=1, n_ppc=nppc, n_x=ntot, n_x_ini=nx, n_y=ny, u_inject=0.2,
                mass_ratio=4.0, sigma_e=0.1, omega_pe=0.125, v_the=0.05, v_thi=0.05,
                theta_bn=90.0, phi_bn=90.0, l_damp_ini=4.0)
    p, up, np2, cc, uf = wall_state(BC_SHOCK, nx, ny, ppc, 73, vth=0.05, extra=extra,
                                    u0=-0.2, np_cap=nppc * ntot * 5)
    rng = np.random.default_rng(73)
    uf[2] = rng.normal(0, 0.05, uf.shape[1:])
    runs = {"one": (it0 + 1, 100000), "grow": (it0 + 16, 4)}
    for tag, (it1, expand) in runs.items():
        d = str(tmp_path / tag) + "/"
        os.makedirs(d)
        _resume_files(p, up, np2, uf, d, it0)
        cfg = {"config": {"verbose": 1, "datadir": d, "max_elapsed": 1.0e6, "max_it": it1,
                          "intvl_ptcl": 100000, "intvl_mom": 100000, "intvl_orb": 100000,
                          "intvl_expand": expand, "restart_file": "%07d_restart" % it0},
               "parameter": base}
        json.dump(cfg, open(d + "config.json", "w"), indent=2)
        r = subprocess.run([SHOCK_APP], cwd=d, capture_output=True, text=True, timeout=300)
        print(tag, r.stdout[-1200:], r.stderr[-1200:])
        assert r.returncode == 0
        assert r.stdout.count("Time step:") == it1 - it0
        final = "%07d_restart" % (it1 + 1)
        gup, gnp2, guf = (_dataset(d, final, k)[..., 0] for k in ("up", "np2", "uf"))
        nxe1 = int(_attr(d, final, "nxe")[0])
        if tag == "one":
            pa = _app_constants(p, d, final)
            assert nxe1 == p["nxe"]
            o_up, o_np2, o_cc, o_uf = (a.copy(order="F") for a in (up, np2, cc, uf))
            orc = Oracle(pa)
            gp = o_up.copy(order="F")
            orc.step(o_up, gp, o_uf, o_cc, o_np2)
            inj = gnp2 - o_np2
            assert (inj >= 0).all() and np.array_equal(inj[:, 0], inj[:, 1]) and inj.sum() > 0
            for s in range(2):
                for jl in range(ny):
                    n = o_np2[jl, s]
                    assert np.array_equal(gup[5, :n, jl, s].view(np.int64),
                                          o_up[5, :n, jl, s].view(np.int64))
                    assert np.allclose(gup[:5, :n, jl, s], o_up[:5, :n, jl, s], rtol=1e-12,
                                       atol=1e-12)
            keep = np.ones(guf.shape, bool)
            c0 = p["nxe"] - p["nxgs"] + 2
            keep[[1, 2, 4, 5], c0 - 1:c0 + 1, :] = False      # inject()'s columns
            assert np.abs((guf - o_uf)[keep]).max() <= 1e-11 * np.abs(o_uf).max()
        else:
            # relocate() at every step it with mod(it, intvl_expand) = 0
            assert nxe1 == min(p["nxge"], p["nxe"] + it1 // expand - it0 // expand) > p["nxe"]
            for s in range(2):
                for jl in range(ny):
                    x = gup[0, :gnp2[jl, s], jl, s]
                    assert (x >= p["nxs"]).all() and (x < nxe1 + 1).all()
            assert np.isfinite(guf).all() and gnp2.sum() > np2.sum()
