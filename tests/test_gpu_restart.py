"""Restart files (common/paraio.f90:93-426) through the C ABI: the JSON
layout and raw offsets the reference writes, host-array round trips, the
device-resident state written and read row by row in both particle forms,
and a run resumed from a file against the oracle resumed the same way (a
fresh process: the CG warm start starts at zero, field.f90:105-119).

The reference's paraio.f90 needs MPI-IO and is not buildable here, so the
layout is checked against the order and metadata paraio.f90:143-211 and
put_metadata (:956-1002) write, read back with Python's json and numpy."""
import json
import os

import numpy as np
import pytest

import wumingpic2d_amd as wm
from oracle import Oracle
from wmhelpers import grid_args, make_ctx, random_fields, state

pytestmark = pytest.mark.gpu

ATTRS = [("dummy_attribute", "i4", 1), ("it", "i4", 1), ("nxs", "i4", 1), ("nxe", "i4", 1),
         ("ndim", "i4", 1), ("np", "i4", 1), ("nxgs", "i4", 1), ("nxge", "i4", 1),
         ("nygs", "i4", 1), ("nyge", "i4", 1), ("nsp", "i4", 1), ("nproc", "i4", 1),
         ("delx", "f8", 1), ("delt", "f8", 1), ("c", "f8", 1), ("r", "f8", None), ("q", "f8", None)]


def _io_ctx(p, d, form=None):
    ctx = make_ctx(p, form=form)
    ctx.paraio__init(*grid_args(p), 1, 0, p["delx"], p["delt"], p["c"], p["q"], p["r"], d + "/")
    return ctx


def _read(d, name):
    meta = json.load(open(os.path.join(d, name + ".json")))
    raw = open(os.path.join(d, name + ".raw"), "rb").read()
    return meta, raw


def _get(raw, m):
    dt = {"i4": np.int32, "f8": np.float64}[m["datatype"]]
    n = m["size"] // np.dtype(dt).itemsize
    return np.frombuffer(raw, dtype=dt, count=n, offset=m["offset"])


def test_host_array_output_layout_and_round_trip(tmp_path):
    d = str(tmp_path)
    p, up, np2, cc, uf = state(16, 8, 4, 61, vth=0.1)
    random_fields(uf, 3)
    ctx = _io_ctx(p, d)
    ctx.paraio__output(up, uf, np2, p["nxgs"], p["nxge"], 123, "0000123_restart")
    meta, raw = _read(d, "0000123_restart")
    assert list(meta) == ["meta", "attribute", "dataset"]
    assert meta["meta"] == {"endian": 1, "rawfile": "0000123_restart.raw"}
    # attributes in paraio.f90's order, packed from offset 0
    off = 0
    want = {"dummy_attribute": 8, "it": 123, "nxs": p["nxgs"], "nxe": p["nxge"], "ndim": 6,
            "np": p["np"], "nxgs": p["nxgs"], "nxge": p["nxge"], "nygs": p["nygs"],
            "nyge": p["nyge"], "nsp": 2, "nproc": 1, "delx": 1.0, "delt": p["delt"], "c": 1.0,
            "r": p["r"], "q": p["q"]}
    assert list(meta["attribute"]) == [a[0] for a in ATTRS]
    for name, dt, n in ATTRS:
        m = meta["attribute"][name]
        n = n or p["nsp"]
        assert (m["datatype"], m["offset"], m["ndim"], m["shape"]) == (dt, off, 1, [n])
        assert m["size"] == n * (4 if dt == "i4" else 8)
        v = _get(raw, m)
        assert np.array_equal(v, np.atleast_1d(want[name])), name
        assert np.array_equal(np.atleast_1d(m["data"]), v)
        off += m["size"]
    ds = meta["dataset"]
    assert list(ds) == ["up", "np2", "uf"]
    assert ds["up"]["shape"] == [6, p["np"], p["ny"], 2, 1] and ds["up"]["offset"] == off
    assert ds["np2"]["shape"] == [p["ny"], 2, 1] and ds["np2"]["offset"] == off + ds["up"]["size"]
    assert ds["uf"]["shape"] == [6, p["nx"] + 4, p["ny"] + 4, 1]
    assert ds["uf"]["description"] == "electromagnetic fields including ghost cells"
    # bit patterns: the IDs in up(6) are integers whose bits may read as NaN
    assert np.array_equal(_get(raw, ds["up"]).view(np.int64), up.ravel(order="F").view(np.int64))
    assert np.array_equal(_get(raw, ds["np2"]), np2.ravel(order="F"))
    assert np.array_equal(_get(raw, ds["uf"]), uf.ravel(order="F"))
    assert len(raw) == ds["uf"]["offset"] + ds["uf"]["size"]
    # paraio__input reads it back bit for bit
    u2, f2, n2 = np.zeros_like(up), np.zeros_like(uf), np.zeros_like(np2)
    assert ctx.paraio__input(u2, f2, n2, "0000123_restart") == (p["nxgs"], p["nxge"], 123)
    assert np.array_equal(u2.view(np.int64), up.view(np.int64))
    assert np.array_equal(f2, uf) and np.array_equal(n2, np2)


def test_input_checks_and_stops(tmp_path):
    d = str(tmp_path)
    p, up, np2, cc, uf = state(16, 8, 4, 62)
    ctx = make_ctx(p)
    with pytest.raises(wm.WmError) as e:
        ctx.paraio__output(up, uf, np2, p["nxgs"], p["nxge"], 1, "x")
    assert e.value.code == wm.WM_ERR_NOT_INIT and "paraio__init" in str(e.value)
    ctx.close()
    ctx = _io_ctx(p, d)
    ctx.paraio__output(up, uf, np2, p["nxgs"], p["nxge"], 1, "a")
    # another grid: the reference's "invalid input parameters" stop
    q, *_ = state(24, 8, 4, 62)
    other = _io_ctx(q, d)
    with pytest.raises(wm.WmError) as e:
        other.paraio__input(np.zeros((6, q["np"], q["ny"], 2), order="F"),
                            np.zeros((6, q["nx"] + 4, q["ny"] + 4), order="F"),
                            np.zeros((q["ny"], 2), dtype=np.int32, order="F"), "a")
    assert "invalid input parameters" in str(e.value)
    # a dataset of the wrong shape: "Fatal error in reading particle data"
    meta = json.load(open(os.path.join(d, "a.json")))
    meta["dataset"]["up"]["shape"][1] += 1
    json.dump(meta, open(os.path.join(d, "b.json"), "w"))
    os.link(os.path.join(d, "a.raw"), os.path.join(d, "b.raw"))
    with pytest.raises(wm.WmError) as e:
        ctx.paraio__input(np.zeros_like(up), np.zeros_like(uf), np.zeros_like(np2), "b")
    assert "Fatal error in reading particle data" in str(e.value)


def test_reads_a_file_formatted_as_json_fortran_writes_it(tmp_path):
    """The reference's files come from json-fortran's printer (reals such as
    0.5E+0, its own spacing): the reader takes any JSON with the keys."""
    d = str(tmp_path)
    p, up, np2, cc, uf = state(16, 8, 4, 63)
    ctx = _io_ctx(p, d)
    ctx.paraio__output(up, uf, np2, p["nxgs"], p["nxge"], 7, "c")
    meta = json.load(open(os.path.join(d, "c.json")))
    txt = json.dumps(meta, indent=4).replace(": 1.0", ": 0.1E+1").replace(": 0.5", ": 0.5E+0")
    open(os.path.join(d, "c.json"), "w").write(txt)
    u2, f2, n2 = np.zeros_like(up), np.zeros_like(uf), np.zeros_like(np2)
    assert ctx.paraio__input(u2, f2, n2, "c")[2] == 7
    assert np.array_equal(u2.view(np.int64), up.view(np.int64))


@pytest.mark.parametrize("form", [wm.FORM_FP64, wm.FORM_FP32])
def test_device_state_round_trip(tmp_path, form):
    """The state written row by row and read into a fresh context gives the
    same rows back bit for bit (sorted by cell); the file's up holds each
    row's records in its first np2 slots and zeros after them."""
    d = str(tmp_path)
    p, up, np2, cc, uf = state(32, 16, 8, 64, vth=0.2)
    a = _io_ctx(p, d, form)
    a.state_upload(up, np2, cc, uf)
    a.step(3)
    a.save_restart(3, "s")
    ref = [np.zeros_like(up), np.zeros_like(np2), np.zeros_like(cc), np.zeros_like(uf)]
    a.state_download(*ref)
    b = _io_ctx(p, d, form)
    assert b.load_restart("s") == 3
    got = [np.zeros_like(up), np.zeros_like(np2), np.zeros_like(cc), np.zeros_like(uf)]
    b.state_download(*got)
    assert np.array_equal(got[1], ref[1]) and np.array_equal(got[2], ref[2])
    assert np.array_equal(got[3], ref[3])
    for s in range(2):
        for jl in range(p["ny"]):
            n = ref[1][jl, s]
            assert np.array_equal(got[0][:, :n, jl, s].view(np.int64), ref[0][:, :n, jl, s].view(np.int64))
    meta, raw = _read(d, "s")
    fu = _get(raw, meta["dataset"]["up"]).reshape(up.shape, order="F")
    for s in range(2):
        for jl in range(p["ny"]):
            n = ref[1][jl, s]
            assert np.array_equal(fu[:, :n, jl, s].view(np.int64), ref[0][:, :n, jl, s].view(np.int64))
            assert not fu[:, n:, jl, s].any()


def test_resumed_run_matches_the_oracle_resumed_the_same_way(tmp_path):
    """FP64: 4 steps, the state saved, a fresh context loads it and runs 4
    more; the oracle does the same with a fresh instance (CG warm start at
    zero).  Fields and particles at the FP64 bars, row order exact."""
    d = str(tmp_path)
    p, up, np2, cc, uf = state(32, 16, 8, 65, vth=0.1)
    a = _io_ctx(p, d)
    a.state_upload(up, np2, cc, uf)
    a.step(4)
    a.save_restart(4, "r")
    a.close()
    b = _io_ctx(p, d)
    assert b.load_restart("r") == 4
    b.step(4)
    orc = Oracle(p)
    gp = up.copy(order="F")
    for _ in range(4):
        orc.step(up, gp, uf, cc, np2)
    orc = Oracle(p)                                   # the restarted process
    for _ in range(4):
        orc.step(up, gp, uf, cc, np2)
    gup, gnp2, guf = np.zeros_like(up), np.zeros_like(np2), np.zeros_like(uf)
    b.state_download(up=gup, np2=gnp2, uf=guf)
    assert np.array_equal(gnp2, np2)
    for s in range(2):
        for jl in range(p["ny"]):
            n = np2[jl, s]
            assert np.array_equal(gup[5, :n, jl, s].view(np.int64), up[5, :n, jl, s].view(np.int64))
            assert np.allclose(gup[:5, :n, jl, s], up[:5, :n, jl, s], rtol=1e-12, atol=1e-12)
    assert np.abs(guf - uf).max() <= 1e-11 * np.abs(uf).max()


def test_param_mom_ptcl_orb_files(tmp_path):
    """paraio__param, __mom (velocity and temperature divided by density in
    place; cell-centred fields with the reference's Bz average), __ptcl (all
    active particles, rows in order) and __orb (positive IDs only)."""
    d = str(tmp_path)
    p, up, np2, cc, uf = state(16, 8, 4, 66, vth=0.1)
    random_fields(uf, 4)
    ctx = _io_ctx(p, d)
    ctx.paraio__param(4, 0.125, 0.025, 0.01, 0.0004, 0.01, 0.05, "init_param")
    meta, raw = _read(d, "init_param")
    names = ["dummy_attribute", "nx", "ny", "np", "delx", "delt", "c", "r", "q", "wpe", "wge",
             "wpi", "wgi", "vti", "vte", "vai", "vae", "n0"]
    assert list(meta["attribute"]) == names
    v = {k: _get(raw, meta["attribute"][k]) for k in names}
    assert v["nx"][0] == 16 and v["ny"][0] == 8 and v["n0"][0] == 4
    assert v["vai"][0] == 1.0 * 0.0004 / 0.025 and v["vae"][0] == 1.0 * 0.01 / 0.125
    rng = np.random.default_rng(6)
    mom = np.asfortranarray(rng.uniform(0.5, 2.0, (7, p["nx"] + 2, p["ny"] + 2, 2)))
    m0 = mom.copy(order="F")
    ctx.paraio__mom(mom, uf, 9)
    assert np.array_equal(mom[1:], m0[1:] / m0[:1]) and np.array_equal(mom[0], m0[0])
    meta, raw = _read(d, "0000009_mom")
    ds = meta["dataset"]
    assert list(ds) == ["den", "vel", "temp", "uf"]
    den = _get(raw, ds["den"]).reshape(ds["den"]["shape"], order="F")
    vel = _get(raw, ds["vel"]).reshape(ds["vel"]["shape"], order="F")
    assert np.array_equal(den, mom[0, 1:-1, 1:-1, :]) and np.array_equal(vel, mom[1:4, 1:-1, 1:-1, :])
    fc = _get(raw, ds["uf"]).reshape(ds["uf"]["shape"], order="F")
    i, j = 5, 3                                             # interior cell (i, j), 0-based
    U = lambda k, a, b: uf[k, a + 2, b + 2]
    assert fc[2, i, j] == (U(2, i, j) + U(2, i + 1, j) + U(2, i + 1, j) + U(2, i + 1, j + 1)) / 4
    assert fc[4, i, j] == (U(4, i, j) + U(4, i, j + 1)) / 2
    # tracers: make every third particle of species 1 positive
    ids = up[5].view(np.int64)
    ids[::3, :, 0] = -ids[::3, :, 0]
    ctx.paraio__ptcl(up, uf, np2, 10)
    ctx.paraio__orb(up, uf, np2, 10)
    for suffix, sel in (("_ptcl", lambda r: np.ones(r.shape[1], bool)),
                        ("_orb", lambda r: r[5].view(np.int64) > 0)):
        meta, raw = _read(d, "0000010" + suffix)
        ds = meta["dataset"]
        assert list(ds) == ["uf", "up01", "up02"]
        for s in range(2):
            rows = np.concatenate([up[:, :np2[jl, s], jl, s] for jl in range(p["ny"])], axis=1)
            want = rows[:, sel(rows)]
            got = _get(raw, ds["up%02d" % (s + 1)]).reshape(ds["up%02d" % (s + 1)]["shape"], order="F")
            assert np.array_equal(got.view(np.int64), want.view(np.int64))
