"""Moment diagnostics on the device vs the oracle, through the C ABI.

mom_calc__accl, mom_calc__nvt and boundary_periodic__mom as host-array
calls, and fused on the device-resident state (wm_moments) in both particle
forms.  Bars: positions copied bit-exactly and the fold bit-exact (the same
additions in the same order); the velocity update and the moment sums to
1e-12 of each moment's largest magnitude (the device contracts multiply-adds
into FMAs and adds per cell, then per tile, where the reference adds in row
order).
"""
import numpy as np
import pytest

import wumingpic2d_amd as wm
from oracle import Oracle
from wmhelpers import bits, grid_args, make_ctx, random_fields, state

pytestmark = pytest.mark.gpu

RTOL = 1e-12


def _mom_init(ctx, p):
    ctx.mom_calc__init(*grid_args(p), p["delx"], p["delt"], p["c"], p["q"], p["r"])


def _mom_close(a, b, rtol=RTOL):
    for s in range(b.shape[3]):
        for k in range(7):
            scale = max(np.abs(b[k, ..., s]).max(), 1e-300)
            err = np.abs(a[k, ..., s] - b[k, ..., s]).max()
            assert err <= rtol * scale, (k, s, err, scale)


def _oracle_moments(p, up, uf, cc, np2):
    orc = Oracle(p)
    gp = np.zeros_like(up)
    orc.mom_calc_accl(gp, up, uf, cc)
    mom = np.zeros((7, p["nx"] + 2, p["ny"] + 2, p["nsp"]), order="F")
    orc.mom_calc_nvt(mom, gp, np2)
    orc.bc_mom(mom)
    return mom


@pytest.mark.parametrize("nx,ny,ppc,seed,amp", [(32, 16, 8, 1, 0.05), (40, 11, 13, 2, 0.4),
                                                (70, 3, 5, 3, 1.0)])
def test_mom_calc_accl_matches_oracle(nx, ny, ppc, seed, amp):
    p, up, np2, cc, uf = state(nx, ny, ppc, seed, vth=0.4)
    random_fields(uf, seed, amp)
    g1 = np.zeros_like(up)
    Oracle(p).mom_calc_accl(g1, up, uf, cc)
    ctx = make_ctx(p, init=())
    _mom_init(ctx, p)
    g2 = np.full_like(up, 7.25)                  # what accl must not touch keeps this
    ctx.mom_calc__accl(g2, up, uf, cc, p["nxgs"], p["nxge"])
    for s in range(p["nsp"]):
        for jl in range(ny):
            n = cc[-1, jl, s]
            assert np.array_equal(bits(g2[0:2, :n, jl, s]), bits(up[0:2, :n, jl, s]))
            np.testing.assert_allclose(g2[2:5, :n, jl, s], g1[2:5, :n, jl, s], rtol=RTOL,
                                       atol=1e-15)
            assert (g2[5, :, jl, s] == 7.25).all() and (g2[:, n:, jl, s] == 7.25).all()


@pytest.mark.parametrize("nx,ny,ppc,seed", [(32, 16, 8, 4), (45, 9, 11, 5), (8, 1, 3, 6)])
def test_mom_calc_nvt_matches_oracle(nx, ny, ppc, seed):
    p, up, np2, cc, uf = state(nx, ny, ppc, seed, vth=0.3)
    rng = np.random.default_rng(seed)
    for s in range(p["nsp"]):                    # any order in a row, ragged and empty rows
        for jl in range(ny):
            n = int(rng.integers(0, np2[jl, s] + 1)) if jl % 4 == 3 else int(np2[jl, s])
            perm = rng.permutation(np2[jl, s])[:n]
            up[:, :n, jl, s] = up[:, perm, jl, s]
            np2[jl, s] = n
    m1 = np.zeros((7, nx + 2, ny + 2, p["nsp"]), order="F")
    Oracle(p).mom_calc_nvt(m1, up, np2)
    # mom_calc__init alone, as an app that only adds the diagnostics calls it
    ctx = make_ctx(p, init=("sort",) if seed % 2 else ())
    _mom_init(ctx, p)
    m2 = np.full_like(m1, np.nan)
    ctx.mom_calc__nvt(m2, up, np2)
    _mom_close(m2, m1)


def test_mom_calc_nvt_with_first_cell_at_zero():
    """int(x - 1/2) truncates to 0 below x = 1/2: the weights extrapolate."""
    p, up, np2, cc, uf = state(24, 6, 9, 14, vth=0.3)
    p = dict(p, nxgs=0, nxge=p["nx"] - 1, nygs=0, nyge=p["ny"] - 1, nys=0, nye=p["ny"] - 1)
    up = up.copy(order="F")
    up[0] -= 2.0
    up[1] -= 2.0
    assert (up[0, :np2[0, 0], :, 0] < 0.5).any()
    m1 = np.zeros((7, p["nx"] + 2, p["ny"] + 2, p["nsp"]), order="F")
    Oracle(p).mom_calc_nvt(m1, up, np2)
    ctx = make_ctx(p, init=("sort",))
    _mom_init(ctx, p)
    m2 = np.full_like(m1, np.nan)
    ctx.mom_calc__nvt(m2, up, np2)
    _mom_close(m2, m1)


@pytest.mark.parametrize("nx,ny", [(10, 7), (4, 1), (33, 2)])
def test_bcp_mom_bitexact(nx, ny):
    p, up, np2, cc, uf = state(nx, ny, 1, 7)
    rng = np.random.default_rng(nx + ny)
    m1 = rng.normal(size=(7, nx + 2, ny + 2, p["nsp"])).copy(order="F")
    m2 = m1.copy(order="F")
    Oracle(p).bc_mom(m1)
    ctx = make_ctx(p, init=("bc",))
    ctx.boundary_periodic__mom(m2)
    assert np.array_equal(bits(m1), bits(m2))


@pytest.mark.parametrize("nx,ny,ppc,seed,steps", [(32, 16, 8, 8, 3), (40, 11, 6, 9, 2)])
def test_resident_moments_fp64_match_oracle(nx, ny, ppc, seed, steps):
    """proj/weibel/app.f90:119-121 on the state after some steps of the loop."""
    p, up, np2, cc, uf = state(nx, ny, ppc, seed, vth=0.2)
    ctx = make_ctx(p)
    _mom_init(ctx, p)
    ctx.state_upload(up, np2, cc, uf)
    ctx.step(steps)
    gup, gnp2 = np.zeros_like(up), np.zeros_like(np2)
    gcc, guf = np.zeros_like(cc), np.zeros_like(uf)
    ctx.state_download(up=gup, np2=gnp2, cumcnt=gcc, uf=guf)
    want = _oracle_moments(p, gup, guf, gcc, gnp2)
    got = ctx.moments()
    _mom_close(got, want)
    # one order for every sum: a second call gives the same bits
    assert np.array_equal(bits(ctx.moments()), bits(got))
    # N sums to the particle count over the interior after the fold
    for s in range(p["nsp"]):
        assert abs(got[0, 1:-1, 1:-1, s].sum() - gnp2[:, s].sum()) < 1e-9 * gnp2[:, s].sum()
    # the diagnostic leaves the state as it was
    ctx.step(1)
    e = ctx.energy()
    assert np.isfinite(e).all()


@pytest.mark.parametrize("nx,ny,ppc,seed,steps", [(32, 16, 8, 10, 3), (24, 9, 12, 11, 2)])
def test_resident_moments_fp32_match_oracle(nx, ny, ppc, seed, steps):
    """The FP32 form's bins (movers of the last push settled first)."""
    p, up, np2, cc, uf = state(nx, ny, ppc, seed, vth=0.2)
    ctx = wm.Context(0)
    ctx.set_form(wm.FORM_FP32)
    g = grid_args(p)
    ctx.boundary_periodic__init(*g, 0, 0, 0, 0, 0, 0, [0], p["delx"], p["delt"], p["c"])
    ctx.particle__init(*g, p["delx"], p["delt"], p["c"], p["q"], p["r"])
    ctx.field__init(*g, 0, 0, 0, 0, p["delx"], p["delt"], p["c"], p["q"], p["r"], p["gfac"])
    ctx.sort__init(*g)
    _mom_init(ctx, p)
    ctx.state_upload(up, np2, cc, uf)
    ctx.step(steps)
    got = ctx.moments()                         # before anything else reads the bins
    gup, gnp2 = np.zeros_like(up), np.zeros_like(np2)
    gcc, guf = np.zeros_like(cc), np.zeros_like(uf)
    ctx.state_download(up=gup, np2=gnp2, cumcnt=gcc, uf=guf)
    want = _oracle_moments(p, gup, guf, gcc, gnp2)
    _mom_close(got, want)


@pytest.mark.parametrize("nsp", [1, 3])
def test_moments_other_species_counts(nsp):
    """One and three species: the species index of the launch grid and of mom."""
    p, up, np2, cc, uf = state(24, 7, 6, 15 + nsp, vth=0.2, nsp=nsp)
    random_fields(uf, nsp, 0.1)
    ctx = make_ctx(p)
    _mom_init(ctx, p)
    ctx.state_upload(up, np2, cc, uf)
    _mom_close(ctx.moments(), _oracle_moments(p, up, uf, cc, np2))
    m1 = np.zeros((7, p["nx"] + 2, p["ny"] + 2, nsp), order="F")
    Oracle(p).mom_calc_nvt(m1, up, np2)
    m2 = np.full_like(m1, np.nan)
    host = make_ctx(p)                   # the host-array form on a context of its own
    _mom_init(host, p)
    host.mom_calc__nvt(m2, up, np2)
    _mom_close(m2, m1)


def test_host_array_calls_refuse_on_a_resident_context():
    """The host-array entry points stage their arguments through the buffers
    of the device-resident state: on a context holding one they refuse and
    change nothing, so the steps after the refused calls are bit-identical
    to a run without them."""
    p, up, np2, cc, uf = state(32, 8, 8, 17, vth=0.2)
    outs = []
    for mix in (False, True):
        ctx = make_ctx(p)
        _mom_init(ctx, p)
        ctx.state_upload(up, np2, cc, uf)
        ctx.step(2)
        if mix:
            mom = np.zeros((7, p["nx"] + 2, p["ny"] + 2, p["nsp"]), order="F")
            calls = [lambda: ctx.mom_calc__nvt(mom, up, np2),
                     lambda: ctx.mom_calc__accl(np.zeros_like(up), up, uf, cc, p["nxgs"], p["nxge"]),
                     lambda: ctx.particle__solv(np.zeros_like(up), up, uf, cc, p["nxgs"], p["nxge"]),
                     lambda: ctx.field__fdtd_i(uf.copy(order="F"), up, up, cc, p["nxgs"], p["nxge"]),
                     lambda: ctx.sort__bucket(np.zeros_like(up), up, cc.copy(order="F"), np2,
                                              p["nxgs"], p["nxge"]),
                     lambda: ctx.boundary_periodic__particle_x(up.copy(order="F"), np2),
                     lambda: ctx.boundary_periodic__particle_y(up.copy(order="F"),
                                                               np2.copy(order="F"))]
            for call in calls:
                with pytest.raises(wm.WmError) as e:
                    call()
                assert e.value.code == wm.WM_ERR_ARG
                assert "device-resident state" in str(e.value)
        ctx.step(3)
        o, n, f = np.zeros_like(up), np.zeros_like(np2), np.zeros_like(uf)
        ctx.state_download(up=o, np2=n, uf=f)
        outs.append((o, n, f, ctx.moments()))
        ctx.close()
    assert np.array_equal(outs[0][1], outs[1][1])
    assert np.array_equal(outs[0][2], outs[1][2])
    assert np.array_equal(outs[0][3], outs[1][3])
    for s in range(p["nsp"]):
        for jl in range(p["ny"]):
            k = outs[0][1][jl, s]
            assert np.array_equal(outs[0][0][:, :k, jl, s].view(np.int64),
                                  outs[1][0][:, :k, jl, s].view(np.int64))


def test_nvt_particle_outside_its_row_is_reported():
    p, up, np2, cc, uf = state(16, 4, 3, 12)
    up[1, 0, 1, 0] += 2.0                       # two rows above the row that holds it
    ctx = make_ctx(p, init=("sort",))
    _mom_init(ctx, p)
    mom = np.zeros((7, 18, 6, p["nsp"]), order="F")
    with pytest.raises(wm.WmError) as e:
        ctx.mom_calc__nvt(mom, up, np2)
    assert e.value.code == wm.WM_ERR_RANGE


def test_moments_need_mom_calc_init():
    p, up, np2, cc, uf = state(16, 4, 3, 13)
    ctx = make_ctx(p)
    ctx.state_upload(up, np2, cc, uf)
    with pytest.raises(wm.WmError) as e:
        ctx.moments()
    assert e.value.code == wm.WM_ERR_NOT_INIT
