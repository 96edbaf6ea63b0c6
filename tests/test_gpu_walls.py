"""The wall boundary families on the device (FP64 form) against the oracle:
the shock's (proj/shock/boundary_shock.f90; conducting wall at nxs, an
injection edge moving with the upstream flow, an active range nxs..nxe
short of the grid) and the reconnection's (proj/reconnection/
boundary_reconnection.f90; conducting walls at nxs and nxe).

Bars as in test_gpu_parity: bit-exact for the ghost maps, folds and mirrors
(index work and exact negations/differences), 1e-12 relative for the push,
deposit and field solve, CG iteration counts equal."""
import numpy as np
import pytest

import wumingpic2d_amd as wm
from oracle import BC_RECONNECTION, BC_SHOCK, Oracle
from wmhelpers import active, bits, make_wall_ctx, random_fields, wall_state

pytestmark = pytest.mark.gpu

RTOL = 1e-12
FAMILIES = [BC_SHOCK, BC_RECONNECTION]


def _fam(ctx, bc):
    return ctx.boundary_shock if bc == BC_SHOCK else ctx.boundary_reconnection


def _close(a, b, rtol=RTOL):
    return np.abs(a - b).max() <= rtol * max(np.abs(b).max(), 1e-300)


@pytest.mark.parametrize("bc", FAMILIES)
def test_wall_ghost_maps_bitexact(bc):
    p, up, np2, cc, uf = wall_state(bc, 20, 8, 2, 31, extra=5)
    nxs, nxe, nys, nye = p["nxs"], p["nxe"], p["nys"], p["nye"]
    orc = Oracle(p)
    ctx = make_wall_ctx(p)
    W = _fam(ctx, bc)
    rng = np.random.default_rng(32)
    nx, ny = p["nx"], p["ny"]
    df = np.asfortranarray(rng.normal(size=(6, nx + 4, ny + 4)))
    a, b = df.copy(order="F"), df.copy(order="F")
    orc.bc_dfield(a)
    W.dfield(b, nxs, nxe, nys, nye, p["nxgs"], p["nxge"])
    assert np.array_equal(a, b)
    uj = np.asfortranarray(rng.normal(size=(3, nx + 4, ny + 4)))
    a, b = uj.copy(order="F"), uj.copy(order="F")
    orc.bc_curre(a)
    W.curre(b, nxs, nxe, nys, nye, p["nxgs"], p["nxge"])
    assert np.array_equal(a, b)
    # phi(nxs-1:nxe+1, nys-1:nye+1): the oracle keeps nxgs-1..nxge+1
    w = nxe - nxs + 3
    for l in (1, 2, 3):
        full = np.asfortranarray(rng.normal(size=(nx + 2, ny + 2)))
        a = full.copy(order="F")
        orc.bc_phi(a, l)
        c0 = nxs - p["nxgs"]
        b = np.asfortranarray(full[c0:c0 + w].copy())
        W.phi(b, nxs, nxe, nys, nye, l)
        assert np.array_equal(a[c0:c0 + w], b), l
    mom = np.asfortranarray(rng.normal(size=(7, nx + 2, ny + 2, p["nsp"])))
    a, b = mom.copy(order="F"), mom.copy(order="F")
    orc.bc_mom(a)
    W.mom(b)
    assert np.array_equal(a, b)


@pytest.mark.parametrize("bc", FAMILIES)
def test_wall_particle_edges_bitexact(bc):
    p, up, np2, cc, uf = wall_state(bc, 24, 8, 8, 33, vth=0.3, extra=4)
    random_fields(uf, 33, amp=0.2)
    orc = Oracle(p)
    gp = np.zeros_like(up)
    orc.particle__solv(gp, up, uf, cc)        # particles cross the edges
    ctx = make_wall_ctx(p)
    W = _fam(ctx, bc)
    nxs, nxe = p["nxs"], p["nxe"]
    edges = [("particle_x", lambda o, a, n: o.particle_x(a, n),
              lambda a, n: W.particle_x(a, n, nxs, nxe))]
    if bc == BC_SHOCK:
        edges.append(("injection", lambda o, a, n: o.bc_injection(a, n, p["u0"]),
                      lambda a, n: ctx.boundary_shock__injection(a, n, nxs, nxe, p["u0"])))
    for name, ref, dev in edges:
        a, b = gp.copy(order="F"), gp.copy(order="F")
        ref(orc, a, np2)
        dev(b, np2)
        moved = 0
        for s in range(p["nsp"]):
            for jl in range(p["ny"]):
                ra, rb = active(a, np2, jl, s, 6), active(b, np2, jl, s, 6)
                assert np.array_equal(bits(ra), bits(rb)), name
                moved += int((bits(ra) != bits(active(gp, np2, jl, s, 6))).any(axis=0).sum())
        assert moved > 0, name
    a, b = gp.copy(order="F"), gp.copy(order="F")
    na, nb = np2.copy(order="F"), np2.copy(order="F")
    orc.particle_y(a, na)
    W.particle_y(b, nb)
    assert np.array_equal(na, nb)
    for s in range(p["nsp"]):
        for jl in range(p["ny"]):
            assert np.array_equal(bits(active(a, na, jl, s, 6)), bits(active(b, nb, jl, s, 6)))


def _host_array_step(ctx, bc, p, up, gp, uf, cc, np2):
    """One pass of the family's loop through the host-array entry points
    (proj/shock/app.f90:112-117, proj/reconnection/app.f90:101-106)."""
    nxs, nxe = p["nxs"], p["nxe"]
    W = _fam(ctx, bc)
    ctx.particle__solv(gp, up, uf, cc, nxs, nxe)
    if bc == BC_SHOCK:
        ctx.boundary_shock__injection(gp, np2, nxs, nxe, p["u0"])
    ctx.field__fdtd_i(uf, up, gp, cc, nxs, nxe, bc)
    if bc != BC_SHOCK:
        W.particle_x(gp, np2, nxs, nxe)
    W.particle_y(gp, np2)
    ctx.sort__bucket(up, gp, cc, np2, nxs, nxe)


@pytest.mark.parametrize("bc", FAMILIES)
def test_wall_host_array_loop_matches_oracle(bc):
    p, up, np2, cc, uf = wall_state(bc, 24, 12, 8, 34, vth=0.15, extra=6)
    rng = np.random.default_rng(34)
    uf[2, :, :] = rng.normal(0, 0.05, uf.shape[1:])       # a random Bz drives the fields
    orc = Oracle(p)
    ctx = make_wall_ctx(p)
    o = [a.copy(order="F") for a in (up, uf, cc, np2)]
    d = [a.copy(order="F") for a in (up, uf, cc, np2)]
    ogp, dgp = o[0].copy(order="F"), d[0].copy(order="F")
    for it in range(5):
        orc.step(o[0], ogp, o[1], o[2], o[3])
        _host_array_step(ctx, bc, p, d[0], dgp, d[1], d[2], d[3])
        assert np.array_equal(orc.cg_iters(), ctx.cg_iters()), it
    assert np.array_equal(o[3], d[3])
    a, b = p["nxs"] - p["nxgs"], p["nxe"] + 1 - p["nxgs"]
    assert np.array_equal(o[2][a:b + 1], d[2][a:b + 1])
    for s in range(p["nsp"]):
        for jl in range(p["ny"]):
            assert np.array_equal(bits(o[0][5, :o[3][jl, s], jl, s]), bits(d[0][5, :d[3][jl, s], jl, s]))
            assert np.allclose(active(d[0], d[3], jl, s), active(o[0], o[3], jl, s), rtol=RTOL,
                               atol=1e-13)
    assert _close(d[1], o[1], rtol=1e-11)


@pytest.mark.parametrize("bc", FAMILIES)
def test_wall_resident_steps_match_oracle(bc):
    p, up, np2, cc, uf = wall_state(bc, 32, 16, 8, 35, vth=0.15, extra=8)
    rng = np.random.default_rng(35)
    uf[2, :, :] = rng.normal(0, 0.05, uf.shape[1:])
    ctx = make_wall_ctx(p)
    ctx.set_walls(p["nxs"], p["nxe"], p["u0"])
    ctx.state_upload(up, np2, cc, uf)
    orc = Oracle(p)
    gp = up.copy(order="F")
    for it in range(8):
        orc.step(up, gp, uf, cc, np2)
        ctx.step(1)
        assert np.array_equal(orc.cg_iters(), ctx.cg_iters()), it
    gup, gnp2, guf = np.zeros_like(up), np.zeros_like(np2), np.zeros_like(uf)
    ctx.state_download(up=gup, np2=gnp2, uf=guf)
    assert np.array_equal(gnp2, np2)
    for s in range(p["nsp"]):
        for jl in range(p["ny"]):
            assert np.array_equal(bits(up[5, :np2[jl, s], jl, s]), bits(gup[5, :np2[jl, s], jl, s]))
            assert np.allclose(active(gup, np2, jl, s), active(up, np2, jl, s), rtol=RTOL, atol=1e-13)
    assert _close(guf, uf, rtol=1e-11)


def test_wall_families_refuse_the_fp32_form_and_a_second_family():
    p, up, np2, cc, uf = wall_state(BC_RECONNECTION, 16, 8, 4, 36)
    ctx = make_wall_ctx(p)
    with pytest.raises(wm.WmError):
        ctx.boundary_periodic__init(p["ndim"], p["np"], p["nsp"], p["nxgs"], p["nxge"], p["nygs"],
                                    p["nyge"], p["nys"], p["nye"], 0, 0, 0, 0, 0, 0, [0],
                                    p["delx"], p["delt"], p["c"])
    ctx.close()
    from wmhelpers import make_ctx, grid_args
    ctx = make_ctx(p, init=("particle", "field", "sort"), form=wm.FORM_FP32)
    ctx.boundary_reconnection__init(*grid_args(p), 0, 0, 0, 0, 0, 0, [0], p["delx"], p["delt"],
                                    p["c"])
    ctx.state_upload(up, np2, cc, uf)
    with pytest.raises(wm.WmError) as e:
        ctx.step(1)
    assert e.value.code == wm.WM_ERR_ARG
