"""Shared helpers of the test suite: contexts built the way the reference
apps initialise their modules (proj/weibel/app.f90:331-347)."""
import numpy as np

import wumingpic2d_amd as wm
from wumingpic2d_amd import weibel


def grid_args(p):
    return (p["ndim"], p["np"], p["nsp"], p["nxgs"], p["nxge"], p["nygs"], p["nyge"], p["nys"],
            p["nye"])


def make_ctx(p, cap=0, mcap=0, init=("particle", "field", "sort", "bc"), form=None):
    ctx = wm.Context(0)
    if form is not None:
        ctx.set_form(form)
    if cap or mcap:
        ctx.set_capacity(cap, mcap)
    g = grid_args(p)
    if "bc" in init:
        ctx.boundary_periodic__init(*g, 0, 0, 0, 0, 0, 0, [0], p["delx"], p["delt"], p["c"])
    if "particle" in init:
        ctx.particle__init(*g, p["delx"], p["delt"], p["c"], p["q"], p["r"])
    if "field" in init:
        ctx.field__init(*g, 0, 0, 0, 0, p["delx"], p["delt"], p["c"], p["q"], p["r"], p["gfac"])
    if "sort" in init:
        ctx.sort__init(*g)
    return ctx


def state(nx, ny, ppc, seed, **kw):
    p = weibel.grid_params(nx, ny, ppc, **{k: v for k, v in kw.items() if k in ("np_cap", "delt", "wpe", "nsp")})
    up, np2, cc, uf = weibel.init_state(p, seed, **{k: v for k, v in kw.items() if k in ("vth", "udrift")})
    return p, up, np2, cc, uf


def random_fields(uf, seed, amp=0.05):
    rng = np.random.default_rng(seed)
    uf[...] = rng.normal(0.0, amp, uf.shape)
    return uf


def bits(a):
    return np.ascontiguousarray(a).view(np.int64)


def active(up, np2, jl, s, ndim=5):
    return up[:ndim, :np2[jl, s], jl, s]


def rho_tsc(p, up, np2):
    """Charge on cells from the TSC shape of ele_cur (field.f90:224-236):
    q * s0x * s0y around the particle's cell, periodic wrap."""
    nx, ny = p["nx"], p["ny"]
    rho = np.zeros((nx, ny))
    for s in range(p["nsp"]):
        for jl in range(ny):
            n = np2[jl, s]
            x, y = up[0, :n, jl, s], up[1, :n, jl, s]
            i, j = x.astype(int), y.astype(int)
            for sx, dxo in zip(tsc_weights(x - 0.5 - i), (-1, 0, 1)):
                for sy, dyo in zip(tsc_weights(y - 0.5 - j), (-1, 0, 1)):
                    np.add.at(rho, ((i + dxo - p["nxgs"]) % nx, (j + dyo - p["nys"]) % ny),
                              p["q"][s] * sx * sy)
    return rho


def tsc_weights(dh):
    return (0.5 * (0.5 - dh) ** 2, 0.75 - dh * dh, 0.5 * (0.5 + dh) ** 2)


def gauss_residual(p, uf, up, np2):
    """div+ E - 4 pi rho at the cell centres (interior)."""
    ex = uf[3, 2:-2, 2:-2]
    ey = uf[4, 2:-2, 2:-2]
    div = (np.roll(ex, -1, 0) - ex) + (np.roll(ey, -1, 1) - ey)
    return div - 4 * np.pi * rho_tsc(p, up, np2)


def wall_state(bc, nx, ny, ppc, seed, vth=0.1, extra=0, u0=-0.2, np_cap=None):
    """A state for a wall family (oracle.BC_SHOCK / BC_RECONNECTION): the
    Weibel pairs, on a grid `extra` columns wider than they fill for the
    shock (its box nxs..nxe grows into the rest, proj/shock/app.f90:613-623;
    its particles fill [nxs, nxe), :436), with none left deep in cell nxs
    (the shock mirrors a particle pushed below x = nxs two cells up, a move
    its deposit does not index).  p carries bc, nxs, nxe and u0."""
    from oracle import BC_SHOCK, Oracle

    kw = {"np_cap": np_cap} if np_cap else {}
    p, up, np2, cc, uf = state(nx, ny, ppc, seed, vth=vth, **kw)
    p = dict(p, bc=bc, nxs=p["nxgs"], nxe=p["nxge"], u0=u0 if bc == BC_SHOCK else 0.0)
    if bc != BC_SHOCK:
        return p, up, np2, cc, uf
    q = dict(p, nx=nx + extra + 1, nxge=p["nxge"] + extra + 1)
    q["nxe"] = q["nxge"] - extra
    cc = np.asfortranarray(np.concatenate([cc, np.repeat(cc[-1:], extra + 1, axis=0)], axis=0))
    uf2 = np.zeros((6, q["nx"] + 4, ny + 4), order="F")
    uf2[:, :nx + 4] = uf
    x = up[0]
    x[x < q["nxs"] + 1] += 1.0
    up2 = np.zeros_like(up)
    Oracle(q).sort__bucket(up2, up, cc, np2)
    return q, up2, np2, cc, uf2


def make_wall_ctx(p, init=("particle", "field", "sort")):
    """A context with the wall family of p["bc"] initialised (its __init
    takes boundary_periodic__init's arguments)."""
    from oracle import BC_SHOCK

    ctx = make_ctx(p, init=init)
    init_bc = ctx.boundary_shock__init if p["bc"] == BC_SHOCK else ctx.boundary_reconnection__init
    init_bc(*grid_args(p), 0, 0, 0, 0, 0, 0, [0], p["delx"], p["delt"], p["c"])
    return ctx
