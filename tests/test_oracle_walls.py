"""The oracle's wall boundary families (CPU, no GPU): the shock's
(proj/shock/boundary_shock.f90) and the reconnection's
(proj/reconnection/boundary_reconnection.f90) conducting walls in x with
periodic y, and the shock's moving injection edge.

Pins:
  * CG with wall ghost columns (bc_phi, :551-625 / :505-579): the discrete
    Helmholtz operator with phi(nxs-1) = -phi(nxs), phi(nxe+1) = 0 (Bx) has
    the eigenvectors sin(k (i - nxs + 1/2)), k = m pi / (N + 1/2), and with
    phi(nxs-1) = phi(nxs+1), phi(nxe+1) = 0 (By, Bz) the eigenvectors
    cos(k (i - nxs)), k = (m + 1/2) pi / N, each with eigenvalue
    f4 - 2 cos k - 2 cos ky: an eigenvector right-hand side must come back
    divided by it.  A periodic, sign-flipped or shifted ghost rule fails.
  * Whole steps of both families keep div E - 4 pi rho at roundoff in the
    cells the walls do not reach (the curl terms of dE have zero divergence
    and the deposit conserves charge), so the wall field solve, the y-only
    current fold and the injection edge are consistent with Maxwell's
    equations away from the walls.
  * The particle walls are mirrors: x + x' = 2 x_wall with u' = -u
    (boundary_shock.f90:77-89); the injection edge mirrors x about xend and
    u_x about u0 (:279-290).  These statements are the reference's own; the
    mirror identities and the ensemble's bounds are what the test checks.
  * Field and current ghosts: the conducting-wall relations of
    boundary_shock.f90:396-405 and boundary_reconnection.f90:350-359, y rows
    copied for the active columns only, the current folded in y and not in x.
"""
import math

import numpy as np
import pytest

from oracle import BC_PERIODIC, BC_RECONNECTION, BC_SHOCK, Oracle
from wmhelpers import state, tsc_weights
from wumingpic2d_amd import weibel


def _consts(p):
    t = 1.0 / (p["c"] * p["delt"] * p["gfac"])
    return 4 + t * t, t * t          # f4, f5 (field.f90:56-57)


@pytest.mark.parametrize("bc,nact", [(BC_RECONNECTION, 20), (BC_SHOCK, 14)])
@pytest.mark.parametrize("m,my", [(1, 0), (3, 2), (7, 1)])
def test_wall_cg_eigenvectors(bc, nact, m, my):
    nx, ny = 20, 8
    p = dict(weibel.grid_params(nx, ny, 1), bc=bc)
    nxs = p["nxgs"]
    p["nxs"], p["nxe"] = nxs, nxs + nact - 1
    orc = Oracle(p)
    f4, f5 = _consts(p)
    I = np.arange(nact)[:, None]
    J = np.arange(ny)[None, :]
    ky = 2 * math.pi * my / ny
    modes = [np.sin(m * math.pi / (nact + 0.5) * (I + 0.5)),       # Bx: odd wall, zero beyond
             np.cos((m + 0.5) * math.pi / nact * I),              # By, Bz: even about nxs
             np.cos((m + 0.5) * math.pi / nact * I)]
    ks = [m * math.pi / (nact + 0.5), (m + 0.5) * math.pi / nact, (m + 0.5) * math.pi / nact]
    gkl = np.zeros((3, nx, ny), order="F")
    for l in range(3):
        gkl[l, :nact] = modes[l] * np.cos(ky * J + 0.3 * l) / f5
    db = np.zeros((6, nx + 4, ny + 4), order="F")
    st, it = orc.cgm(db, gkl)
    assert st == 0
    for l in range(3):
        lam = f4 - 2 * math.cos(ks[l]) - 2 * math.cos(ky)
        want = f5 * gkl[l, :nact] / lam
        assert np.abs(db[l, 2:2 + nact, 2:-2] - want).max() < 1e-13 * np.abs(want).max()
        assert not db[l, 2 + nact:, :].any()      # beyond the active range: untouched
    # the same right-hand side with periodic ghosts is not an eigenvector
    per = Oracle(dict(p, bc=BC_PERIODIC))
    db2 = np.zeros_like(db)
    per.cgm(db2, gkl)
    assert np.abs(db2[:3, 2:2 + nact, 2:-2] - db[:3, 2:2 + nact, 2:-2]).max() > 1e-6


def _ensemble(p, n, rng, lo, hi):
    """n particles per row of species 0 spread over [lo, hi) in x."""
    ny = p["ny"]
    up = np.zeros((p["ndim"], p["np"], ny, p["nsp"]), order="F")
    np2 = np.zeros((ny, p["nsp"]), dtype=np.int32, order="F")
    for jl in range(ny):
        up[0, :n, jl, 0] = rng.uniform(lo, hi, n)
        up[1, :n, jl, 0] = p["nys"] + jl + rng.uniform(0, 1, n)
        up[2:5, :n, jl, 0] = rng.normal(0, 0.3, (3, n))
        np2[jl, 0] = n
    return up, np2


@pytest.mark.parametrize("bc", [BC_SHOCK, BC_RECONNECTION])
def test_wall_particle_x_mirrors(bc):
    p = dict(weibel.grid_params(16, 4, 4), bc=bc)
    nxs, nxe = p["nxgs"], p["nxge"]
    orc = Oracle(p)
    rng = np.random.default_rng(3)
    up, np2 = _ensemble(p, 60, rng, nxs, nxe + 1)
    before = up.copy(order="F")
    orc.particle_x(up, np2)
    x0, x1 = before[0, :60, :, 0].ravel(), up[0, :60, :, 0].ravel()
    u0 = before[2:5, :60, :, 0].reshape(3, -1)
    u1 = up[2:5, :60, :, 0].reshape(3, -1)
    left = x0.astype(int) < nxs + 1
    right = x0.astype(int) >= nxe - 1
    assert left.any() and right.any() and (~left & ~right).any()
    assert np.array_equal(x0[left] + x1[left], np.full(left.sum(), 2.0 * (nxs + 1)))
    assert np.allclose(x0[right] + x1[right], 2.0 * (nxe - 1), rtol=0, atol=1e-14)
    flip = left | right
    assert np.array_equal(u1[:, flip], -u0[:, flip])
    assert np.array_equal(x1[~flip], x0[~flip]) and np.array_equal(u1[:, ~flip], u0[:, ~flip])
    assert (x1 >= nxs + 1).all() and (x1 < nxe).all()


def test_injection_edge_mirrors_about_xend_and_u0():
    p = dict(weibel.grid_params(16, 4, 4), bc=BC_SHOCK)
    nxs, nxe = p["nxgs"], p["nxgs"] + 11
    p["nxs"], p["nxe"] = nxs, nxe
    orc = Oracle(p)
    u0 = -0.3
    xend = nxe * p["delx"] + u0 / math.sqrt(1 + u0 * u0 / p["c"] ** 2) * p["delt"]
    rng = np.random.default_rng(4)
    up, np2 = _ensemble(p, 80, rng, nxs, nxe + 1)
    before = up.copy(order="F")
    orc.bc_injection(up, np2, u0)
    x0, x1 = before[0, :80, :, 0].ravel(), up[0, :80, :, 0].ravel()
    left = x0.astype(int) < nxs + 1
    beyond = ~left & (x0 > xend)
    assert left.any() and beyond.any()
    assert np.array_equal(x0[left] + x1[left], np.full(left.sum(), 2.0 * (nxs + 1)))
    assert np.allclose(x0[beyond] + x1[beyond], 2 * xend, rtol=0, atol=1e-14)
    ux0, ux1 = before[2, :80, :, 0].ravel(), up[2, :80, :, 0].ravel()
    assert np.allclose((ux1 - u0)[beyond], -(ux0 - u0)[beyond], rtol=0, atol=1e-15)
    v0 = before[:, :80, :, 0].reshape(p["ndim"], -1)
    v1 = up[:, :80, :, 0].reshape(p["ndim"], -1)
    assert np.array_equal(v1[3:5][:, beyond], -v0[3:5][:, beyond])
    keep = ~left & ~beyond
    assert np.array_equal(v1[:, keep], v0[:, keep])


@pytest.mark.parametrize("bc", [BC_SHOCK, BC_RECONNECTION])
def test_wall_field_ghosts(bc):
    nx, ny = 12, 6
    p = dict(weibel.grid_params(nx, ny, 1), bc=bc)
    nact = 9 if bc == BC_SHOCK else nx
    nxs, nxe = p["nxgs"], p["nxgs"] + nact - 1
    p["nxs"], p["nxe"] = nxs, nxe
    orc = Oracle(p)
    rng = np.random.default_rng(5)
    df = np.asfortranarray(rng.normal(size=(6, nx + 4, ny + 4)))
    d0 = df.copy(order="F")
    orc.bc_dfield(df)
    c = lambda i: i - p["nxgs"] + 2                    # reference column -> array index
    rows = slice(0, ny + 4)
    # y ghost rows of the active columns are periodic copies of the interior
    a = slice(c(nxs), c(nxe) + 1)
    for k in range(6):
        if bc == BC_RECONNECTION and k in (0, 4, 5):
            a = slice(c(nxs), c(nxe))                  # column nxe is rewritten below
        else:
            a = slice(c(nxs), c(nxe) + 1)
        assert np.array_equal(df[k, a, ny + 2], d0[k, a, 2])
        assert np.array_equal(df[k, a, 0], d0[k, a, ny])
    # left wall: Bx, Ey, Ez odd about the wall, By, Bz, Ex even about column nxs
    L = c(nxs)
    for k in (0, 4, 5):
        assert np.array_equal(df[k, L - 1, rows], -df[k, L, rows])
    for k in (1, 2, 3):
        assert np.array_equal(df[k, L - 1, rows], df[k, L + 1, rows])
    R = c(nxe)
    if bc == BC_SHOCK:
        assert not df[:, R + 1, rows].any()
        assert np.array_equal(df[:, R + 2:, :], d0[:, R + 2:, :])     # beyond: untouched
    else:
        for k in (0, 4, 5):
            assert np.array_equal(df[k, R, rows], -df[k, R - 1, rows])
        for k in (1, 2, 3):
            assert np.array_equal(df[k, R + 1, rows], df[k, R - 1, rows])
    # the outermost ghost columns are never written by a wall family
    assert np.array_equal(df[:, L - 2, :], d0[:, L - 2, :])


def test_wall_current_folds_in_y_only():
    nx, ny = 10, 6
    p = dict(weibel.grid_params(nx, ny, 1), bc=BC_RECONNECTION)
    orc = Oracle(p)
    per = Oracle(dict(p, bc=BC_PERIODIC))
    rng = np.random.default_rng(6)
    uj = np.asfortranarray(rng.normal(size=(3, nx + 4, ny + 4)))
    a, b = uj.copy(order="F"), uj.copy(order="F")
    orc.bc_curre(a)
    per.bc_curre(b)
    # per column, the y fold conserves the sum over the interior rows of
    # what the ghost rows held; nothing moves between columns
    tot = uj[:, :, 2:ny + 2].sum(axis=2) + uj[:, :, [0, 1, ny + 2, ny + 3]].sum(axis=2)
    assert np.allclose(a[:, :, 2:ny + 2].sum(axis=2), tot, rtol=0, atol=1e-12)
    # the columns the periodic x fold does not touch agree with it; the four
    # it folds into (nxs, nxs+1, nxe-1, nxe) keep only their own y fold
    assert np.array_equal(a[:, 4:nx, :], b[:, 4:nx, :])
    for c in (2, 3, nx, nx + 1):
        assert not np.allclose(a[:, c, 2:ny + 2], b[:, c, 2:ny + 2])
    # the ghost columns keep their own (y-folded) values instead of images
    assert not np.allclose(a[:, 0:2, 2:ny + 2], a[:, nx:nx + 2, 2:ny + 2])


def test_wall_moments_fold_into_their_own_edge():
    nx, ny, nsp = 8, 5, 2
    p = dict(weibel.grid_params(nx, ny, 1), bc=BC_SHOCK)
    orc = Oracle(p)
    rng = np.random.default_rng(7)
    mom = np.asfortranarray(rng.normal(size=(7, nx + 2, ny + 2, nsp)))
    m0 = mom.copy(order="F")
    orc.bc_mom(mom)
    # the x ghosts land on their own edge column, then rows fold in y
    e = m0.copy(order="F")
    e[:, 1] += e[:, 0]
    e[:, nx] += e[:, nx + 1]
    e[:, :, ny] += e[:, :, 0]
    e[:, :, 1] += e[:, :, ny + 1]
    assert np.allclose(mom, e, rtol=0, atol=1e-14)


def _gauss(p, uf, up, np2):
    """div E - 4 pi rho (forward differences, rho from the deposit's TSC
    shape, no wrap in x) on the interior rows."""
    nx, ny = p["nx"], p["ny"]
    rho = np.zeros((nx + 4, ny))
    for s in range(p["nsp"]):
        for jl in range(ny):
            n = np2[jl, s]
            x, y = up[0, :n, jl, s], up[1, :n, jl, s]
            i, j = x.astype(int), y.astype(int)
            for sx, dxo in zip(tsc_weights(x - 0.5 - i), (-1, 0, 1)):
                for sy, dyo in zip(tsc_weights(y - 0.5 - j), (-1, 0, 1)):
                    np.add.at(rho, (i + dxo - p["nxgs"] + 2, (j + dyo - p["nys"]) % ny),
                              p["q"][s] * sx * sy)
    ex, ey = uf[3, :, 2:-2], uf[4, :, 2:-2]
    div = (np.roll(ex, -1, 0) - ex) + (np.roll(ey, -1, 1) - ey)
    return div - 4 * math.pi * rho


def _widen(p, cc, uf, extra):
    """The same state on a grid `extra` columns wider in x (empty columns)."""
    q = dict(p, nx=p["nx"] + extra, nxge=p["nxge"] + extra)
    cc2 = np.concatenate([cc, np.repeat(cc[-1:], extra, axis=0)], axis=0)
    uf2 = np.zeros((6, q["nx"] + 4, p["ny"] + 4))
    uf2[:, :p["nx"] + 4] = uf
    return q, np.asfortranarray(cc2), np.asfortranarray(uf2)


@pytest.mark.parametrize("bc", [BC_RECONNECTION, BC_SHOCK])
def test_wall_steps_keep_gauss_law_away_from_the_walls(bc):
    """Pairs start co-located (rho = 0) in a random B; six whole steps of the
    family's loop (proj/reconnection/app.f90:101-106, proj/shock/app.f90:
    112-117).  Cells at least 6 from a wall or the injection edge (the
    mirrors move charge without a current within 5) keep div E - 4 pi rho
    at roundoff.  (div B is not a pin here: the wall ghost rules of the B
    solve do not commute with the divergence, so the CG's 1e-6 residual
    leaves a div B that decays from the walls.)"""
    nx, ny = 24, 12
    p, up, np2, cc, uf = state(nx, ny, 6, 8, vth=0.3)
    p = dict(p, bc=bc)
    nxs, nxe = p["nxgs"], p["nxge"]
    if bc == BC_SHOCK:
        # the shock's particles fill [nxs, nxe) (proj/shock/app.f90:436):
        # two more columns on the right, beyond the active range
        p, cc, uf = _widen(p, cc, uf, 2)
        nx = p["nx"]
        nxe = p["nxe"] = p["nxge"] - 1
        p["u0"] = -0.2
        # and none deep in cell nxs: one pushed below x = nxs would be
        # mirrored two cells up, a move the deposit does not index
        x = up[0]
        x[x < nxs + 1] += 1.0
        Oracle(p).sort__bucket(up2 := np.zeros_like(up), up, cc, np2)
        up = up2
    rng = np.random.default_rng(8)
    bz = rng.normal(0, 0.05, (nx + 4, ny))
    uf[2, :, 2:-2] = bz
    uf[2, :, :2], uf[2, :, -2:] = uf[2, :, -4:-2], uf[2, :, 2:4]
    orc = Oracle(p)
    gp = up.copy(order="F")
    g0 = _gauss(p, uf, up, np2)
    # a particle in cell nxe is mirrored about nxe - 1 into cell nxe - 3 and
    # its shape reaches nxe - 4: keep 6 cells from either edge
    keep = slice(nxs - p["nxgs"] + 2 + 6, nxe - p["nxgs"] + 2 - 6 + 1)
    emax = 0.0
    for _ in range(6):
        orc.step(up, gp, uf, cc, np2)
        g = _gauss(p, uf, up, np2)
        emax = max(emax, np.abs(uf[3:5]).max())
        assert np.abs(g[keep] - g0[keep]).max() < 1e-12 * max(emax, 1e-3)
    assert emax > 1e-4
