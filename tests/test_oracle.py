"""Pins of the CPU oracle (FP64 C restatement of the reference's hot path).

Each oracle function is checked against what fixes it independently of its
own formula: the reference's own particle/sort modules compiled unchanged
(oracle/_ref) and the golden vectors they produced (tests/golden), closed
forms of the Boris rotation and of TSC interpolation, the charge-conservation
identity of the Esirkepov scheme, exact Helmholtz solutions, Gauss's law over
whole steps, and brute-force index maps.  The migration ORDER of
boundary_periodic__particle_y has no such pin (the file cannot be built here);
its test checks invariants and agreement with a second, closed-form
formulation only.
"""
import math
import os
from fractions import Fraction

import numpy as np
import pytest

from oracle import OR_ERR_CG_ITEMAX, Oracle, RefBuild, ref_available
from wumingpic2d_amd import weibel
from wmhelpers import bits, random_fields, state
from wmhelpers import rho_tsc as _rho, tsc_weights as _w

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "ref_particle_sort.npz")


# ------------------------------------------------------- reference build
@pytest.mark.skipif(not ref_available(), reason="oracle/_ref not built (no reference tree)")
@pytest.mark.parametrize("nx,ny,ppc,seed,amp", [(32, 16, 8, 1, 0.05), (20, 7, 13, 2, 0.4),
                                                (64, 4, 3, 3, 1.0)])
def test_push_bitexact_vs_reference_build(nx, ny, ppc, seed, amp):
    p, up, np2, cc, uf = state(nx, ny, ppc, seed, vth=0.4)
    random_fields(uf, seed, amp)
    g1, g2 = np.zeros_like(up), np.zeros_like(up)
    Oracle(p).particle__solv(g1, up, uf, cc)
    RefBuild(p).particle__solv(g2, up, uf, cc)
    assert np.array_equal(bits(g1), bits(g2))


@pytest.mark.skipif(not ref_available(), reason="oracle/_ref not built (no reference tree)")
@pytest.mark.parametrize("nx,ny,ppc,seed", [(32, 16, 8, 4), (9, 5, 11, 5)])
def test_sort_bitexact_vs_reference_build(nx, ny, ppc, seed):
    p, up, np2, cc, uf = state(nx, ny, ppc, seed)
    rng = np.random.default_rng(seed)
    for s in range(2):
        for jl in range(ny):
            n = int(rng.integers(0, np2[jl, s] + 1)) if jl % 3 == 1 else int(np2[jl, s])
            perm = rng.permutation(np2[jl, s])[:n]
            up[:, :n, jl, s] = up[:, perm, jl, s]
            np2[jl, s] = n
    g1, g2 = np.zeros_like(up), np.zeros_like(up)
    c1, c2 = np.zeros_like(cc), np.zeros_like(cc)
    Oracle(p).sort__bucket(g1, up, c1, np2)
    RefBuild(p).sort__bucket(g2, up, c2, np2)
    assert np.array_equal(c1, c2)
    for s in range(2):
        for jl in range(ny):
            n = np2[jl, s]
            assert np.array_equal(bits(g1[:, :n, jl, s]), bits(g2[:, :n, jl, s]))


# ------------------------------------------------------- golden vectors
@pytest.mark.parametrize("case", ["a", "b"])
def test_push_and_sort_golden(case):
    z = np.load(GOLDEN)
    nx, ny, ppc, npc = (int(v) for v in z[f"{case}_grid"])
    p = weibel.grid_params(nx, ny, ppc, np_cap=npc)
    f = lambda k: np.asfortranarray(z[f"{case}_{k}"])
    orc = Oracle(p)
    gp = np.zeros_like(f("up"))
    orc.particle__solv(gp, f("up"), f("uf"), f("cumcnt"))
    assert np.array_equal(bits(gp), bits(f("push_gp")))
    gs = np.zeros_like(f("up"))
    cs = np.zeros_like(f("cumcnt"))
    n2 = f("sort_np2")
    orc.sort__bucket(gs, f("sort_in"), cs, n2)
    assert np.array_equal(cs, f("sort_cumcnt"))
    ref = f("sort_gp")
    for s in range(2):
        for jl in range(ny):
            n = n2[jl, s]
            assert np.array_equal(bits(gs[:, :n, jl, s]), bits(ref[:, :n, jl, s]))


def test_sort_is_a_stable_argsort():
    p, up, np2, cc, uf = state(24, 6, 9, 6)
    rng = np.random.default_rng(6)
    for s in range(2):
        for jl in range(6):
            perm = rng.permutation(np2[jl, s])
            up[:, :np2[jl, s], jl, s] = up[:, perm, jl, s]
    g, c = np.zeros_like(up), np.zeros_like(cc)
    Oracle(p).sort__bucket(g, up, c, np2)
    for s in range(2):
        for jl in range(6):
            n = np2[jl, s]
            key = up[0, :n, jl, s].astype(np.int64) - p["nxgs"]
            order = np.argsort(key, kind="stable")
            assert np.array_equal(bits(g[:, :n, jl, s]), bits(up[:, order, jl, s]))
            cnt = np.bincount(key, minlength=p["nx"])
            assert np.array_equal(c[:, jl, s], np.concatenate([[0], np.cumsum(cnt)]))


# --------------------------------------------------- push: closed forms
def _single(p, x, y, u, s=0):
    """One particle of species s in an otherwise empty state."""
    up = np.zeros((6, p["np"], p["ny"], p["nsp"]), order="F")
    np2 = np.zeros((p["ny"], p["nsp"]), dtype=np.int32, order="F")
    cc = np.zeros((p["nx"] + 1, p["ny"], p["nsp"]), dtype=np.int32, order="F")
    jl = int(y) - p["nys"]
    i = int(x) - p["nxgs"]
    up[:5, 0, jl, s] = [x, y, *u]
    np2[jl, s] = 1
    cc[i + 1:, jl, s] = 1
    return up, np2, cc, jl


def test_boris_rotation_in_uniform_bz():
    p = weibel.grid_params(16, 16, 1)
    uf = np.zeros((6, 20, 20), order="F")
    B0 = 0.7
    uf[2] = B0
    u = np.array([0.3, -0.2, 0.1])
    up, np2, cc, jl = _single(p, 7.3, 9.6, u)
    gp = np.zeros_like(up)
    Oracle(p).particle__solv(gp, up, uf, cc)
    un = gp[2:5, 0, jl, 0]
    assert abs(np.linalg.norm(un) - np.linalg.norm(u)) < 1e-15
    # rotation by 2 atan(t), t = (q/m) (dt/2) B / gamma (the Boris angle)
    q, r = p["q"][0], p["r"][0]
    gam = math.sqrt(p["c"] ** 2 + u @ u)
    t = q / r * 0.5 * p["delt"] * B0 / gam
    th = -2.0 * math.atan(t)
    rot = np.array([u[0] * math.cos(th) - u[1] * math.sin(th),
                    u[0] * math.sin(th) + u[1] * math.cos(th), u[2]])
    assert np.allclose(un, rot, rtol=0, atol=1e-15)
    ig = 1.0 / math.sqrt(1 + un @ un)
    assert abs(gp[0, 0, jl, 0] - (7.3 + un[0] * p["delt"] * ig)) < 1e-14
    assert abs(gp[1, 0, jl, 0] - (9.6 + un[1] * p["delt"] * ig)) < 1e-14


def test_electric_acceleration_reproduces_linear_fields():
    """B = 0: u_new = u + (q/m) dt E(x_p).  Linear E at the Yee positions
    Ex(i, j+1/2), Ey(i+1/2, j), Ez(i+1/2, j+1/2) is interpolated exactly."""
    p = weibel.grid_params(16, 16, 1)
    nxg = p["nx"] + 4
    I, J = np.meshgrid(np.arange(nxg) + p["nxgs"] - 2, np.arange(nxg) + p["nys"] - 2, indexing="ij")
    uf = np.zeros((6, nxg, nxg), order="F")
    lin = lambda X, Y, a, b, c: a * X + b * Y + c
    coef = [(0.01, -0.02, 0.3), (-0.015, 0.005, -0.1), (0.02, 0.01, 0.05)]
    uf[3] = lin(I, J + 0.5, *coef[0])
    uf[4] = lin(I + 0.5, J, *coef[1])
    uf[5] = lin(I + 0.5, J + 0.5, *coef[2])
    rng = np.random.default_rng(3)
    orc = Oracle(p)
    for _ in range(20):
        x, y = rng.uniform(5, 14, 2)
        up, np2, cc, jl = _single(p, x, y, [0.0, 0.0, 0.0])
        gp = np.zeros_like(up)
        orc.particle__solv(gp, up, uf, cc)
        qm_dt = p["q"][0] / p["r"][0] * p["delt"]
        E = gp[2:5, 0, jl, 0] / qm_dt
        want = [lin(x, y, *c) for c in coef]
        assert np.allclose(E, want, rtol=0, atol=1e-14)


@pytest.mark.parametrize("comp", [0, 1, 2])
def test_magnetic_gather_reproduces_linear_fields(comp):
    """E = 0: the Boris step satisfies u+ - u- = (u+ + u-) x t with
    t = (q/m)(dt/2) B / gamma, which recovers B at the particle from a push
    with u perpendicular to the tested component.  Linear B at
    Bx(i+1/2, j), By(i, j+1/2), Bz(i, j) is interpolated exactly."""
    p = weibel.grid_params(16, 16, 1)
    nxg = p["nx"] + 4
    I, J = np.meshgrid(np.arange(nxg) + p["nxgs"] - 2, np.arange(nxg) + p["nys"] - 2, indexing="ij")
    uf = np.zeros((6, nxg, nxg), order="F")
    a, b, c = 0.02, -0.03, 0.4
    pos = [(I + 0.5, J), (I, J + 0.5), (I, J)][comp]
    uf[comp] = a * pos[0] + b * pos[1] + c
    rng = np.random.default_rng(comp)
    orc = Oracle(p)
    for _ in range(10):
        x, y = rng.uniform(5, 14, 2)
        u = np.array([0.2, 0.3, -0.1])
        u[comp] = 0.0
        up, np2, cc, jl = _single(p, x, y, u)
        gp = np.zeros_like(up)
        orc.particle__solv(gp, up, uf, cc)
        un = gp[2:5, 0, jl, 0]
        gam = math.sqrt(1.0 + u @ u)
        d, s_ = un - u, un + u
        # d = s x t with t along the tested component: least squares for t
        tv = np.zeros(3)
        tv[comp] = 1.0
        unit = np.cross(s_, tv)
        tc = (d @ unit) / (unit @ unit)
        B = tc * gam / (p["q"][0] / p["r"][0] * 0.5 * p["delt"])
        assert abs(B - (a * x + b * y + c)) < 1e-12


# ------------------------------------------------ deposit: invariants
def _fold(p, uj3):
    """Periodic sum of a ghost-inclusive (nx+4, ny+4) array onto the interior."""
    nx, ny = p["nx"], p["ny"]
    out = np.zeros((nx, ny))
    for a in range(nx + 4):
        for b in range(ny + 4):
            out[(a - 2) % nx, (b - 2) % ny] += uj3[a, b]
    return out


def test_deposit_charge_continuity_and_moments():
    p, up, np2, cc, uf = state(12, 10, 5, 7, vth=0.45)
    random_fields(uf, 7, 0.3)
    orc = Oracle(p)
    gp = np.zeros_like(up)
    orc.particle__solv(gp, up, uf, cc)
    uj = np.zeros((3, p["nx"] + 4, p["ny"] + 4), order="F")
    orc.ele_cur(uj, up, gp, cc)
    # discrete continuity: rho1 - rho0 + dt/dx (Jx(i+1)-Jx(i) + Jy(j+1)-Jy(j)) = 0
    jx, jy = _fold(p, uj[0]), _fold(p, uj[1])
    r0 = _rho(p, up, np2)
    r1 = _rho(p, gp, np2)
    div = (np.roll(jx, -1, 0) - jx) + (np.roll(jy, -1, 1) - jy)
    scale = np.abs(r1 - r0).max()
    assert scale > 1e-4
    assert np.abs(r1 - r0 + p["delt"] * div).max() < 1e-13 * scale
    # first moments: sum Jx = sum q dx/dt, sum Jz = sum q vz
    tot = np.zeros(3)
    for s in range(2):
        for jl in range(p["ny"]):
            n = np2[jl, s]
            tot[0] += p["q"][s] * (gp[0, :n, jl, s] - up[0, :n, jl, s]).sum() / p["delt"]
            tot[1] += p["q"][s] * (gp[1, :n, jl, s] - up[1, :n, jl, s]).sum() / p["delt"]
            u = gp[2:5, :n, jl, s]
            tot[2] += p["q"][s] * (u[2] / np.sqrt(1 + (u * u).sum(0))).sum()
    for k in range(3):
        assert abs(uj[k].sum() - tot[k]) < 1e-12 * max(1.0, np.abs(uj[k]).sum())


def test_deposit_jz_of_a_resting_particle():
    """No motion in x, y: Jz = q gvz s0x s0y exactly (field.f90:292-298 with ds = 0)."""
    p = weibel.grid_params(12, 12, 1)
    up, np2, cc, jl = _single(p, 6.3, 7.8, [0.0, 0.0, 0.4])
    gp = up.copy(order="F")
    gp[2:5, 0, jl, 0] = [0.0, 0.0, 0.4]
    uj = np.zeros((3, 16, 16), order="F")
    Oracle(p).ele_cur(uj, up, gp, cc)
    gvz = 0.4 / math.sqrt(1 + 0.16)
    sx, sy = _w(6.3 - 0.5 - 6), _w(7.8 - 0.5 - 7)
    for a in range(3):
        for b in range(3):
            assert uj[2, 6 - 1 + a - p["nxgs"] + 2, 7 - 1 + b - p["nys"] + 2] == pytest.approx(
                p["q"][0] * gvz * sx[a] * sy[b], rel=1e-15, abs=1e-18)
    assert np.abs(uj[0]).max() == 0.0 and np.abs(uj[1]).max() == 0.0


# ------------------------------------------------------------- CG (cgm)
def _cg_case(nx, ny, b_fn, warm=None, err=1e-6):
    p = weibel.grid_params(nx, ny, 1)
    orc = Oracle(p)
    f1 = p["c"] * p["delt"]
    t = 1.0 / (p["c"] * p["delt"] * p["gfac"])
    f5 = t * t
    f4 = 4 + f5
    gkl = np.zeros((3, nx, ny), order="F")
    I, J = np.meshgrid(np.arange(nx), np.arange(ny), indexing="ij")
    for l in range(3):
        gkl[l] = b_fn(I, J, l) / f5
    db = np.zeros((6, nx + 4, ny + 4), order="F")
    if warm is not None:
        db[:3, 2:-2, 2:-2] = warm
    st, it = orc.cgm(db, gkl, err)
    return st, it, db[:3, 2:-2, 2:-2], f4, f5, gkl * f5


def test_cg_fourier_mode_exact_and_extra_iteration():
    nx, ny = 32, 16
    kx, ky = 2 * math.pi * 3 / nx, 2 * math.pi * 2 / ny
    st, it, phi, f4, f5, b = _cg_case(nx, ny, lambda I, J, l: math.cos(l) * np.cos(kx * I + ky * J))
    assert st == 0
    lam = f5 + 4 * math.sin(kx / 2) ** 2 + 4 * math.sin(ky / 2) ** 2
    assert np.abs(phi - b / lam).max() < 1e-13
    # an eigenvector converges in one step; the stop test on the pre-update
    # residual (field.f90:426 before the loop test :387) costs one more
    assert list(it) == [2, 2, 2]


def test_cg_random_rhs_solves_helmholtz():
    rng = np.random.default_rng(5)
    B = rng.normal(size=(3, 24, 20))
    st, it, phi, f4, f5, b = _cg_case(24, 20, lambda I, J, l: B[l], err=1e-10)
    assert st == 0
    Aphi = f4 * phi - (np.roll(phi, 1, 1) + np.roll(phi, -1, 1) + np.roll(phi, 1, 2) +
                       np.roll(phi, -1, 2))
    for l in range(3):
        assert np.linalg.norm(Aphi[l] - b[l]) <= 1e-10 * np.linalg.norm(b[l]) * 1.01
    assert all(3 <= k <= 20 for k in it)


def test_cg_tiny_rhs_skips_loop():
    """sum_g starts as |b|^2 (field.f90:362) but eps = |b| err (:364): for
    |b| < err the loop never runs and the warm start comes back unchanged."""
    rng = np.random.default_rng(6)
    warm = rng.normal(size=(3, 16, 8))
    st, it, phi, f4, f5, b = _cg_case(16, 8, lambda I, J, l: 1e-9 * np.cos(0.3 * I + l), warm=warm)
    assert st == 0 and list(it) == [0, 0, 0]
    assert np.array_equal(phi, warm)


def test_cg_ite_max_is_fatal():
    st, it, *_ = _cg_case(16, 8, lambda I, J, l: np.cos(0.7 * I * J + l), err=1e-300)
    assert st == OR_ERR_CG_ITEMAX
    assert it[0] == 100


# ------------------------------------------------- field step: Gauss law
def _gauss_residual(p, uf, up, np2):
    ex = uf[3, 2:-2, 2:-2]
    ey = uf[4, 2:-2, 2:-2]
    div = (np.roll(ex, -1, 0) - ex) + (np.roll(ey, -1, 1) - ey)
    return div - 4 * math.pi * _rho(p, up, np2)


def test_whole_steps_keep_gauss_law():
    """Co-located pairs start with rho = 0 and E = 0; the curl terms of dE have
    zero discrete divergence and the deposit is charge conserving, so
    div E - 4 pi rho stays at roundoff (random B exercises the curl stencils)."""
    p, up, np2, cc, uf = state(24, 16, 6, 8, vth=0.3)
    rng = np.random.default_rng(8)
    bz = rng.normal(0, 0.05, (p["nx"], p["ny"]))
    uf[2, 2:-2, 2:-2] = bz
    uf[2, :2, 2:-2], uf[2, -2:, 2:-2] = bz[-2:], bz[:2]
    uf[2, :, :2], uf[2, :, -2:] = uf[2, :, -4:-2], uf[2, :, 2:4]
    orc = Oracle(p)
    gp = up.copy(order="F")
    emax = 0.0
    for _ in range(6):
        orc.step(up, gp, uf, cc, np2)
        g = _gauss_residual(p, uf, up, np2)
        emax = max(emax, np.abs(uf[3:5]).max())
        assert np.abs(g).max() < 1e-12 * max(emax, 1e-3)
    assert emax > 1e-4   # the fields did evolve


# ------------------------------------------------------ boundary maps
def test_bc_dfield_and_phi_are_periodic_copies():
    p = weibel.grid_params(10, 7, 1)
    orc = Oracle(p)
    rng = np.random.default_rng(1)
    df = rng.normal(size=(6, 14, 11)).copy(order="F")
    inner = df[:, 2:-2, 2:-2].copy()
    orc.bc_dfield(df)
    for a in range(14):
        for b in range(11):
            assert np.array_equal(df[:, a, b], inner[:, (a - 2) % 10, (b - 2) % 7])
    phi = rng.normal(size=(12, 9)).copy(order="F")
    inner = phi[1:-1, 1:-1].copy()
    orc.bc_phi(phi)
    for a in range(12):
        for b in range(9):
            assert phi[a, b] == inner[(a - 1) % 10, (b - 1) % 7]


def test_bc_curre_folds_every_ghost_once():
    p = weibel.grid_params(10, 7, 1)
    orc = Oracle(p)
    rng = np.random.default_rng(2)
    uj = rng.integers(-50, 50, size=(3, 14, 11)).astype(np.float64).copy(order="F")
    want = np.zeros((3, 10, 7))
    for a in range(14):
        for b in range(11):
            want[:, (a - 2) % 10, (b - 2) % 7] += uj[:, a, b]
    orc.bc_curre(uj)
    assert np.array_equal(uj[:, 2:-2, 2:-2], want)       # integer-valued: exact
    for a in range(14):
        for b in range(11):
            assert np.array_equal(uj[:, a, b], want[:, (a - 2) % 10, (b - 2) % 7])


# ------------------------------------------------- particle boundaries
def test_round_down_add_is_exact():
    rng = np.random.default_rng(3)
    orc = Oracle(weibel.grid_params(8, 8, 1))
    for _ in range(2000):
        a = float(rng.uniform(-3, 300)) * (1 + float(rng.uniform()) * 1e-3)
        b = float(rng.choice([128.0, -128.0, 2048.0, 1e-17, -3.3]))
        got = orc.add_rd(a, b)
        exact = Fraction(a) + Fraction(b)
        assert Fraction(got) <= exact
        assert Fraction(math.nextafter(got, math.inf)) > exact


def test_particle_x_wrap_never_reaches_upper_bound():
    p = weibel.grid_params(128, 4, 1)
    up, np2, cc, jl = _single(p, 2.5, 3.5, [0, 0, 0])
    xs = [2.0 - 2.0 ** -50, 1.75, 130.0, 130.0 + 2 ** -44, 2.0 - 2.0 ** -52]
    orc = Oracle(p)
    for x in xs:
        up[0, 0, jl, 0] = x
        orc.particle_x(up, np2)
        v = up[0, 0, jl, 0]
        assert 2 <= int(v) <= 129
        exact = Fraction(x) + (128 if x < 2 else -128)
        assert Fraction(v) <= exact < Fraction(math.nextafter(v, math.inf))


def _fill_closed_form(rows_before, leave_dir, ny):
    """Second formulation of the single-rank hole filling (see wm_bc.hip)."""
    out = []
    for j in range(ny):
        ids = rows_before[j]
        holes = [k for k in range(len(ids)) if leave_dir[j][k] != 0]
        up_ = lambda jj: [rows_before[jj][k] for k in range(len(rows_before[jj])) if leave_dir[jj][k] == 1]
        dn_ = lambda jj: [rows_before[jj][k] for k in range(len(rows_before[jj])) if leave_dir[jj][k] == -1]
        below, above = (j - 1) % ny, (j + 1) % ny
        arr = dn_(above) + up_(below) if j == 0 else up_(below) + dn_(above)
        new = list(ids)
        N, H, A = len(ids), len(holes), len(arr)
        for k in range(min(H, A)):
            new[holes[k]] = arr[k]
        if A >= H:
            new += arr[H:]
        else:
            m = H - A
            L = N - m
            rem = [h for h in holes[A:] if h < L]
            tail = [k for k in range(N - 1, L - 1, -1) if k not in holes[A:]]
            for h, t in zip(rem, tail):
                new[h] = new[t]
            new = new[:L]
        out.append(new)
    return out


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_particle_y_invariants_and_closed_form(seed):
    p, up, np2, cc, uf = state(12, 7, 6, seed, vth=0.4)
    rng = np.random.default_rng(seed)
    # move particles up/down one row at random (a few rows get many movers)
    for s in range(2):
        for jl in range(p["ny"]):
            n = np2[jl, s]
            d = rng.choice([-1, 0, 0, 0, 1], size=n) * (rng.uniform(size=n) < (0.9 if jl == 3 else 0.3))
            up[1, :n, jl, s] += d
    before = {s: [list(bits(up[5, :np2[jl, s], jl, s])) for jl in range(p["ny"])] for s in range(2)}
    dirs = {s: [list((up[1, :np2[jl, s], jl, s].astype(int) - (jl + p["nys"]))) for jl in range(p["ny"])]
            for s in range(2)}
    orc = Oracle(p)
    assert orc.particle_y(up, np2) == 0
    for s in range(2):
        after = [list(bits(up[5, :np2[jl, s], jl, s])) for jl in range(p["ny"])]
        assert sorted(sum(after, [])) == sorted(sum(before[s], []))
        for jl in range(p["ny"]):
            y = up[1, :np2[jl, s], jl, s]
            assert np.all(y.astype(int) == jl + p["nys"])
        ny = p["ny"]
        dd = [[(1 if (v == 1 or v == -(ny - 1)) else (-1 if (v == -1 or v == ny - 1) else 0)) for v in row]
              for row in dirs[s]]
        assert after == _fill_closed_form(before[s], dd, ny)


# ------------------------------------------------------------- energy
def test_energy_sums_closed_form():
    p, up, np2, cc, uf = state(8, 6, 3, 9)
    for s in range(2):
        for jl in range(6):
            up[2:5, :np2[jl, s], jl, s] = [[0.3], [0.4], [1.2]]
    uf[...] = 0.0
    uf[3, 2:-2, 2:-2] = 0.1
    uf[2, 2:-2, 2:-2] = -0.2
    e = Oracle(p).energy(up, uf, np2)
    n = np2.sum(axis=0)
    gam = math.sqrt(1 + 0.09 + 0.16 + 1.44)
    assert e[0] == pytest.approx(n[0] * p["r"][0] * (gam - 1), rel=1e-13)
    assert e[2] == pytest.approx(8 * 6 * 0.01 / (8 * math.pi), rel=1e-14)
    assert e[3] == pytest.approx(8 * 6 * 0.04 / (8 * math.pi), rel=1e-14)
