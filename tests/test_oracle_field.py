"""Pins of the oracle's implicit field solve and of the moving-particle Jz
weights, from what the mathematics fixes (no reference build of field.f90
exists here: it needs an MPI module this image does not have).

The scheme.  The push's cell-centring (particle.f90:68-81, pinned bit-exact
against the reference build) fixes the staggering, in cells:
Bx(i+1/2, j), By(i, j+1/2), Bz(i, j), Ex(i, j+1/2), Ey(i+1/2, j),
Ez(i+1/2, j+1/2), each current component with its E component.  Maxwell's
equations in Gaussian units on that staggering, time-centred with the weight
gfac on the new fields (field.f90:22-63 f1 = c dt/dx), are

    dB = -f1 curl-(E + gfac dE)
    dE = +f1 curl+(B + gfac dB) - 4 pi dt J

with the Yee differences D+ f(i) = f(i+1) - f(i), D- f(i) = f(i) - f(i-1):

    curl-(E) = ( D-y Ez, -D-x Ez, D-x Ey - D-y Ex )   (at the B positions)
    curl+(B) = ( D+y Bz, -D+x Bz, D+x By - D+y Bx )   (at the E positions)

For one Fourier mode exp(i(kx i + ky j)) the differences become the
symbols D+ -> e^{ik} - 1 and D- -> 1 - e^{-ik}; the scheme is then a 6x6
complex linear system per mode, solved here with numpy.  The oracle solves
it differently: it eliminates dE, uses div- B = 0 to turn curl curl into
the 5-point Laplacian, and runs CG on (f5 - Lap) dB = f5 gkl.  Agreement
therefore pins the gkl stencils (field.f90:125-146), the CG solve and the
dE stencils (:154-171), every sign and coefficient, B, E and J terms alike.

Closed form of the vacuum modes.  With J = 0 a TE mode (Bz, Ex, Ey) or TM
mode (Ez, Bx, By) is an eigenvector with amplification

    lambda = (1 + i (1 - g) w) / (1 - i g w),
    w = +/- 2 f1 sqrt(sin^2(kx/2) + sin^2(ky/2)),

from (lambda - 1)^2 = -w^2 (1 + g (lambda - 1))^2; |lambda| = 1 exactly at
g = 1/2 and the phase per step tends to c dt |k| / dx as k -> 0.

Moving-particle Jz (field.f90:292-298): the Esirkepov weight of Jz is the
time integral over the step of the product of the x and y shapes, each
interpolated linearly in time between its old and new TSC values:
W = int_0^1 (S0x + t dSx)(S0y + t dSy) dt.  Two-point Gauss-Legendre
quadrature is exact for this quadratic in t; the shapes come from the TSC
function itself, evaluated at the old and the new position (also across a
cell boundary), not from the deposit's index shifts.
"""
import math

import numpy as np
import pytest

from oracle import Oracle
from wumingpic2d_amd import weibel
from wmhelpers import state

PI = math.pi


def _grid_index(p):
    """Reference indices (i, j) of the ghost-inclusive (nx+4, ny+4) arrays."""
    I = np.arange(p["nx"] + 4) + p["nxgs"] - 2
    J = np.arange(p["ny"] + 4) + p["nys"] - 2
    return np.meshgrid(I, J, indexing="ij")


def _f1(p):
    return p["c"] * p["delt"] / p["delx"]


def _curl_minus(bx, by):
    """curl-(v) as a 3x3 matrix acting on (vx, vy, vz) amplitudes."""
    return np.array([[0, 0, by], [0, 0, -bx], [-by, bx, 0]], dtype=complex)


def _curl_plus(ax, ay):
    return np.array([[0, 0, ay], [0, 0, -ax], [-ay, ax, 0]], dtype=complex)


def fourier_step(p, kx, ky, B, E, J):
    """One step of the implicit scheme for complex mode amplitudes B, E, J."""
    f1, g, dt = _f1(p), p["gfac"], p["delt"]
    ax, ay = np.exp(1j * kx) - 1, np.exp(1j * ky) - 1
    bx, by = 1 - np.exp(-1j * kx), 1 - np.exp(-1j * ky)
    Cm, Cp = _curl_minus(bx, by), _curl_plus(ax, ay)
    A = np.block([[np.eye(3), g * f1 * Cm], [-g * f1 * Cp, np.eye(3)]])
    rhs = np.concatenate([-f1 * Cm @ E, f1 * Cp @ B - 4 * PI * dt * J])
    d = np.linalg.solve(A, rhs)
    return B + d[:3], E + d[3:]


def _real(amp, theta):
    return np.real(amp * np.exp(1j * theta))


def _fields(p, theta, B, E):
    uf = np.zeros((6, p["nx"] + 4, p["ny"] + 4), order="F")
    for c in range(3):
        uf[c] = _real(B[c], theta)
        uf[3 + c] = _real(E[c], theta)
    return uf


def _current(p, theta, J):
    uj = np.zeros((3, p["nx"] + 4, p["ny"] + 4), order="F")
    for c in range(3):
        uj[c] = _real(J[c], theta)
    return uj


def _mode(p, m, n):
    kx, ky = 2 * PI * m / p["nx"], 2 * PI * n / p["ny"]
    I, J = _grid_index(p)
    return kx, ky, kx * I + ky * J


def check_single_mode(orc_factory, seed, m, n, gfac, with_fields=True, with_current=True):
    """The oracle's field step on one mode against fourier_step; returns the
    largest deviation relative to the field scale."""
    p = weibel.grid_params(16, 12, 1, gfac=gfac)
    kx, ky, theta = _mode(p, m, n)
    rng = np.random.default_rng(seed)
    cn = lambda k: rng.normal(size=k) + 1j * rng.normal(size=k)
    bx, by = 1 - np.exp(-1j * kx), 1 - np.exp(-1j * ky)
    alpha = cn(1)[0]
    B = np.array([by * alpha, -bx * alpha, cn(1)[0]]) if with_fields else np.zeros(3, complex)
    E = cn(3) if with_fields else np.zeros(3, complex)
    J = 0.05 * cn(3) if with_current else np.zeros(3, complex)
    uf = _fields(p, theta, B, E)
    uj = _current(p, theta, J)
    orc = orc_factory(p)
    orc.set_cg_err(1e-14)
    orc.field_advance(uf, uj)
    B1, E1 = fourier_step(p, kx, ky, B, E, J)
    want = _fields(p, theta, B1, E1)
    scale = np.abs(want).max()
    return np.abs(uf - want).max() / scale


@pytest.mark.parametrize("m,n,gfac,seed", [(1, 1, 0.501, 1), (3, -2, 0.501, 2), (5, 4, 0.5, 3),
                                           (0, 3, 0.6, 4), (7, 0, 0.501, 5), (8, 6, 0.55, 6)])
def test_single_mode_matches_implicit_maxwell(m, n, gfac, seed):
    assert check_single_mode(Oracle, seed, m, n, gfac) < 1e-11


@pytest.mark.parametrize("m,n", [(2, 1), (-1, 3)])
def test_single_mode_current_alone(m, n):
    """Zero fields, a single-mode J: only the f3 J terms of gkl and the
    4 pi dt J terms of dE act."""
    assert check_single_mode(Oracle, 11 + m, m, n, 0.501, with_fields=False) < 1e-11


def amplification(p, kx, ky, branch):
    f1, g = _f1(p), p["gfac"]
    w = branch * 2 * f1 * math.sqrt(math.sin(kx / 2) ** 2 + math.sin(ky / 2) ** 2)
    return (1 + 1j * (1 - g) * w) / (1 - 1j * g * w)


def eigenmode(p, kx, ky, lam, kind):
    """Eigenvector (B, E) of the vacuum scheme for amplification lam."""
    f1, g = _f1(p), p["gfac"]
    mu = 1 + g * (lam - 1)
    ax, ay = np.exp(1j * kx) - 1, np.exp(1j * ky) - 1
    bx, by = 1 - np.exp(-1j * kx), 1 - np.exp(-1j * ky)
    if kind == "TE":       # Bz = 1; (lam-1) Ex = f1 mu D+y Bz, (lam-1) Ey = -f1 mu D+x Bz
        B = np.array([0, 0, 1], dtype=complex)
        E = np.array([f1 * mu * ay / (lam - 1), -f1 * mu * ax / (lam - 1), 0])
    else:                  # Ez = 1; (lam-1) Bx = -f1 mu D-y Ez, (lam-1) By = f1 mu D-x Ez
        E = np.array([0, 0, 1], dtype=complex)
        B = np.array([-f1 * mu * by / (lam - 1), f1 * mu * bx / (lam - 1), 0])
    return B, E


def check_vacuum_mode(orc_factory, kind, m, n, gfac, branch, steps=3):
    p = weibel.grid_params(16, 12, 1, gfac=gfac)
    kx, ky, theta = _mode(p, m, n)
    lam = amplification(p, kx, ky, branch)
    B, E = eigenmode(p, kx, ky, lam, kind)
    uf = _fields(p, theta, B, E)
    uj = np.zeros((3, p["nx"] + 4, p["ny"] + 4), order="F")
    orc = orc_factory(p)
    orc.set_cg_err(1e-14)
    dev = 0.0
    for s in range(1, steps + 1):
        orc.field_advance(uf, uj)
        want = _fields(p, theta, lam ** s * B, lam ** s * E)
        dev = max(dev, np.abs(uf - want).max() / np.abs(want).max())
    return dev


@pytest.mark.parametrize("kind", ["TE", "TM"])
@pytest.mark.parametrize("m,n,gfac,branch", [(1, 2, 0.5, 1), (3, -1, 0.501, -1), (6, 5, 0.6, 1),
                                             (2, 0, 0.75, -1)])
def test_vacuum_eigenmodes_closed_form(kind, m, n, gfac, branch):
    assert check_vacuum_mode(Oracle, kind, m, n, gfac, branch) < 1e-11


@pytest.mark.parametrize("kind", ["TE", "TM"])
def test_vacuum_wave_speed_and_energy(kind):
    """Measured from the oracle alone: at gfac = 1/2 a long wave keeps its
    amplitude (|lambda| = 1 to roundoff) and advances by c dt |k| per step
    (phase speed c, within the scheme's O(k^2) dispersion); at gfac > 1/2 it
    is damped."""
    for gfac, nx in ((0.5, 64), (0.6, 64)):
        p = weibel.grid_params(nx, 8, 1, gfac=gfac)
        kx, ky, theta = _mode(p, 1, 0)
        lam = amplification(p, kx, ky, 1)
        B, E = eigenmode(p, kx, ky, lam, kind)
        uf = _fields(p, theta, B, E)
        orc = Oracle(p)
        orc.set_cg_err(1e-14)
        orc.field_advance(uf, np.zeros((3, nx + 4, 12), order="F"))
        # complex amplitude of the mode in the output over the input's (both 1):
        # Bz of a TE mode, Ez of a TM mode
        out = uf[2 if kind == "TE" else 5, 2:-2, 2:-2]
        lam_meas = 2 * (out * np.exp(-1j * theta[2:-2, 2:-2])).mean()
        if gfac == 0.5:
            assert abs(abs(lam_meas) - 1) < 1e-12
            speed = np.angle(lam_meas) / (p["c"] * p["delt"] * abs(kx) / p["delx"])
            assert abs(speed - 1) < 2e-3
        else:
            assert abs(lam_meas) < 1 - 1e-6


def check_uniform_current(orc_factory):
    p = weibel.grid_params(10, 8, 1)
    uf = np.zeros((6, 14, 12), order="F")
    uj = np.zeros((3, 14, 12), order="F")
    J0 = (0.013, -0.021, 0.034)
    for c in range(3):
        uj[c] = J0[c]
    orc = orc_factory(p)
    orc.field_advance(uf, uj)
    want = [-4 * PI * p["delt"] * J0[c] for c in range(3)]
    dev = np.abs(uf[:3]).max()
    for c in range(3):
        dev = max(dev, np.abs(uf[3 + c] - want[c]).max() / abs(want[c]))
    return dev


def test_uniform_current_gives_dE_and_no_dB():
    assert check_uniform_current(Oracle) < 1e-15


def _div_b(uf):
    """div- B = D-x Bx + D-y By at the Bz nodes, interior."""
    bx, by = uf[0], uf[1]
    return (bx[2:-2, 2:-2] - bx[1:-3, 2:-2]) + (by[2:-2, 2:-2] - by[2:-2, 1:-3])


def check_div_b(orc_factory, steps=20):
    """Whole steps with particles from a divergence-free B: div- B stays at
    roundoff.  The CG tolerance is tightened: the x and y solves are
    independent, so at the reference's 1e-6 div B would drift at 1e-6."""
    p, up, np2, cc, uf = state(16, 12, 6, 21, vth=0.3)
    rng = np.random.default_rng(21)
    nx, ny = p["nx"], p["ny"]
    A = rng.normal(0, 0.05, (nx, ny))
    bz = rng.normal(0, 0.05, (nx, ny))
    per = lambda a: np.pad(a, 2, mode="wrap")
    Ag = per(A)
    bx = Ag - np.roll(Ag, 1, 1)          # D-y A
    by = -(Ag - np.roll(Ag, 1, 0))       # -D-x A
    uf[0] = per(bx[2:-2, 2:-2])
    uf[1] = per(by[2:-2, 2:-2])
    uf[2] = per(bz)
    uf[3:6] = np.stack([per(rng.normal(0, 0.02, (nx, ny))) for _ in range(3)])
    orc = orc_factory(p)
    orc.set_cg_err(1e-14)
    gp = up.copy(order="F")
    bmax = np.abs(uf[:2]).max()
    assert np.abs(_div_b(uf)).max() < 1e-16
    worst = 0.0
    b0 = uf[:3].copy()
    for _ in range(steps):
        orc.step(up, gp, uf, cc, np2)
        worst = max(worst, np.abs(_div_b(uf)).max() / bmax)
    assert np.abs(uf[:3] - b0).max() > 1e-3 * bmax   # B did evolve
    return worst


def test_div_b_stays_at_roundoff():
    assert check_div_b(Oracle) < 1e-13


def _tsc(d):
    """Quadratic B-spline (TSC) weight of a node at distance d (cells)."""
    a = np.abs(d)
    return np.where(a <= 0.5, 0.75 - a * a, np.where(a <= 1.5, 0.5 * (1.5 - a) ** 2, 0.0))


JZ_MOVES = [(6.31, 7.82, 0.23, -0.17), (6.9, 7.1, 0.35, 0.0), (6.1, 7.95, -0.4, 0.3),
            (6.95, 7.05, 0.2, -0.22), (6.5, 7.5, -0.49, -0.49), (6.02, 7.61, 0.0, 0.6)]


def check_moving_jz(orc_factory, x0, y0, dx, dy):
    p = weibel.grid_params(12, 12, 1)
    up = np.zeros((6, p["np"], p["ny"], p["nsp"]), order="F")
    np2 = np.zeros((p["ny"], p["nsp"]), dtype=np.int32, order="F")
    cc = np.zeros((p["nx"] + 1, p["ny"], p["nsp"]), dtype=np.int32, order="F")
    jl, il = int(y0) - p["nys"], int(x0) - p["nxgs"]
    u = np.array([0.3, -0.2, 0.5])
    up[:5, 0, jl, 0] = [x0, y0, *u]
    np2[jl, 0] = 1
    cc[il + 1:, jl, 0] = 1
    gp = up.copy(order="F")
    gp[0, 0, jl, 0], gp[1, 0, jl, 0] = x0 + dx, y0 + dy
    uj = np.zeros((3, 16, 16), order="F")
    orc_factory(p).ele_cur(uj, up, gp, cc)
    i, j = int(x0), int(y0)
    a = np.arange(-2, 3)
    s0x, s1x = _tsc(x0 - 0.5 - (i + a)), _tsc(x0 + dx - 0.5 - (i + a))
    s0y, s1y = _tsc(y0 - 0.5 - (j + a)), _tsc(y0 + dy - 0.5 - (j + a))
    W = np.zeros((5, 5))
    for t in (0.5 - 0.5 / math.sqrt(3), 0.5 + 0.5 / math.sqrt(3)):
        W += 0.5 * np.outer(s0x + t * (s1x - s0x), s0y + t * (s1y - s0y))
    gvz = u[2] / math.sqrt(1 + u @ u)
    got = uj[2, i - 2 - p["nxgs"] + 2:i + 3 - p["nxgs"] + 2, j - 2 - p["nys"] + 2:j + 3 - p["nys"] + 2]
    want = p["q"][0] * gvz * W
    assert abs(uj[2].sum() - want.sum()) < 1e-15 * abs(want.sum()) + 1e-18
    return np.abs(got - want).max() / np.abs(want).max()


@pytest.mark.parametrize("x0,y0,dx,dy", JZ_MOVES)
def test_moving_particle_jz_is_time_integral_of_shapes(x0, y0, dx, dy):
    assert check_moving_jz(Oracle, x0, y0, dx, dy) < 1e-14
