"""Pins of the oracle's moment diagnostics (mom_calc__accl, mom_calc__nvt,
boundary_periodic__mom).

mom_calc.f90 builds here unchanged (oracle/_ref), so accl and nvt are checked
bit-exactly against the reference's own code, including a grid whose first
cell starts at x = 0, where int() truncation makes the CIC weights
extrapolate.  Independently of that build: accl equals particle__solv's
velocity update at half the step (itself pinned against the reference
build), zero fields leave u unchanged, one particle's four node weights have
a closed form, N sums to the particle count.  boundary_periodic.f90 does not
build here (it uses MPI); the fold is pinned by where every ghost entry of
an integer-valued array must land.
"""
import math

import numpy as np
import pytest

from oracle import Oracle, RefBuild, ref_available
from wumingpic2d_amd import weibel
from wmhelpers import bits, random_fields, state


def mom_shape(p):
    return (7, p["nx"] + 2, p["ny"] + 2, p["nsp"])


def shifted(p, nxgs, nygs):
    """The same grid with its first cell at (nxgs, nygs)."""
    q = dict(p)
    q.update(nxgs=nxgs, nxge=nxgs + p["nx"] - 1, nygs=nygs, nyge=nygs + p["ny"] - 1, nys=nygs,
             nye=nygs + p["ny"] - 1)
    return q


def moved_state(p, up, dx0, dy0):
    up = up.copy(order="F")
    up[0] += dx0
    up[1] += dy0
    return up


@pytest.mark.skipif(not ref_available(), reason="oracle/_ref not built (no reference tree)")
@pytest.mark.parametrize("nx,ny,ppc,seed,amp,org", [(32, 16, 8, 1, 0.05, 2), (20, 7, 13, 2, 0.4, 2),
                                                    (12, 5, 9, 3, 1.0, 0)])
def test_mom_accl_and_nvt_bitexact_vs_reference_build(nx, ny, ppc, seed, amp, org):
    p, up, np2, cc, uf = state(nx, ny, ppc, seed, vth=0.4)
    if org != 2:                     # first cell at x = y = 0: int() truncation below 0.5
        p = shifted(p, org, org)
        up = moved_state(p, up, org - 2, org - 2)
    random_fields(uf, seed, amp)
    orc, ref = Oracle(p), RefBuild(p)
    g1, g2 = np.zeros_like(up), np.zeros_like(up)
    orc.mom_calc_accl(g1, up, uf, cc)
    ref.mom_calc__accl(g2, up, uf, cc)
    assert np.array_equal(bits(g1), bits(g2))
    m1, m2 = np.zeros(mom_shape(p), order="F"), np.zeros(mom_shape(p), order="F")
    orc.mom_calc_nvt(m1, g1, np2)
    ref.mom_calc__nvt(m2, g1, np2)
    assert np.array_equal(bits(m1), bits(m2))
    if org == 0:                     # the truncated cells were exercised
        assert (up[0, :np2[0, 0], :, 0] < 0.5).any()


def test_accl_is_the_push_velocity_at_half_step():
    p, up, np2, cc, uf = state(24, 8, 10, 4, vth=0.5)
    random_fields(uf, 4, 0.3)
    g1 = np.zeros_like(up)
    Oracle(p).mom_calc_accl(g1, up, uf, cc)
    half = dict(p, delt=p["delt"] * 0.5)
    g2 = np.zeros_like(up)
    Oracle(half).particle__solv(g2, up, uf, cc)
    for s in range(p["nsp"]):
        for jl in range(p["ny"]):
            n = np2[jl, s]
            assert np.array_equal(bits(g1[2:5, :n, jl, s]), bits(g2[2:5, :n, jl, s]))
            assert np.array_equal(bits(g1[0:2, :n, jl, s]), bits(up[0:2, :n, jl, s]))
            assert not np.array_equal(g1[0:2, :n, jl, s], g2[0:2, :n, jl, s])   # no move


def test_accl_with_zero_fields_keeps_u():
    p, up, np2, cc, uf = state(16, 6, 7, 5, vth=0.6)
    uf[...] = 0.0
    g = np.zeros_like(up)
    Oracle(p).mom_calc_accl(g, up, uf, cc)
    for s in range(p["nsp"]):
        for jl in range(p["ny"]):
            n = np2[jl, s]
            assert np.array_equal(bits(g[:5, :n, jl, s]), bits(up[:5, :n, jl, s]))


@pytest.mark.parametrize("org,x,y,nodes", [
    # x - 1/2 = 3.75 -> node 3 with weight 1/4, node 4 with 3/4; y - 1/2 = 2.125
    (2, 4.25, 2.625, ((3, 0.25), (4, 0.75), (2, 0.875), (3, 0.125))),
    # first cell at 0: int(-0.375) = 0, so the weights are 1.375 and -0.375
    (0, 0.125, 1.75, ((0, 1.375), (1, -0.375), (1, 0.75), (2, 0.25))),
])
def test_nvt_single_particle_node_weights(org, x, y, nodes):
    p = shifted(weibel.grid_params(6, 4, 1), org, org)
    up = np.zeros((p["ndim"], p["np"], p["ny"], p["nsp"]), order="F")
    np2 = np.zeros((p["ny"], p["nsp"]), dtype=np.int32, order="F")
    u = (0.5, -0.25, 0.75)
    jl = int(y) - org
    up[:5, 0, jl, 1] = (x, y, *u)
    np2[jl, 1] = 1
    mom = np.full(mom_shape(p), np.nan, order="F")
    Oracle(p).mom_calc_nvt(mom, up, np2)
    (i0, wx0), (i1, wx1), (j0, wy0), (j1, wy1) = nodes
    g = 1.0 / math.sqrt(1.0 + sum(v * v for v in u) / p["c"] ** 2)
    want = np.zeros(mom_shape(p))
    for i, wx in ((i0, wx0), (i1, wx1)):
        for j, wy in ((j0, wy0), (j1, wy1)):
            a, b = i - (org - 1), j - (org - 1)
            want[:, a, b, 1] = [wx * wy, u[0] * g * wx * wy, u[1] * g * wx * wy,
                                u[2] * g * wx * wy, u[0] ** 2 * g * wx * wy,
                                u[1] ** 2 * g * wx * wy, u[2] ** 2 * g * wx * wy]
    np.testing.assert_allclose(mom, want, rtol=1e-15, atol=1e-16)
    assert np.array_equal(mom[0], want[0])        # the weights are exact binary fractions


def test_nvt_partition_of_unity_and_momentum_sum():
    p, up, np2, cc, uf = state(20, 9, 11, 6, vth=0.5)
    mom = np.zeros(mom_shape(p), order="F")
    Oracle(p).mom_calc_nvt(mom, up, np2)
    for s in range(p["nsp"]):
        n = int(np2[:, s].sum())
        assert abs(mom[0, ..., s].sum() - n) < 1e-10 * n
        u = np.concatenate([up[2:5, :np2[jl, s], jl, s] for jl in range(p["ny"])], axis=1)
        g = 1.0 / np.sqrt(1.0 + (u * u).sum(axis=0) / p["c"] ** 2)
        for k in range(3):
            np.testing.assert_allclose(mom[1 + k, ..., s].sum(), (u[k] * g).sum(), rtol=1e-11)
            np.testing.assert_allclose(mom[4 + k, ..., s].sum(), (u[k] ** 2 * g).sum(), rtol=1e-11)


@pytest.mark.parametrize("nx,ny", [(10, 7), (1, 1), (3, 1)])
def test_bc_mom_lands_every_ghost_on_its_periodic_image(nx, ny):
    p = weibel.grid_params(nx, ny, 1)
    rng = np.random.default_rng(nx * 31 + ny)
    mom = rng.integers(-50, 50, size=mom_shape(p)).astype(np.float64).copy(order="F")
    want = np.zeros((7, nx, ny, p["nsp"]))
    for a in range(nx + 2):
        for b in range(ny + 2):
            want[:, (a - 1) % nx, (b - 1) % ny] += mom[:, a, b]
    before = mom.copy()
    Oracle(p).bc_mom(mom)
    assert np.array_equal(mom[:, 1:-1, 1:-1], want)       # integer-valued: exact
    # the corners are read, never written
    for a in (0, -1):
        for b in (0, -1):
            assert np.array_equal(mom[:, a, b], before[:, a, b])
