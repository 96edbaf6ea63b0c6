"""Synthetic Weibel-instability states in the reference's array layout.

A counter-streaming pair plasma (the two-beam Weibel set-up): in every row
``ppc*nx`` particles per species at uniform random positions, every species
at the same positions (zero net charge, the pairing of
proj/weibel/app.f90:410-411), momenta ``(0, 0, +/-udrift) + N(0, vth)`` per
component with the drift sign alternating by particle index, IDs
``-(global index + 1)`` bit-cast into ``up(6)`` (proj/weibel/app.f90:462-468),
``E = B = 0``.  Charges follow the normalisation of proj/weibel/app.f90:299-303:
``q = +/- wpe * sqrt(r / (4 pi n0 / delx**2))`` with ``n0 = ppc``.

The draws come from a counter-based generator keyed by (seed, particle
index), the same one ``wm_state_init_weibel`` runs on the device, so host and
device states agree for any launch shape.
"""
from __future__ import annotations

import math

import numpy as np

_M1 = np.uint64(0x9E3779B97F4A7C15)
_M2 = np.uint64(0xBF58476D1CE4E5B9)
_M3 = np.uint64(0x94D049BB133111EB)
_MS = np.uint64(0xD1B54A32D192ED03)
_VS = np.uint64(0x5851F42D4C957F2D)


def _mix64(z: np.ndarray) -> np.ndarray:
    z = z + _M1
    z = (z ^ (z >> np.uint64(30))) * _M2
    z = (z ^ (z >> np.uint64(27))) * _M3
    return z ^ (z >> np.uint64(31))


def uniform(seed: int, ctr: np.ndarray) -> np.ndarray:
    """U[0,1) from (seed, counter): 53 high bits of a mixed 64-bit hash."""
    with np.errstate(over="ignore"):
        h = _mix64((np.uint64(seed) * _MS) ^ _mix64(ctr.astype(np.uint64)))
    return (h >> np.uint64(11)).astype(np.float64) * 2.0 ** -53


def gauss(seed: int, ctr: np.ndarray) -> np.ndarray:
    """N(0,1) by Box-Muller from the counter pair (2*ctr, 2*ctr+1)."""
    c = ctr.astype(np.uint64)
    u1 = 1.0 - uniform(seed, np.uint64(2) * c)
    u2 = uniform(seed, np.uint64(2) * c + np.uint64(1))
    return np.sqrt(-2.0 * np.log(u1)) * np.cos(6.283185307179586 * u2)


def pair_charges(ppc: int, wpe: float, delx: float = 1.0, nsp: int = 2):
    """q, r of an equal-mass pair plasma (proj/weibel/app.f90:299-303)."""
    r = np.ones(nsp)
    q = np.array([(1.0 if s % 2 == 0 else -1.0) * math.sqrt(r[s] / (4 * math.pi * ppc / delx ** 2))
                  * wpe for s in range(nsp)])
    return q, r


def grid_params(nx: int, ny: int, ppc: int, *, np_cap: int | None = None, nsp: int = 2,
                wpe: float = 0.125, delt: float = 0.5, c: float = 1.0, gfac: float = 0.501):
    """Reference-style grid (nxgs = nygs = 2, proj/weibel/app.f90:250-253)."""
    nxgs, nygs = 2, 2
    q, r = pair_charges(ppc, wpe, nsp=nsp)
    return dict(ndim=6, np=np_cap if np_cap else 5 * ppc * nx, nsp=nsp, nxgs=nxgs,
                nxge=nxgs + nx - 1, nygs=nygs, nyge=nygs + ny - 1, nys=nygs, nye=nygs + ny - 1,
                nx=nx, ny=ny, delx=1.0, delt=delt, c=c, gfac=gfac, q=q, r=r, ppc=ppc)


def init_state(p: dict, seed: int, vth: float = 0.05, udrift: float = 0.5):
    """Reference-layout arrays (up, np2, cumcnt, uf), rows sorted by x-cell.

    ``up`` is (ndim, np, ny, nsp) Fortran-ordered; ``cumcnt`` is the
    exclusive prefix of per-cell counts (sort.f90:64-69)."""
    nx, ny, nsp, ppc = p["nx"], p["ny"], p["nsp"], p["ppc"]
    n = ppc * nx
    if n > p["np"]:
        raise ValueError("Too large number of particles")
    ndim, npc = p["ndim"], p["np"]
    up = np.zeros((ndim, npc, ny, nsp), order="F")
    np2 = np.full((ny, nsp), n, dtype=np.int32, order="F")
    cumcnt = np.zeros((nx + 1, ny, nsp), dtype=np.int32, order="F")
    xhi = np.nextafter(float(p["nxge"] + 1), 0.0)
    k = np.arange(n, dtype=np.uint64)
    vs = int(np.uint64(seed) ^ _VS)
    for jl in range(ny):
        j = p["nys"] + jl
        pidx = np.uint64(jl) * np.uint64(n) + k
        x = p["nxgs"] + nx * uniform(seed, np.uint64(4) * pidx)
        y = j + uniform(seed, np.uint64(4) * pidx + np.uint64(1))
        x = np.minimum(x, xhi)
        y = np.minimum(y, np.nextafter(float(j + 1), 0.0))
        for s in range(nsp):
            gidx = (np.uint64(s) * np.uint64(ny) + np.uint64(jl)) * np.uint64(n) + k
            ux = vth * gauss(vs, np.uint64(3) * gidx)
            uy = vth * gauss(vs, np.uint64(3) * gidx + np.uint64(1))
            uz = np.where(k % np.uint64(2) == 1, -udrift, udrift) + \
                vth * gauss(vs, np.uint64(3) * gidx + np.uint64(2))
            ids = -(gidx.astype(np.int64) + 1)
            cell = x.astype(np.int64) - p["nxgs"]
            order = np.argsort(cell, kind="stable")
            rec = np.stack([x, y, ux, uy, uz, ids.view(np.float64)])[:ndim, order]
            up[:, :n, jl, s] = rec
            cnt = np.bincount(cell, minlength=nx)
            cumcnt[1:, jl, s] = np.cumsum(cnt)
    uf = np.zeros((6, nx + 4, ny + 4), order="F")
    return up, np2, cumcnt, uf
