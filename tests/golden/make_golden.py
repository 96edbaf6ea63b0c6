"""Writes the golden vectors of the reference's particle__solv and sort__bucket.

Inputs are seeded synthetic states (wumingpic2d_amd.weibel) with random
fields; outputs come from the reference's own common/particle.f90 and
common/sort.f90, compiled unchanged into oracle/_ref/libwmref.so by
oracle/build_ref.sh.  Run from the repository root:

    python tests/golden/make_golden.py

The fixture holds data only (arrays of inputs and expected outputs).
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

from oracle import RefBuild, ref_available  # noqa: E402
from wumingpic2d_amd import weibel  # noqa: E402

CASES = [
    # name, nx, ny, ppc, seed, field amplitude, thermal speed
    ("a", 16, 8, 4, 101, 0.2, 0.3),
    ("b", 12, 5, 7, 202, 0.5, 0.6),
]


def main() -> None:
    if not ref_available():
        raise SystemExit("oracle/_ref/libwmref.so is not built (run oracle/build_ref.sh)")
    out = {}
    for name, nx, ny, ppc, seed, amp, vth in CASES:
        p = weibel.grid_params(nx, ny, ppc)
        up, np2, cc, uf = weibel.init_state(p, seed, vth=vth)
        rng = np.random.default_rng(seed)
        uf[...] = rng.normal(0.0, amp, uf.shape)
        ref = RefBuild(p)
        gp = np.zeros_like(up)
        ref.particle__solv(gp, up, uf, cc)
        # sort input: each row scrambled, some rows shortened
        us = up.copy(order="F")
        n2 = np2.copy(order="F")
        for s in range(p["nsp"]):
            for jl in range(ny):
                n = int(rng.integers(0, np2[jl, s] + 1)) if jl % 2 else int(np2[jl, s])
                perm = rng.permutation(np2[jl, s])[:n]
                us[:, :n, jl, s] = up[:, perm, jl, s]
                n2[jl, s] = n
        gs = np.zeros_like(up)
        cs = np.zeros_like(cc)
        ref.sort__bucket(gs, us, cs, n2)
        for k, v in dict(up=up, uf=uf, cumcnt=cc, np2=np2, push_gp=gp, sort_in=us, sort_np2=n2,
                         sort_gp=gs, sort_cumcnt=cs).items():
            out[f"{name}_{k}"] = v
        out[f"{name}_grid"] = np.array([nx, ny, ppc, p["np"]], dtype=np.int64)
    np.savez_compressed(os.path.join(HERE, "ref_particle_sort.npz"), **out)
    print("wrote", os.path.join(HERE, "ref_particle_sort.npz"))


if __name__ == "__main__":
    main()
