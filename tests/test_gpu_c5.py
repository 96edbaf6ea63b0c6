"""The largest Weibel workload at its full size: the pair plasma on 8192 x 8192 cells with 32
particles per cell and species (4.3e9 particles), in the FP32 form with
flexible CG and the learned preconditioner, on one GPU.  Rows of 1.4 times
their mean occupancy and mover pools of 32768 records per row keep the
device state near 250 GB of the card's 288 (the reference's rows are 5 times
the mean, proj/weibel/app.f90:248).  Two steps after a plain warm-up step
that records the training pairs: the particle count and the IDs' sum and xor
(computed on the device) are conserved exactly, the energies stay finite
and their total within 1e-3."""
import numpy as np
import pytest

import wumingpic2d_amd as wm
from wumingpic2d_amd import weibel
from wmhelpers import make_ctx

pytestmark = pytest.mark.gpu


def test_c5_fp32_learned_pc_fits_and_keeps_its_invariants():
    nx = ny = 8192
    ppc = 32
    p = weibel.grid_params(nx, ny, ppc)
    ctx = make_ctx(p, cap=int(1.4 * ppc * nx), mcap=32768, form=wm.FORM_FP32)
    ctx.init_weibel(5, ppc, 0.05, 0.5)
    n0 = ctx.particle_count()
    assert n0 == 2 * nx * ny * ppc
    ids0 = ctx.id_checksum()
    e0 = ctx.energy()
    ctx.pc_collect(2, 2)
    ctx.step(1)
    loss = ctx.pc_train(20, 3e-3, 5)
    assert np.isfinite(loss).all()
    ctx.pc_collect(0, 0)
    ctx.set_solver(wm.SOLVER_FCG)
    ctx.step(2)
    a, b = ctx.solve_stats()
    print(f"particles {n0}, flexible-CG iterations {a.tolist()}, plain {b.tolist()}, "
          f"capacity {ctx.capacity_stats()}")
    assert ctx.particle_count() == n0
    assert ctx.id_checksum() == ids0
    e = ctx.energy()
    assert np.isfinite(e).all()
    assert abs(e.sum() - e0.sum()) <= 1e-3 * e0.sum()
    ctx.close()
