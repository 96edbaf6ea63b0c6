"""Learned preconditioner of the implicit solve on the GPU: the BF16 MFMA
net's forward pass and its hand-written backward pass against the oracle
(oracle/pc_oracle.py), training on pairs recorded by the plain solver,
the iteration gain of flexible CG over plain CG at the same tolerance with
the same solution, and the plain-solver fallback."""
import numpy as np
import pytest

import pc_oracle as P
import wumingpic2d_amd as wm
from wmhelpers import make_ctx, state

pytestmark = pytest.mark.gpu


def _f4(p):
    return 4 + (p["delx"] / (p["c"] * p["delt"] * p["gfac"])) ** 2


def _random_theta(seed):
    """Weights of unit gain per layer, so the coarse (MFMA) path carries as
    much of the output as the fine one."""
    rng = np.random.default_rng(seed)
    th = np.zeros(P.NPARAM, np.float32)
    for name, (o, ci, co) in P.OFFSETS.items():
        th[o:o + 9 * ci * co] = rng.standard_normal(9 * ci * co) / np.sqrt(9 * ci)
    return th


def _ctx(nx, ny, seed=50):
    p, up, np2, cc, uf = state(nx, ny, 4, seed, vth=0.1)
    ctx = make_ctx(p)
    ctx.state_upload(up, np2, cc, uf)
    return p, ctx


@pytest.mark.parametrize("nx,ny", [(64, 32), (48, 40)])
def test_forward_matches_oracle(nx, ny):
    p, ctx = _ctx(nx, ny)
    th = _random_theta(1)
    ctx.pc_set_params(th)
    r = np.random.default_rng(2).standard_normal((3, ny, nx)) * 3e-4
    z = ctx.pc_apply(r)
    zo, c = P.forward(th, r)
    err = np.linalg.norm(z - zo) / np.linalg.norm(zo)
    assert err < 2e-3, err
    # the coarse levels matter: dropping them changes the output a lot
    th0 = th.copy()
    o = P.OFFSETS["u1"]
    th0[o[0]:o[0] + 9 * o[1] * o[2]] = 0
    z0, _ = P.forward(th0, r)
    assert np.linalg.norm(z0 - zo) > 0.2 * np.linalg.norm(zo)


def test_gradient_matches_oracle():
    p, ctx = _ctx(64, 32)
    th = _random_theta(3)
    ctx.pc_set_params(th)
    rng = np.random.default_rng(4)
    r = rng.standard_normal((3, 32, 64))
    e = rng.standard_normal((3, 32, 64))
    L, g = ctx.pc_grad(r, e)
    Lo, go = P.loss_grad(th, r, e)
    assert abs(L - Lo) <= 1e-3 * Lo, (L, Lo)
    for name, (o, ci, co) in P.OFFSETS.items():
        a, b = g[o:o + 9 * ci * co], go[o:o + 9 * ci * co]
        assert np.linalg.norm(a - b) <= 1e-2 * np.linalg.norm(b), name


def _pair_ctxs(nx, ny, seed, tol):
    p, up, np2, cc, uf = state(nx, ny, 8, seed, vth=0.2)
    out = []
    for _ in range(2):
        c = make_ctx(p)
        c.set_cg_tolerance(tol)
        c.state_upload(up, np2, cc, uf)
        out.append(c)
    return p, out


def test_training_gains_iterations_same_solution():
    p, (a, b) = _pair_ctxs(64, 64, 52, 1e-10)
    a.pc_collect(16, 4)
    a.step(4)
    b.step(4)
    assert a.pc_pair_count() == 16
    loss = a.pc_train(300, 3e-3, 7)
    assert loss[1] < 0.02 * loss[0], loss
    a.set_solver(wm.SOLVER_FCG)
    for _ in range(3):
        a.step(1)
        b.step(1)
        fcg, plain = a.solve_stats()
        _, ref = b.solve_stats()
        assert np.all(2 * (fcg + plain) <= ref + 1), (fcg, plain, ref)
    ua, ub = np.zeros((6, 68, 68), order="F"), np.zeros((6, 68, 68), order="F")
    a.state_download(uf=ua)
    b.state_download(uf=ub)
    assert np.abs(ua - ub).max() <= 1e-7 * np.abs(ub).max()


def test_unusable_preconditioner_falls_back_to_plain():
    p, (a, b) = _pair_ctxs(32, 32, 53, 1e-6)
    a.pc_set_params(np.zeros(P.NPARAM, np.float32))
    a.set_solver(wm.SOLVER_FCG)
    a.step(2)
    b.step(2)
    fcg, plain = a.solve_stats()
    assert np.all(fcg == 0) and np.array_equal(plain, b.cg_iters())
    ua, ub = np.zeros((6, 36, 36), order="F"), np.zeros((6, 36, 36), order="F")
    a.state_download(uf=ua)
    b.state_download(uf=ub)
    assert np.array_equal(ua, ub)


def test_training_is_deterministic():
    p, (a, b) = _pair_ctxs(32, 32, 54, 1e-8)
    for c in (a, b):
        c.pc_collect(8, 3)
        c.step(3)
        c.pc_train(40, 3e-3, 11)
    assert np.array_equal(a.pc_params().view(np.int32), b.pc_params().view(np.int32))


def test_solver_needs_parameters():
    p, ctx = _ctx(32, 16)
    with pytest.raises(wm.WmError) as e:
        ctx.set_solver(wm.SOLVER_FCG)
    assert e.value.code == wm.WM_ERR_ARG
    with pytest.raises(wm.WmError):
        ctx.pc_train(10)
