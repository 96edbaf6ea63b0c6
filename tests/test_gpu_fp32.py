"""FP32 particle form (int32 cell + FP32 offsets and momenta) against the FP64
oracle.  Error bars are the FP32 ones: per-particle state after a few steps
within a few 1e-6 of the FP64 trajectory, fields within 1e-4 of the field
scale, energies within 1e-5.  Row counts, the particle order inside rows and
the cell counts can differ only where an FP32 position lands on the other
side of a cell or row boundary, so particles are matched by ID."""
import numpy as np
import pytest

import wumingpic2d_amd as wm
from oracle import Oracle
from wmhelpers import make_ctx, state

pytestmark = pytest.mark.gpu


def _ctx32(p, **kw):
    ctx = wm.Context(0)
    ctx.set_form(wm.FORM_FP32)
    g = (p["ndim"], p["np"], p["nsp"], p["nxgs"], p["nxge"], p["nygs"], p["nyge"], p["nys"], p["nye"])
    ctx.boundary_periodic__init(*g, 0, 0, 0, 0, 0, 0, [0], p["delx"], p["delt"], p["c"])
    ctx.particle__init(*g, p["delx"], p["delt"], p["c"], p["q"], p["r"])
    ctx.field__init(*g, 0, 0, 0, 0, p["delx"], p["delt"], p["c"], p["q"], p["r"], p["gfac"])
    ctx.sort__init(*g)
    return ctx


def _by_id(up, np2):
    rec = {}
    for s in range(up.shape[3]):
        for jl in range(up.shape[2]):
            n = np2[jl, s]
            ids = up[5, :n, jl, s].view(np.int64)
            for k in range(n):
                rec[int(ids[k])] = up[:5, k, jl, s]
    return rec


def test_fp32_initialiser_matches_host_state():
    p, up, np2, cc, uf = state(32, 8, 8, 41, vth=0.05, udrift=0.5)
    ctx = _ctx32(p)
    ctx.init_weibel(41, 8, 0.05, 0.5)
    gup, gnp2, gcc = np.zeros_like(up), np.zeros_like(np2), np.zeros_like(cc)
    ctx.state_download(up=gup, np2=gnp2, cumcnt=gcc)
    assert np.array_equal(gnp2, np2) and np.array_equal(gcc, cc)
    for s in range(2):
        for jl in range(p["ny"]):
            n = np2[jl, s]
            a, b = up[:, :n, jl, s], gup[:, :n, jl, s]
            assert np.array_equal(a[5].view(np.int64), b[5].view(np.int64))
            ix = a[0].astype(np.int64)
            ox = np.minimum(np.float32(a[0] - ix), np.float32(np.nextafter(np.float32(1), np.float32(0))))
            assert np.array_equal(b[0], ix + ox.astype(np.float64))
            assert np.allclose(b[2:5], a[2:5], rtol=2e-7, atol=1e-9)


@pytest.mark.parametrize("nx,ny,ppc,seed,steps,nsp,vth", [(32, 16, 16, 42, 6, 2, 0.1),
                                                         (48, 12, 8, 43, 4, 2, 0.1),
                                                         (8, 4, 96, 44, 3, 1, 0.1),
                                                         (16, 16, 8, 45, 4, 3, 0.1),
                                                         (16, 8, 64, 46, 3, 2, 0.4),
                                                         (24, 6, 16, 47, 4, 2, 0.2)])
def test_fp32_steps_track_fp64_oracle(nx, ny, ppc, seed, steps, nsp, vth):
    """Two species run the push's straight-line pair path; one and three
    species run its species loop.  96 particles per cell put bins above the
    80 the push holds in registers, so its direct-load tail runs too (and
    an 8-cell row is narrower than one workgroup's 16 cells).  The hot plasma
    (u_th = 0.4 c) moves about a quarter of the particles per step, about 15
    per bin: bins with more movers than their summary lists, several arrivals
    from one direction, and holes past the listed ones all occur, so the
    migration's record-reading paths run.  The 6-row grid adds a row count
    that is neither a power of two nor a multiple of four."""
    p, up, np2, cc, uf = state(nx, ny, ppc, seed, vth=vth, nsp=nsp)
    ctx = _ctx32(p)
    ctx.state_upload(up, np2, cc, uf)
    orc = Oracle(p)
    gp = up.copy(order="F")
    for _ in range(steps):
        orc.step(up, gp, uf, cc, np2)
        ctx.step(1)
    gup, gnp2, guf = np.zeros_like(up), np.zeros_like(np2), np.zeros_like(uf)
    ctx.state_download(up=gup, np2=gnp2, uf=guf)
    assert gnp2.sum() == np2.sum()
    a, b = _by_id(up, np2), _by_id(gup, gnp2)
    assert a.keys() == b.keys()
    dpos = max(np.abs(a[k][:2] - b[k][:2]).max() for k in a)
    du = max(np.abs(a[k][2:] - b[k][2:]).max() for k in a)
    assert dpos < 2e-5, dpos
    assert du < 2e-5, du
    scale = np.abs(uf).max()
    assert np.abs(guf - uf).max() < 1e-3 * scale
    e1, e2 = orc.energy(up, uf, np2), ctx.energy()
    assert np.allclose(e1[:2], e2[:2], rtol=1e-5)
    assert np.allclose(e1[2:], e2[2:], rtol=1e-2, atol=1e-9 * e1[:2].sum())


def test_fp32_run_is_deterministic():
    p, up, np2, cc, uf = state(32, 16, 8, 44, vth=0.2)
    outs = []
    for _ in range(2):
        ctx = _ctx32(p)
        ctx.state_upload(up, np2, cc, uf)
        ctx.step(5)
        o, n, f = np.zeros_like(up), np.zeros_like(np2), np.zeros_like(uf)
        ctx.state_download(up=o, np2=n, uf=f)
        outs.append((o, n, f))
        ctx.close()
    assert np.array_equal(outs[0][1], outs[1][1])
    assert np.array_equal(outs[0][2], outs[1][2])
    for s in range(2):
        for jl in range(p["ny"]):
            k = outs[0][1][jl, s]
            assert np.array_equal(outs[0][0][:, :k, jl, s].view(np.int64),
                                  outs[1][0][:, :k, jl, s].view(np.int64))


def test_fp32_settling_between_steps_changes_nothing():
    """A push leaves its movers pending; the next push places them in its
    prologue, or the settle kernel does when the bins are read first (energy
    sums, download).  Both use the same code, so energy sums between steps
    leave the trajectory bit-identical."""
    p, up, np2, cc, uf = state(32, 12, 16, 48, vth=0.3)
    outs = []
    for settle_between in (False, True):
        ctx = _ctx32(p)
        ctx.state_upload(up, np2, cc, uf)
        for _ in range(4):
            ctx.step(1)
            if settle_between:
                ctx.energy()
        o, n, f = np.zeros_like(up), np.zeros_like(np2), np.zeros_like(uf)
        ctx.state_download(up=o, np2=n, uf=f)
        outs.append((o, n, f))
        ctx.close()
    assert np.array_equal(outs[0][1], outs[1][1])
    assert np.array_equal(outs[0][2], outs[1][2])
    for s in range(2):
        for jl in range(p["ny"]):
            k = outs[0][1][jl, s]
            assert np.array_equal(outs[0][0][:, :k, jl, s].view(np.int64),
                                  outs[1][0][:, :k, jl, s].view(np.int64))


def _one_row_state(p, rows):
    """Particles of species 0 only: rows = {jl: [(x, y, ux, uy), ...]}."""
    ndim, npc, ny, nsp = p["ndim"], p["np"], p["ny"], p["nsp"]
    up = np.zeros((ndim, npc, ny, nsp), order="F")
    np2 = np.zeros((ny, nsp), dtype=np.int32, order="F")
    cc = np.zeros((p["nx"] + 1, ny, nsp), dtype=np.int32, order="F")
    gid = 0
    for jl, parts in rows.items():
        parts = sorted(parts, key=lambda t: int(t[0]))
        n = len(parts)
        a = np.array(parts, dtype=np.float64).reshape(n, 4)
        up[0, :n, jl, 0], up[1, :n, jl, 0], up[2, :n, jl, 0], up[3, :n, jl, 0] = a.T
        up[5, :n, jl, 0] = (-(np.arange(gid, gid + n, dtype=np.int64) + 1)).view(np.float64)
        gid += n
        np2[jl, 0] = n
        cnt = np.bincount((a[:, 0] - p["nxgs"]).astype(np.int64), minlength=p["nx"])
        cc[1:, jl, 0] = np.cumsum(cnt)
    return up, np2, cc, np.zeros((6, p["nx"] + 4, ny + 4), order="F")


def test_fp32_full_bin_is_laid_out_anew():
    """Row segments of 176 slots leave two 8-slot blocks of room beyond the
    bins' needs (16 cells of at least 8 slots; 12 particles in cell 4 get 16,
    30 in cell 5 get 32).  Shared out by the running sum of the needs, one of
    them goes to cell 5: 40 slots.  Twelve fast particles of cell 4 arrive in
    cell 5: the fill check after the push sees 42 > 40.  Bins move whole,
    their movers' holes with them, so cell 4 keeps its 12 slots until the
    next push places its movers: the new layout needs 14 x 8 + 16 + 48 = 176
    slots, the whole span, so the block is not re-laid alone but its segment
    is, before the next push.  The state then matches the FP64 oracle's."""
    from wumingpic2d_amd import weibel
    p = weibel.grid_params(16, 8, 4)
    jl, y = 3, p["nys"] + 3 + 0.5
    row = [(p["nxgs"] + 4.95, y, 0.9, 0.0)] * 12 + \
          [(p["nxgs"] + 5 + f, y, 0.0, 0.0) for f in np.linspace(0.1, 0.6, 30)]
    up, np2, cc, uf = _one_row_state(p, {jl: row})
    ctx = _ctx32(p)
    ctx.set_capacity(176)
    ctx.state_upload(up, np2, cc, uf)
    orc = Oracle(p)
    gp = up.copy(order="F")
    for _ in range(3):
        ctx.step(1)
        orc.step(up, gp, uf, cc, np2)
    st = ctx.capacity_stats()
    assert st["relayouts"] >= 1 and st["segment_relayouts"] >= 1 and st["max_bin_fill"] >= 42, st
    gup, gnp2, guf = np.zeros_like(up), np.zeros_like(np2), np.zeros_like(uf)
    ctx.state_download(up=gup, np2=gnp2, uf=guf)
    assert gnp2.sum() == np2.sum() == 42
    a, b = _by_id(up, np2), _by_id(gup, gnp2)
    assert a.keys() == b.keys()
    assert max(np.abs(a[k] - b[k]).max() for k in a) < 2e-5
    assert np.abs(guf - uf).max() < 1e-3 * np.abs(uf).max()


def test_fp32_row_overflow_is_reported():
    """A row of 160 slots holding 30 particles in cell 5 and 5 in cell 6
    (need 14 x 8 + 32 + 8 = 152) receives 12 particles from the row below
    into cell 5: it would need 168 slots, more than its segment holds, and
    the step stops with the reference's 'memory over' (its rows overflow np
    the same way, boundary_periodic.f90:228-236).  A row too full for its
    segment is refused at upload with the same error."""
    from wumingpic2d_amd import weibel
    p = weibel.grid_params(16, 8, 4)
    x5, x6 = p["nxgs"] + 5, p["nxgs"] + 6
    y3, y2 = p["nys"] + 3, p["nys"] + 2
    rows = {3: [(x5 + f, y3 + 0.5, 0.0, 0.0) for f in np.linspace(0.1, 0.6, 30)] +
               [(x6 + 0.5, y3 + 0.5, 0.0, 0.0)] * 5,
            2: [(x5 + 0.5, y2 + 0.95, 0.0, 0.9)] * 12}
    up, np2, cc, uf = _one_row_state(p, rows)
    ctx = _ctx32(p)
    ctx.set_capacity(160)
    ctx.state_upload(up, np2, cc, uf)
    with pytest.raises(wm.WmError) as e:
        ctx.step(1)
    assert e.value.code == wm.WM_ERR_NP_OVER
    ctx.close()
    full = {3: [(x5 + f, y3 + 0.5, 0.0, 0.0) for f in np.linspace(0.1, 0.6, 42)] +
               [(x6 + 0.5, y3 + 0.5, 0.0, 0.0)] * 5}
    up, np2, cc, uf = _one_row_state(p, full)
    ctx = _ctx32(p)
    ctx.set_capacity(160)
    with pytest.raises(wm.WmError) as e:
        ctx.state_upload(up, np2, cc, uf)
    assert e.value.code == wm.WM_ERR_NP_OVER


def test_fp32_tight_rows_relayout_and_track_oracle():
    """Row segments of 1.375 times a row's 512 particles: bins of about 16
    particles take 20 slots on average once rounded up to 8-slot blocks, so
    each bin has a few slots of room, a hot plasma (u_th = 0.2 c) fills bins
    every step and blocks of cells are laid out anew again and again.  The trajectory
    still tracks the FP64 oracle at the FP32 bars, and two runs are
    bit-identical."""
    p, up0, np20, cc0, uf0 = state(32, 16, 16, 50, vth=0.2)
    cap = 704
    outs = []
    for rep in range(2):
        up, np2, cc, uf = (a.copy(order="F") for a in (up0, np20, cc0, uf0))
        ctx = _ctx32(p)
        ctx.set_capacity(cap)
        ctx.state_upload(up, np2, cc, uf)
        ctx.step(6)
        st = ctx.capacity_stats()
        assert st["relayouts"] > 0, st
        o, n, f = np.zeros_like(up), np.zeros_like(np2), np.zeros_like(uf)
        ctx.state_download(up=o, np2=n, uf=f)
        outs.append((o, n, f))
        ctx.close()
    assert np.array_equal(outs[0][1], outs[1][1]) and np.array_equal(outs[0][2], outs[1][2])
    for s in range(2):
        for jl in range(p["ny"]):
            k = outs[0][1][jl, s]
            assert np.array_equal(outs[0][0][:, :k, jl, s].view(np.int64),
                                  outs[1][0][:, :k, jl, s].view(np.int64))
    up, np2, cc, uf = (a.copy(order="F") for a in (up0, np20, cc0, uf0))
    orc = Oracle(p)
    gp = up.copy(order="F")
    for _ in range(6):
        orc.step(up, gp, uf, cc, np2)
    gup, gnp2, guf = outs[0]
    assert gnp2.sum() == np2.sum()
    a, b = _by_id(up, np2), _by_id(gup, gnp2)
    assert a.keys() == b.keys()
    assert max(np.abs(a[k] - b[k]).max() for k in a) < 2e-5
    assert np.abs(guf - uf).max() < 1e-3 * np.abs(uf).max()


def test_fp32_dense_hot_chunks_spill_their_mover_lists():
    """16 cells per row (one push workgroup), 160 particles per cell and
    species, u_th = 0.5 c: about a third of the 5120 particles of a
    workgroup leave their cells each step, more than the 960 its LDS list
    holds, so the list continues in the row's spill pool.  The trajectory
    still tracks the FP64 oracle at the FP32 bars."""
    p, up, np2, cc, uf = state(16, 8, 160, 51, vth=0.5)
    ctx = _ctx32(p)
    ctx.state_upload(up, np2, cc, uf)
    orc = Oracle(p)
    gp = up.copy(order="F")
    for _ in range(3):
        orc.step(up, gp, uf, cc, np2)
        ctx.step(1)
    st = ctx.capacity_stats()
    assert st["max_workgroup_movers"] > st["workgroup_lds_list"] and st["max_row_spill"] > 0, st
    gup, gnp2, guf = np.zeros_like(up), np.zeros_like(np2), np.zeros_like(uf)
    ctx.state_download(up=gup, np2=gnp2, uf=guf)
    assert gnp2.sum() == np2.sum()
    a, b = _by_id(up, np2), _by_id(gup, gnp2)
    assert a.keys() == b.keys()
    assert max(np.abs(a[k] - b[k]).max() for k in a) < 2e-5
    assert np.abs(guf - uf).max() < 1e-3 * np.abs(uf).max()


def test_fp32_host_array_entry_points_refuse():
    p, up, np2, cc, uf = state(16, 4, 2, 45)
    ctx = _ctx32(p)
    with pytest.raises(wm.WmError) as e:
        ctx.particle__solv(np.zeros_like(up), up, uf, cc, p["nxgs"], p["nxge"])
    assert e.value.code == wm.WM_ERR_ARG


def test_fp32_learned_pc_flexible_cg_tracks_fp64_oracle():
    """The headline form: FP32 particles, and from the fourth step on the
    field solve by flexible CG with the learned preconditioner (trained on
    the device on pairs the first three plain solves recorded), against the
    FP64 oracle's plain CG at the FP32 tolerances above."""
    p, up, np2, cc, uf = state(64, 32, 16, 49, vth=0.1)
    ctx = _ctx32(p)
    ctx.state_upload(up, np2, cc, uf)
    ctx.pc_collect(8, 2)
    orc = Oracle(p)
    gp = up.copy(order="F")
    for _ in range(3):
        orc.step(up, gp, uf, cc, np2)
        ctx.step(1)
    assert ctx.pc_pair_count() > 0
    loss = ctx.pc_train(200, 3e-3, 1)
    assert loss[1] < loss[0]
    ctx.pc_collect(0, 0)
    ctx.set_solver(wm.SOLVER_FCG)
    fcg = plain = 0
    for _ in range(5):
        orc.step(up, gp, uf, cc, np2)
        ctx.step(1)
        a, b = ctx.solve_stats()
        fcg += int(a.sum())
        plain += int(b.sum())
    assert fcg > 0, "flexible CG did not run"
    gup, gnp2, guf = np.zeros_like(up), np.zeros_like(np2), np.zeros_like(uf)
    ctx.state_download(up=gup, np2=gnp2, uf=guf)
    assert gnp2.sum() == np2.sum()
    a, b = _by_id(up, np2), _by_id(gup, gnp2)
    assert a.keys() == b.keys()
    assert max(np.abs(a[k][:2] - b[k][:2]).max() for k in a) < 2e-5
    assert max(np.abs(a[k][2:] - b[k][2:]).max() for k in a) < 2e-5
    assert np.abs(guf - uf).max() < 1e-3 * np.abs(uf).max()
    e1, e2 = orc.energy(up, uf, np2), ctx.energy()
    assert np.allclose(e1[:2], e2[:2], rtol=1e-5)
    assert np.allclose(e1[2:], e2[2:], rtol=1e-2, atol=1e-9 * e1[:2].sum())
