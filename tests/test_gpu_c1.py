"""The 128 x 128 Weibel run at full size on the device against the oracle run
on this host's cores: 64 particles per cell per species (2.1e6 particles),
counter-streaming pair beams u_z = +/-0.5 c with 0.05 c thermal spread,
dt = 0.5, 200 steps, identical seeded initial state.

The energy history is the reference's own physics check (energy_history,
proj/weibel/app.f90:479-545): kinetic energy per species, E and B energy,
every 5 steps.  Bars (the north star's): each channel within 1e-3 of the
oracle's over the whole run, relative to that channel's largest value, and
the Weibel growth rate fitted to ln(B energy) over its exponential phase
(steps 60-160: B energy grows from about 20 to about 2000) within 2% of the
oracle's.

The FP32 run uses the headline form: FP32 particles and flexible CG with the
learned preconditioner, trained on the device on pairs the plain solver
records in the first steps."""
import os

import numpy as np
import pytest

import wumingpic2d_amd as wm
from oracle import Oracle
from wmhelpers import make_ctx, state

pytestmark = pytest.mark.gpu

NX = NY = 128
PPC = 64
STEPS = 200
EVERY = 5
FIT = (60, 160)


def _state():
    return state(NX, NY, PPC, 1, vth=0.05, udrift=0.5, np_cap=2 * PPC * NX)


def _growth_rate(hist, dt):
    k = np.arange(hist.shape[0]) * EVERY
    sel = (k >= FIT[0]) & (k <= FIT[1])
    slope = np.polyfit(k[sel] * dt, np.log(hist[sel, 3]), 1)[0]
    return 0.5 * slope


@pytest.fixture(scope="module")
def oracle_history():
    p, up, np2, cc, uf = _state()
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    orc = Oracle(p, threads=max(1, min(16, cores)))
    gp = up.copy(order="F")
    hist = [orc.energy(up, uf, np2)]
    for s in range(STEPS):
        orc.step(up, gp, uf, cc, np2)
        if (s + 1) % EVERY == 0:
            hist.append(orc.energy(up, uf, np2))
    return p, np.array(hist)


def _compare(p, hist, ref, tag):
    dev = np.abs(hist - ref).max(0) / np.abs(ref).max(0)
    tot = np.abs(hist.sum(1) - ref.sum(1)).max() / ref.sum(1).max()
    g, g0 = _growth_rate(hist, p["delt"]), _growth_rate(ref, p["delt"])
    print(f"{tag}: channel deviations {dev}, total {tot:.2e}, growth rate {g:.6f} vs "
          f"oracle {g0:.6f} ({abs(g / g0 - 1):.2e}); B energy {ref[0, 3]:.3e} -> {ref[-1, 3]:.3e}")
    assert ref[-1, 3] > 100 * ref[FIT[0] // EVERY, 3]   # the instability did grow
    assert (dev < 1e-3).all(), dev
    assert tot < 1e-3
    assert abs(g / g0 - 1) < 0.02


def test_c1_fp64_energy_history_and_growth_rate(oracle_history):
    p, ref = oracle_history
    _, up, np2, cc, uf = _state()
    ctx = make_ctx(p)
    ctx.state_upload(up, np2, cc, uf)
    hist = [ctx.energy()]
    for s in range(STEPS // EVERY):
        ctx.step(EVERY)
        hist.append(ctx.energy())
    _compare(p, np.array(hist), ref, "FP64 + CG")


def test_c1_fp32_learned_pc_energy_history_and_growth_rate(oracle_history):
    p, ref = oracle_history
    _, up, np2, cc, uf = _state()
    ctx = make_ctx(p, form=wm.FORM_FP32)
    ctx.state_upload(up, np2, cc, uf)
    ctx.pc_collect(16, 4)
    hist = [ctx.energy()]
    fcg = plain = 0
    for s in range(STEPS // EVERY):
        ctx.step(EVERY)
        if s == 0:
            loss = ctx.pc_train(300, 3e-3, 1)
            assert loss[1] < loss[0]
            ctx.pc_collect(0, 0)
            ctx.set_solver(wm.SOLVER_FCG)
        else:
            a, b = ctx.solve_stats()
            fcg += int(a.sum())
            plain += int(b.sum())
        hist.append(ctx.energy())
    assert fcg > 0, "flexible CG did not run"
    print(f"flexible-CG iterations of the last solves: {fcg}, plain: {plain}")
    _compare(p, np.array(hist), ref, "FP32 + learned PC")
