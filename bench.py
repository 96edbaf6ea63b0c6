#!/usr/bin/env python3
"""Throughput of the WumingPIC2D hot path on one MI355X.

One step = the loop body of proj/weibel/app.f90:102-106 (push, current
deposit, implicit field solve, particle boundaries, sort) over the whole
particle population, resident in HBM.  The default workload is the Weibel
configuration the headline metric is quoted on: 2048 x 2048 cells, 64
particles per cell per species, 2 species (5.37e8 particles), dt = 0.5,
counter-streaming beams u_z = +/-0.5 c with 0.05 c thermal spread,
synthetic state generated on the device.

Prints one JSON line.  ``value`` is the headline form's rate (FP32
particles, flexible CG with the learned preconditioner) over the timed steps
after the warm-up; ``roofline`` covers the dominant kernel (fused
push+deposit), timed with HIP events on the library's stream; ``cpu_baseline``
is the FP64 CPU restatement timed on this host's cores on a bounded sample
(256 x 256 cells at the same density), with the reference's own OpenMP build
of the push and the sort timed beside the restatement's where it was built.

Both forms then run on to the workload's full length (2000 steps at 2048^2)
for the reference's physics check: the energy history every 10 steps
(energy_history, proj/weibel/app.f90:479-545), the Weibel growth rate fitted
to the magnetic energy of the linear phase, the step rate of the nonlinear
phase, and the FP32 store's capacity use (``full_run``).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "particle-steps/s (push+deposit+sort) & implicit-solve ms: FP32+learned PC vs FP64"
HBM_PEAK_GBS = 8000.0

WORKLOADS = {
    # name: (nx, ny, ppc per species, capacity factor, full run length).
    # Capacity: slots per (species, row) = factor x ppc x nx.  The reference
    # sizes rows at 5 x (proj/weibel/app.f90:248); 2 x leaves a cell's bin
    # room for about 6 standard deviations of its count (Poisson, 64 ppc)
    # once the bins are rounded to 8 slots, so bins overflow (and are laid
    # out anew) rarely even where the nonlinear phase moves particles fast.
    "weibel_2048": (2048, 2048, 64, 2.0, 2000),
    "weibel_128": (128, 128, 64, 2.0, 200),
    "weibel_1024": (1024, 1024, 64, 2.0, 2000),
    # 8192 x 8192, 32 ppc/species (4.3e9 particles): rows of 1.4 x the mean,
    # 32768 mover records per row and two training pairs per slot ring keep
    # the FP32 state near 250 GB of the card's 288; the FP64 form of it
    # (412 GB of particle banks) does not fit, so it has no comparison leg.
    # 100 steps (the linear phase): the nonlinear phase moves up to a
    # quarter of a row's particles in one step, which at 8192 cells needs
    # about 120k records per row, more pool than fits beside the bins.
    "weibel_8192": (8192, 8192, 32, 1.4, 100),
}
# per workload: FP32 mover records per row (0: a quarter of the row's
# slots), the learned preconditioner's pair ring and pairs per solve, its
# Adam steps, and whether the FP64 comparison leg fits
WORKLOAD_EXTRA = {"weibel_8192": {"mrow": 32768, "pairs": (2, 2), "pc_steps": 100,
                                  "compare": False, "moments": False}}


def extra(workload: str, key: str, default):
    return WORKLOAD_EXTRA.get(workload, {}).get(key, default)
EVERY = 10            # steps between energy-history entries
FIT = (60, 160)       # linear-phase window (steps) of the growth-rate fit

# Algorithmic bytes per particle of the fused push+deposit kernel.  FP64 form:
# read x, y, ux, uy, uz (40 B) and write the same 40 B.  FP32 form (cell
# bins, in place): read the offsets and momenta (20 B; the cell is the bin's)
# and write them back (20 B); the few particles that change cell go to the
# mover list instead.
PUSHDEP_BYTES_PER_PARTICLE = 80.0
PUSHDEP_BYTES_PER_PARTICLE_FP32 = 40.0
# The moment pass reads each particle once: FP64 x, y, u (40 B) or the FP32
# bin slot (offsets and u, 20 B).
MOM_BYTES_PER_PARTICLE = {"fp64": 40.0, "fp32": 20.0}
MOM_REPEATS = 3


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def host_cores() -> int:
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def cpu_baseline(seconds: float = 12.0) -> dict:
    """FP64 CPU restatement (the oracle) on a 256x256, 64 ppc/species sample,
    OpenMP over every core of this process's affinity mask; beside it, where
    oracle/_ref holds the reference's own OpenMP build of particle.f90 and
    sort.f90, the reference's push and sort against the restatement's on the
    same state and thread count."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np
    import oracle as orm
    from wumingpic2d_amd import weibel

    threads = host_cores()
    os.environ["OMP_NUM_THREADS"] = str(threads)
    nx = ny = 256
    ppc = 64
    p = weibel.grid_params(nx, ny, ppc, np_cap=int(1.25 * ppc * nx))
    up, np2, cc, uf = weibel.init_state(p, seed=2, vth=0.05, udrift=0.5)
    orc = orm.Oracle(p, threads=threads)
    gp = up.copy(order="F")
    npart = int(np2.sum())
    orc.step(up, gp, uf, cc, np2)                    # warm-up
    steps, t0 = 0, time.perf_counter()
    while steps < 20:
        orc.step(up, gp, uf, cc, np2)
        steps += 1
        log(f"cpu baseline step {steps}")
        if time.perf_counter() - t0 > seconds:
            break
    dt = time.perf_counter() - t0
    out = {"value": npart * steps / dt, "unit": "particle-steps/s", "cores": threads,
           "kind": "port",
           "sample": f"Weibel {nx}x{ny}, {ppc} ppc/species, 2 species ({npart} particles), "
                     f"{steps} steps after 1 warm-up, FP64 C restatement with {threads} OpenMP "
                     f"threads"}
    if orm.ref_available(omp=True):
        ref = orm.RefBuild(p, omp=True)
        gp = up.copy(order="F")                      # both outputs start from the same bits
        gpr = up.copy(order="F")

        def timed(f, n=5):
            f()
            t = time.perf_counter()
            for _ in range(n):
                f()
            return (time.perf_counter() - t) / n * 1e3

        push_port = timed(lambda: orc.particle__solv(gp, up, uf, cc))
        push_ref = timed(lambda: ref.particle__solv(gpr, up, uf, cc))
        same = bool(np.array_equal(gp[:5].view(np.int64), gpr[:5].view(np.int64)))
        sort_port = timed(lambda: orc.sort__bucket(gp, up, cc.copy(order="F"), np2))
        sort_ref = timed(lambda: ref.sort__bucket(gpr, up, cc.copy(order="F"), np2))
        out["reference_calibration"] = {
            "what": "the reference's own particle__solv and sort__bucket (particle.f90, sort.f90 "
                    "built unchanged with amdflang -O2 -fopenmp) against the restatement's, same "
                    f"state, {threads} threads",
            "push_ms_reference": round(push_ref, 2), "push_ms_port": round(push_port, 2),
            "sort_ms_reference": round(sort_ref, 2), "sort_ms_port": round(sort_port, 2),
            "push_bits_equal": same,
            "reference_time_over_port_time": round((push_ref + sort_ref) / (push_port + sort_port), 3)}
        log(f"reference calibration: {out['reference_calibration']}")
    # the output-step moment diagnostics on the same sample (serial C port)
    mom = np.zeros((7, nx + 2, ny + 2, p["nsp"]), order="F")
    reps, t1 = 0, time.perf_counter()
    while reps < 5:
        orc.mom_calc_accl(gp, up, uf, cc)
        orc.mom_calc_nvt(mom, gp, np2)
        orc.bc_mom(mom)
        reps += 1
        if time.perf_counter() - t1 > 4.0:
            break
    dm = time.perf_counter() - t1
    out["moments"] = {"value": npart * reps / dm, "unit": "particles/s", "cores": 1,
                      "sample": "mom_calc__accl + mom_calc__nvt + boundary_periodic__mom of the "
                                f"same state, {reps} times, serial C restatement"}
    return out


def copy_bandwidth(nbytes: int = 1 << 31, reps: int = 10) -> float:
    """HBM GB/s of the runtime's device-to-device copy of nbytes (each copy
    reads and writes nbytes), through the HIP runtime the library uses."""
    import ctypes

    hip = ctypes.CDLL("libamdhip64.so.7")
    vp = ctypes.c_void_p
    hip.hipMalloc.argtypes = [ctypes.POINTER(vp), ctypes.c_size_t]
    hip.hipMemcpy.argtypes = [vp, vp, ctypes.c_size_t, ctypes.c_int]
    hip.hipEventCreate.argtypes = [ctypes.POINTER(vp)]
    hip.hipEventRecord.argtypes = [vp, vp]
    hip.hipEventSynchronize.argtypes = [vp]
    hip.hipEventElapsedTime.argtypes = [ctypes.POINTER(ctypes.c_float), vp, vp]
    hip.hipFree.argtypes = [vp]
    hip.hipEventDestroy.argtypes = [vp]
    a, b, e0, e1 = vp(), vp(), vp(), vp()
    if hip.hipMalloc(ctypes.byref(a), nbytes) or hip.hipMalloc(ctypes.byref(b), nbytes):
        return float("nan")
    hip.hipEventCreate(ctypes.byref(e0))
    hip.hipEventCreate(ctypes.byref(e1))
    hip.hipMemcpy(b, a, nbytes, 3)                   # hipMemcpyDeviceToDevice
    hip.hipEventRecord(e0, None)
    for _ in range(reps):
        hip.hipMemcpy(b, a, nbytes, 3)
    hip.hipEventRecord(e1, None)
    hip.hipEventSynchronize(e1)
    ms = ctypes.c_float()
    hip.hipEventElapsedTime(ctypes.byref(ms), e0, e1)
    for x in (a, b):
        hip.hipFree(x)
    for x in (e0, e1):
        hip.hipEventDestroy(x)
    return 2.0 * nbytes * reps / (ms.value * 1e-3) / 1e9


def load_traffic(workload: str, form: str):
    """(HBM bytes per push+deposit launch, launches per step) from the
    committed rocprofv3 --pmc summary (FETCH_SIZE and WRITE_SIZE,
    copy-calibrated), or (None, None)."""
    path = os.path.join(ROOT, "profiles", "pmc_traffic.json")
    try:
        with open(path) as f:
            t = json.load(f)
        w = t["workloads"][workload][form]
        return w["bytes_per_launch"], w.get("launches_per_step", 1)
    except (OSError, KeyError, ValueError):
        return None, None


def make_context(wm, weibel, workload: str, form: str):
    nx, ny, ppc, capf, _ = WORKLOADS[workload]
    p = weibel.grid_params(nx, ny, ppc)
    cap = int(capf * ppc * nx)
    ctx = wm.Context(0)
    ctx.set_form(wm.FORM_FP32 if form == "fp32" else wm.FORM_FP64)
    ctx.set_capacity(cap, extra(workload, "mrow", 0) if form == "fp32" else max(1024, cap // 16))
    g = (p["ndim"], p["np"], p["nsp"], p["nxgs"], p["nxge"], p["nygs"], p["nyge"], p["nys"], p["nye"])
    ctx.boundary_periodic__init(*g, 0, 0, 0, 0, 0, 0, [0], p["delx"], p["delt"], p["c"])
    ctx.particle__init(*g, p["delx"], p["delt"], p["c"], p["q"], p["r"])
    ctx.field__init(*g, 0, 0, 0, 0, p["delx"], p["delt"], p["c"], p["q"], p["r"], p["gfac"])
    ctx.sort__init(*g)
    ctx.mom_calc__init(*g, p["delx"], p["delt"], p["c"], p["q"], p["r"])
    log(f"workload {workload} ({form}): {nx}x{ny}, {ppc} ppc/species, capacity {cap}/row")
    return ctx, p, cap


PHASES = ["cell_fields", "push_deposit", "current_gather_fold", "field_solve", "particle_bc",
          "sort"]


class Steps:
    """Steps of one context with per-phase device times, iteration counts and
    the energy history (every EVERY steps from step 0)."""

    def __init__(self, ctx, np):
        self.ctx, self.np = ctx, np
        self.n = 0
        self.hist = [ctx.energy()]

    def run(self, k: int, tag: str = "") -> dict:
        np = self.np
        phase = np.zeros(6)
        iters = fcg = plain = 0
        t0 = time.perf_counter()
        for i in range(k):
            self.ctx.step(1)
            phase += self.ctx.phase_times()[:6]
            iters += int(self.ctx.cg_iters().sum())
            a, b = self.ctx.solve_stats()
            fcg += int(a.sum())
            plain += int(b.sum())
            self.n += 1
            if self.n % EVERY == 0:
                self.hist.append(self.ctx.energy())
            if tag and ((i + 1) % 100 == 0 or i + 1 == k):
                log(f"{tag}: step {self.n}")
        elapsed = time.perf_counter() - t0
        return {"elapsed": elapsed, "ms_per_step": 1e3 * elapsed / max(k, 1), "steps": k,
                "phase": phase / max(k, 1), "iters": iters / max(k, 1), "fcg": fcg / max(k, 1),
                "plain": plain / max(k, 1)}


def run_leg(wm, weibel, args, form: str, solver: str) -> dict:
    """Warm-up (plain solves; with the learned solver they record pairs and
    the net is trained after them), the timed steps, and (FP32 with the
    learned solver) as many timed steps with plain CG on the same state; then
    on to the full length with the energy history."""
    import numpy as np

    ctx, p, cap = make_context(wm, weibel, args.workload, form)
    ppc, full = WORKLOADS[args.workload][2], args.full_steps
    ctx.init_weibel(args.seed, ppc, 0.05, 0.5)
    npart = ctx.particle_count()
    log(f"{npart} particles generated on the device")
    S = Steps(ctx, np)
    train = {}
    if solver == "fcg":
        ctx.pc_collect(*extra(args.workload, "pairs", (16, 4)))
    S.run(args.warmup, f"{form} warm-up")
    if solver == "fcg":
        t0 = time.perf_counter()
        loss = ctx.pc_train(args.pc_steps, 3e-3, args.seed)
        train = {"pairs": ctx.pc_pair_count(), "adam_steps": args.pc_steps,
                 "seconds": round(time.perf_counter() - t0, 3),
                 "loss_first": float(loss[0]), "loss_last10": float(loss[1])}
        log(f"learned preconditioner trained: {train}")
        ctx.pc_collect(0, 0)
        ctx.set_solver(wm.SOLVER_FCG)
    timed = S.run(args.steps, f"{form} timed")
    plain = None
    if solver == "fcg":
        ctx.set_solver(wm.SOLVER_CG)
        plain = S.run(args.steps, f"{form} timed, plain CG")
        ctx.set_solver(wm.SOLVER_FCG)
    late = None
    if full > S.n:
        tail = min(100, full - S.n)
        S.run(full - S.n - tail, f"{form} full run")
        late = S.run(tail, f"{form} full run, last {tail}")
    energy = ctx.energy()
    caps = ctx.capacity_stats() if form == "fp32" else None
    if form == "fp64":
        n2 = np.zeros((p["ny"], p["nsp"]), dtype=np.int32, order="F")
        ctx.state_download(np2=n2)
        caps = {"max_row_fill": int(n2.max()), "slots_per_row": cap}
    # output-step moment diagnostics (mom_calc__accl + mom_calc__nvt +
    # boundary_periodic__mom) on the final state, outside the timed steps
    mom_ms = np.zeros(2)
    reps = MOM_REPEATS if extra(args.workload, "moments", True) else 0
    for _ in range(reps):
        ctx.moments()
        mom_ms += ctx.phase_times()[6:8]
    mom_ms /= max(reps, 1)
    log(f"moments: {mom_ms[0]:.3f} ms fused pass + gather, {mom_ms[1]:.3f} ms fold")
    ctx.close()
    return {"form": form, "solver": solver, "particles": npart, "nsteps": S.n, "timed": timed,
            "plain": plain,
            "late": late, "hist": np.array(S.hist), "caps": caps, "energy": [float(e) for e in energy],
            "pc_training": train, "p": p, "moments_ms": [float(v) for v in mom_ms]}


def leg_line(leg: dict) -> dict:
    t = leg["timed"]
    return {"ms_per_step": t["ms_per_step"], "value": leg["particles"] / (t["ms_per_step"] * 1e-3),
            "phase_ms_per_step": {n: round(float(v), 4) for n, v in zip(PHASES, t["phase"])},
            "solve_ms_per_step": round(float(t["phase"][3]), 4), "cg_iters_per_step": t["iters"],
            "fcg_iters_per_step": t["fcg"], "plain_iters_per_step": t["plain"]}


def growth_rate(hist, dt: float) -> float:
    """Half the slope of ln(B energy) against time over the linear phase."""
    import numpy as np

    k = np.arange(hist.shape[0]) * EVERY
    sel = (k >= FIT[0]) & (k <= FIT[1])
    b = hist[sel, -1]
    if sel.sum() < 3 or not (b > 0).all():
        return float("nan")
    return float(0.5 * np.polyfit(k[sel] * dt, np.log(b), 1)[0])


def moments_line(leg: dict, form: str, nx: int, ny: int) -> dict:
    """Device time and HBM rate of the output-step moment pass (k_mom over
    the particles, then the tile gather; the fold apart).  Algorithmic bytes:
    the particles once, the six field planes once, the per-tile node sums
    written and read back, mom written."""
    nsp = 2
    tiles = nsp * -(-ny // 4) * -(-nx // 32) * 7 * 6 * 34 * 8   # 4 rows x 32 cells per tile
    fields = 6 * (nx + 4) * (ny + 4) * 8
    mom = 7 * (nx + 2) * (ny + 2) * nsp * 8
    nbytes = leg["particles"] * MOM_BYTES_PER_PARTICLE[form] + fields + 2 * tiles + mom
    ms, fold_ms = leg["moments_ms"]
    if ms <= 0:
        return None                     # not run for this workload
    achieved = nbytes / (ms * 1e-3) / 1e9
    return {"kernel": "k_mom (fused half-step push + CIC moments) + k_mom_gather",
            "ms": round(ms, 4), "fold_ms": round(fold_ms, 4), "bytes": nbytes,
            "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4)}


def full_run_line(head: dict, ref: dict | None, dt: float) -> dict:
    """Energy histories of the full run: FP32 + learned PC against the plain
    FP64 GPU run from the same initial state, growth rates, nonlinear-phase
    step rates, capacity use."""
    import numpy as np

    h = head["hist"]
    out = {"steps": head["nsteps"], "every": EVERY,
           "fit_window_steps": list(FIT), "growth_rate": {"fp32_fcg": growth_rate(h, dt)},
           "b_energy_first_last": [float(h[0, -1]), float(h[-1, -1])],
           "capacity": head["caps"]}
    if head["late"]:
        out["late_ms_per_step"] = {"fp32_fcg": round(head["late"]["ms_per_step"], 3)}
        out["late_phase_ms_per_step"] = {
            "fp32_fcg": {n: round(float(v), 4) for n, v in zip(PHASES, head["late"]["phase"])}}
    if head["plain"] is not None:
        out["pc_training_amortized_ms_per_step"] = round(
            head["pc_training"]["seconds"] * 1e3 / max(out["steps"], 1), 4)
    if ref is not None:
        r = ref["hist"]
        n = min(h.shape[0], r.shape[0])
        dev = np.abs(h[:n] - r[:n]).max(0) / np.abs(r[:n]).max(0)
        tot = np.abs(h[:n].sum(1) - r[:n].sum(1)).max() / r[:n].sum(1).max()
        out["growth_rate"]["fp64_cg"] = growth_rate(r, dt)
        g32, g64 = out["growth_rate"]["fp32_fcg"], out["growth_rate"]["fp64_cg"]
        out["growth_rate"]["rel_diff"] = abs(g32 / g64 - 1) if g64 == g64 and g64 else None
        out["energy_hist_max_rel"] = {"channels": [float(v) for v in dev], "total": float(tot),
                                      "channel_names": ["kinetic e", "kinetic p", "E", "B"],
                                      "relative_to": "each channel's largest value over the run"}
        out["capacity_fp64"] = ref["caps"]
        if ref["late"]:
            out["late_ms_per_step"]["fp64_cg"] = round(ref["late"]["ms_per_step"], 3)
            out["late_phase_ms_per_step"]["fp64_cg"] = {
                n: round(float(v), 4) for n, v in zip(PHASES, ref["late"]["phase"])}
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="weibel_2048", choices=sorted(WORKLOADS))
    ap.add_argument("--full-steps", type=int, default=None,
                    help="run length of the energy-history check (default: the workload's; "
                         "0: stop after the timed steps)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-compare", action="store_true",
                    help="skip the plain FP64 comparison leg")
    ap.add_argument("--seed", type=int, default=2)
    ap.add_argument("--form", default="fp32", choices=["fp64", "fp32"])
    ap.add_argument("--solver", default=None, choices=["cg", "fcg"],
                    help="implicit solve: plain CG or flexible CG with the learned "
                         "preconditioner (default: fcg for fp32, cg for fp64)")
    ap.add_argument("--pc-steps", type=int, default=None,
                    help="Adam steps of the preconditioner (default: 1000; 100 at 8192^2)")
    ap.add_argument("--copy-calib", action="store_true",
                    help="only run a 1 GiB device-to-device copy (HBM counter calibration)")
    args = ap.parse_args()
    if args.copy_calib:
        print(json.dumps({"copy_gbs": copy_bandwidth(1 << 30, 3),
                          "copy_calibration_bytes_read": 1 << 30,
                          "copy_calibration_bytes_written": 1 << 30}))
        return 0
    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build has no multi-GPU path (asked for --gpus "
              f"{args.gpus}); the domain would split into y-slabs of rows, one per GPU",
              file=sys.stderr)
        return 2
    solver = args.solver or ("fcg" if args.form == "fp32" else "cg")
    nx, ny, ppc, _, full = WORKLOADS[args.workload]
    if args.full_steps is None:
        args.full_steps = full
    if args.pc_steps is None:
        args.pc_steps = extra(args.workload, "pc_steps", 1000)

    import wumingpic2d_amd as wm
    from wumingpic2d_amd import weibel

    head = run_leg(wm, weibel, args, args.form, solver)
    hl = leg_line(head)
    npart = head["particles"]
    bpp = PUSHDEP_BYTES_PER_PARTICLE if args.form == "fp64" else PUSHDEP_BYTES_PER_PARTICLE_FP32
    # The profile records the push's launches per step (one; a push split in
    # several launches would cover its share of the particles in each, so the
    # per-launch rate is the step's rate).
    traffic, launches = load_traffic(args.workload, args.form)
    launches = launches or 1
    launch_ms = hl["phase_ms_per_step"]["push_deposit"] / launches
    achieved = bpp * npart / launches / (launch_ms * 1e-3) / 1e9
    copy_gbs = copy_bandwidth()
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": traffic,
                "kernel": ("k_particle<PUSH,DEPOSIT>" if args.form == "fp64" else "k_push_deposit32")
                + " (fused push+deposit)",
                "bytes_per_particle": bpp, "launches_per_step": launches,
                "launch_ms": round(launch_ms, 4),
                "copy_gbs_measured": round(copy_gbs, 1),
                "frac_of_copy": round(achieved / copy_gbs, 4) if copy_gbs == copy_gbs else None}
    p = head["p"]
    out = {
        "metric": METRIC,
        "value": hl["value"],
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": hl["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.form == "fp64" else "f32 particles, f64 current/solve/diagnostics",
        "data": "synthetic",
        "config": {"workload": f"{args.workload}: Weibel pair plasma {nx}x{ny}, {ppc} ppc/species, "
                               f"{args.form.upper()} form, {solver.upper()} solve",
                   "nx": nx, "ny": ny, "ppc_per_species": ppc, "species": 2, "particles": npart,
                   "dt": p["delt"], "form": args.form, "solver": solver,
                   "slots_per_row": int(WORKLOADS[args.workload][3] * ppc * nx),
                   "parallelism": "single GPU"},
        "solve_ms_per_step": hl["solve_ms_per_step"],
        "cg_iters_per_step": hl["cg_iters_per_step"],
        "fcg_iters_per_step": hl["fcg_iters_per_step"],
        "plain_iters_per_step": hl["plain_iters_per_step"],
        "pc_training": head["pc_training"],
        "phase_ms_per_step": hl["phase_ms_per_step"],
        "energy_last": head["energy"],
        "roofline": roofline,
        "moments": moments_line(head, args.form, nx, ny),
    }
    if head["plain"] is not None:
        pl = head["plain"]
        run_len = WORKLOADS[args.workload][4]
        train_ms = head["pc_training"]["seconds"] * 1e3 / run_len
        eff = hl["ms_per_step"] + train_ms
        out["solve_fp32_same_state"] = {
            "solve_ms_fcg": hl["solve_ms_per_step"], "solve_ms_plain": round(float(pl["phase"][3]), 4),
            "iters_fcg": hl["fcg_iters_per_step"], "iters_plain": pl["iters"],
            "ms_per_step_fcg": round(hl["ms_per_step"], 3), "ms_per_step_plain": round(pl["ms_per_step"], 3),
            # the preconditioner's training spread over the workload's whole run
            "training_ms_per_step": round(train_ms, 4), "run_steps": run_len,
            "effective_ms_per_step_fcg": round(eff, 3),
            "learned_pc_pays": bool(eff < pl["ms_per_step"]),
            "what": f"{args.steps} steps of each, one after the other on the same FP32 state; "
                    "the second stream holds only the bins' fill check"}
    ref = None
    if (not args.no_compare and extra(args.workload, "compare", True)
            and not (args.form == "fp64" and solver == "cg")):
        ref = run_leg(wm, weibel, args, "fp64", "cg")
        rl = leg_line(ref)
        e1, e0 = head["energy"], ref["energy"]
        tot0 = sum(e0)
        out["comparison_fp64_cg"] = dict(rl, **{
            "moments": moments_line(ref, "fp64", nx, ny),
            "step_rate_ratio": round(rl["ms_per_step"] / hl["ms_per_step"], 3),
            "solve_iters_ratio": round(rl["cg_iters_per_step"] / max(hl["cg_iters_per_step"], 1e-9), 3),
            "energy_rel_diff_total_final": abs(sum(e1) - tot0) / tot0,
            "energy_rel_diff_final": [abs(a - b) / max(abs(b), 1e-300) for a, b in zip(e1, e0)]})
    if args.full_steps > 0:
        out["full_run"] = full_run_line(head, ref, p["delt"])
    if not args.no_cpu_baseline:
        log("timing the CPU restatement")
        out["cpu_baseline"] = cpu_baseline()
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
