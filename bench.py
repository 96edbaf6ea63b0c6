#!/usr/bin/env python3
"""Throughput of the WumingPIC2D hot path on one MI355X.

One step = the loop body of proj/weibel/app.f90:102-106 (push, current
deposit, implicit field solve, particle boundaries, sort) over the whole
particle population, resident in HBM.  The default workload is the Weibel
configuration the headline metric is quoted on: 2048 x 2048 cells, 64
particles per cell per species, 2 species (5.37e8 particles), dt = 0.5,
counter-streaming beams u_z = +/-0.5 c with 0.05 c thermal spread,
synthetic state generated on the device.

Prints one JSON line.  ``roofline`` covers the dominant kernel (fused
push+deposit), timed with HIP events on the library's stream;
``cpu_baseline`` is the FP64 CPU restatement timed on this host's cores on a
bounded sample (256 x 256 cells at the same density).
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
    # name: (nx, ny, ppc per species, capacity factor)
    "weibel_2048": (2048, 2048, 64, 1.25),
    "weibel_128": (128, 128, 64, 1.5),
    "weibel_1024": (1024, 1024, 64, 1.25),
}

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


def cpu_baseline(seconds: float = 12.0) -> dict:
    """FP64 CPU restatement (the oracle) on a 256x256, 64 ppc/species sample."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np
    from oracle import Oracle
    from wumingpic2d_amd import weibel

    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    threads = max(1, min(16, cores))
    nx = ny = 256
    ppc = 64
    p = weibel.grid_params(nx, ny, ppc, np_cap=int(1.25 * ppc * nx))
    up, np2, cc, uf = weibel.init_state(p, seed=2, vth=0.05, udrift=0.5)
    orc = Oracle(p, threads=threads)
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
    return {"value": npart * steps / dt, "unit": "particle-steps/s", "cores": threads,
            "kind": "port",
            "sample": f"Weibel {nx}x{ny}, {ppc} ppc/species, 2 species ({npart} particles), "
                      f"{steps} steps after 1 warm-up, FP64 C restatement with {threads} OpenMP threads",
            "moments": {"value": npart * reps / dm, "unit": "particles/s", "cores": 1,
                        "sample": "mom_calc__accl + mom_calc__nvt + boundary_periodic__mom of the "
                                  f"same state, {reps} times, serial C restatement"}}


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
    nx, ny, ppc, capf = WORKLOADS[workload]
    p = weibel.grid_params(nx, ny, ppc)
    cap = int(capf * ppc * nx)
    ctx = wm.Context(0)
    ctx.set_form(wm.FORM_FP32 if form == "fp32" else wm.FORM_FP64)
    ctx.set_capacity(cap, max(1024, cap // 16))
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


def run_leg(wm, weibel, args, form: str, solver: str, steps: int, warmup: int) -> dict:
    """Warm-up (plain solves; with the learned solver they record pairs and
    the net is trained after them), then `steps` timed steps."""
    import numpy as np

    ctx, p, cap = make_context(wm, weibel, args.workload, form)
    ppc = WORKLOADS[args.workload][2]
    ctx.init_weibel(args.seed, ppc, 0.05, 0.5)
    npart = ctx.particle_count()
    log(f"{npart} particles generated on the device")
    train = {}
    if solver == "fcg":
        ctx.pc_collect(16, 4)
    for w in range(warmup):
        ctx.step(1)
        log(f"warm-up step {w + 1}/{warmup}")
    if solver == "fcg":
        t0 = time.perf_counter()
        loss = ctx.pc_train(args.pc_steps, 3e-3, args.seed)
        train = {"pairs": ctx.pc_pair_count(), "adam_steps": args.pc_steps,
                 "seconds": round(time.perf_counter() - t0, 3),
                 "loss_first": float(loss[0]), "loss_last10": float(loss[1])}
        log(f"learned preconditioner trained: {train}")
        ctx.pc_collect(0, 0)
        ctx.set_solver(wm.SOLVER_FCG)
    phase = np.zeros(6)
    iters = fcg_iters = plain_iters = 0
    t0 = time.perf_counter()
    for k in range(steps):
        ctx.step(1)
        phase += ctx.phase_times()[:6]
        iters += int(ctx.cg_iters().sum())
        a, b = ctx.solve_stats()
        fcg_iters += int(a.sum())
        plain_iters += int(b.sum())
        if (k + 1) % 5 == 0 or k + 1 == steps:
            log(f"timed step {k + 1}/{steps} ({form}, {solver})")
    elapsed = time.perf_counter() - t0
    energy = ctx.energy()
    # output-step moment diagnostics (mom_calc__accl + mom_calc__nvt +
    # boundary_periodic__mom) on the final state, outside the timed steps
    mom_ms = np.zeros(2)
    for _ in range(MOM_REPEATS):
        ctx.moments()
        mom_ms += ctx.phase_times()[6:8]
    mom_ms /= MOM_REPEATS
    log(f"moments: {mom_ms[0]:.3f} ms fused pass + gather, {mom_ms[1]:.3f} ms fold")
    ctx.close()
    return {"form": form, "solver": solver, "particles": npart, "steps": steps,
            "elapsed": elapsed, "ms_per_step": 1e3 * elapsed / steps,
            "value": npart * steps / elapsed,
            "phase_ms_per_step": {n: round(float(v) / steps, 4) for n, v in zip(PHASES, phase)},
            "push_deposit_ms": float(phase[1]) / steps,
            "solve_ms_per_step": round(float(phase[3]) / steps, 4),
            "cg_iters_per_step": iters / steps,
            "fcg_iters_per_step": fcg_iters / steps, "plain_iters_per_step": plain_iters / steps,
            "energy": [float(e) for e in energy], "pc_training": train, "p": p,
            "moments_ms": [float(v) for v in mom_ms]}


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
    achieved = nbytes / (ms * 1e-3) / 1e9
    return {"kernel": "k_mom (fused half-step push + CIC moments) + k_mom_gather",
            "ms": round(ms, 4), "fold_ms": round(fold_ms, 4), "bytes": nbytes,
            "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4)}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="weibel_2048", choices=sorted(WORKLOADS))
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-compare", action="store_true",
                    help="skip the plain FP64 comparison leg")
    ap.add_argument("--seed", type=int, default=2)
    ap.add_argument("--form", default="fp32", choices=["fp64", "fp32"])
    ap.add_argument("--solver", default=None, choices=["cg", "fcg"],
                    help="implicit solve: plain CG or flexible CG with the learned "
                         "preconditioner (default: fcg for fp32, cg for fp64)")
    ap.add_argument("--pc-steps", type=int, default=1000,
                    help="Adam steps of the preconditioner (at 2048^2: 200 give 3 flexible-CG "
                         "iterations per component, 1000 give 2)")
    ap.add_argument("--copy-calib", action="store_true",
                    help="only run a 1 GiB device-to-device copy (HBM counter calibration)")
    args = ap.parse_args()
    if args.copy_calib:
        import torch

        n = 1 << 27                                   # 2^27 doubles = 1 GiB
        a = torch.ones(n, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)                                # runtime copy kernel
            torch.add(a, 1.0, out=b)                  # elementwise kernel: 1 GiB in, 1 GiB out
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes_read": 8 * n, "copy_calibration_bytes_written": 8 * n}))
        return 0
    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build has no multi-GPU path (asked for --gpus "
              f"{args.gpus}); the domain would split into y-slabs of rows, one per GPU",
              file=sys.stderr)
        return 2
    solver = args.solver or ("fcg" if args.form == "fp32" else "cg")

    import wumingpic2d_amd as wm
    from wumingpic2d_amd import weibel

    head = run_leg(wm, weibel, args, args.form, solver, args.steps, args.warmup)
    nx, ny, ppc, _ = WORKLOADS[args.workload]
    npart = head["particles"]
    bpp = PUSHDEP_BYTES_PER_PARTICLE if args.form == "fp64" else PUSHDEP_BYTES_PER_PARTICLE_FP32
    # The profile records the push's launches per step (one; a push split in
    # several launches would cover its share of the particles in each, so the
    # per-launch rate is the step's rate).
    traffic, launches = load_traffic(args.workload, args.form)
    launches = launches or 1
    launch_ms = head["push_deposit_ms"] / launches
    achieved = bpp * npart / launches / (launch_ms * 1e-3) / 1e9
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": traffic,
                "kernel": ("k_particle<PUSH,DEPOSIT>" if args.form == "fp64" else "k_push_deposit32")
                + " (fused push+deposit)",
                "bytes_per_particle": bpp, "launches_per_step": launches,
                "launch_ms": round(launch_ms, 4)}
    p = head["p"]
    out = {
        "metric": METRIC,
        "value": head["value"],
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.form == "fp64" else "f32 particles, f64 current/solve/diagnostics",
        "data": "synthetic",
        "config": {"workload": f"{args.workload}: Weibel pair plasma {nx}x{ny}, {ppc} ppc/species, "
                               f"{args.form.upper()} form, {solver.upper()} solve",
                   "nx": nx, "ny": ny, "ppc_per_species": ppc, "species": 2, "particles": npart,
                   "dt": p["delt"], "form": args.form, "solver": solver,
                   "parallelism": "single GPU"},
        "solve_ms_per_step": head["solve_ms_per_step"],
        "cg_iters_per_step": head["cg_iters_per_step"],
        "fcg_iters_per_step": head["fcg_iters_per_step"],
        "plain_iters_per_step": head["plain_iters_per_step"],
        "pc_training": head["pc_training"],
        "phase_ms_per_step": head["phase_ms_per_step"],
        "energy_last": head["energy"],
        "roofline": roofline,
        "moments": moments_line(head, args.form, nx, ny),
    }
    if not args.no_compare and not (args.form == "fp64" and solver == "cg"):
        ref = run_leg(wm, weibel, args, "fp64", "cg", args.steps, args.warmup)
        e1, e0 = head["energy"], ref["energy"]
        tot0 = sum(e0)
        out["comparison_fp64_cg"] = {
            "ms_per_step": ref["ms_per_step"], "value": ref["value"],
            "solve_ms_per_step": ref["solve_ms_per_step"],
            "cg_iters_per_step": ref["cg_iters_per_step"],
            "phase_ms_per_step": ref["phase_ms_per_step"],
            "moments": moments_line(ref, "fp64", nx, ny),
            "step_rate_ratio": round(ref["ms_per_step"] / head["ms_per_step"], 3),
            "solve_iters_ratio": round(ref["cg_iters_per_step"] / max(head["cg_iters_per_step"], 1e-9), 3),
            "energy_rel_diff_total": abs(sum(e1) - tot0) / tot0,
            "energy_rel_diff": [abs(a - b) / max(abs(b), 1e-300) for a, b in zip(e1, e0)],
        }
    if not args.no_cpu_baseline:
        log("timing the CPU restatement")
        out["cpu_baseline"] = cpu_baseline()
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
