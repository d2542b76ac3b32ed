#!/usr/bin/env python3
"""Benchmark of the FP64 RK4 Hall-MHD step on one MI355X.

Workload (default): the tanh-stretched 512x256x512 Harris sheet with Hall
term and guide field (beta 0.1, By 0.1, eta 0.002, d_i 0.25), 200 steps: the
largest single-GPU configuration and the one the 60% HBM-roofline target is
quoted on.  A "step" is one full RK4 step of the reference's stepon (four
right-hand sides with every smoothing sweep and boundary fill, the post-step
perturbation smoothing) plus its setdt, with the state resident in HBM.

Prints ONE JSON line:
  value          cell-updates/s = mx*ny*mz*steps / T (physical cells)
  roofline       dominant kernel family: algorithmic bytes / HIP-event time
  step_roofline  whole step: value * 6256 B (5104 B without Hall) / 8 TB/s
  divb_drift     max and rms of div B (diverB + bndrydiv) before and after the run
  cpu_baseline   the CPU oracle (C restatement, OpenMP) on a bounded sample
                 of the same workload (one step), with the host's CPU model
  C2             the same timed loop on 256x128x256 (200 steps)
  magnetosphere  the reference's own case (201x51x105 dipole, mask, inflow)
  ensemble       48 uneven Harris runs advanced together vs one after another
  adjoint        discrete adjoint of 500 steps at 128x64x128 (sqrt(N)
                 checkpointing): reverse sweep incl. recomputation / forward
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
# the ensemble runs each use a stream of their own: with 16 hardware queues
# (HIP's default is 4) more of them execute at once (read at HIP start-up)
os.environ["GPU_MAX_HW_QUEUES"] = os.environ.get("MHD_HW_QUEUES", "16")

HBM_PEAK = 8.0e12           # MI355X HBM3E, bytes/s
B_ALG_HALL = 6256           # algorithmic bytes per cell per RK4 step (Hall on)
B_ALG_NOHALL = 5104         # d_i = 0


def per_cell_bytes(case):
    """algorithmic HBM bytes per physical cell of one launch of each kernel
    family: every input field of every smoothing pass read once and every
    output written once (the per-sweep table of the reference's step), so a
    launch that fuses two passes counts both"""
    hall = case.hall
    rk2 = (256 + 320 + 320 + 192) // 4        # avrg pass 2 + RK update, mean of the 4 stages
    return {
        "curl": 8 * (6 + 3),                    # B 3 + B_i 3 in, J 3 out
        "avrg_radial_x2": 2 * 8 * (3 + 3),      # two avrgc / avrgE passes
        "avrg_radial_x1": 8 * (3 + 3),
        "avrgh_x2": 2 * 8 * (3 + 3),
        "avrgh_x1": 8 * (3 + 3),
        "avrgh_x3_addE": 3 * 8 * (3 + 3),       # the last three avrgh passes (+ E, as in avrgh_x1)
        "avrg_state_x2": 2 * 8 * (8 + 8),
        "avrg_state_x1": 8 * (8 + 8),
        "avrg_state_x2_rk": 8 * (8 + 8) + rk2,  # avrg pass 1 + (pass 2 fused with the RK update)
        "avrg_state_x1_rk": rk2,
        "rk_update": rk2,                       # stand-alone RK update (same fields)
        "rhs_flux_E_hall": 8 * (15 + 10) if hall else 8 * (15 + 7),   # U8,p_i,J3,xresd3 -> xdif1-4,E3(,xdh3)
        "rhs_curlE_pressure": 8 * ((18 if hall else 15) + 8),         # E3(,xdh3),U5,J3,xdif1-4 -> xdif8
        "avrgtt": 8 * ((8 + 5 + 8) + (8 + 8)) // 2,                   # mean of the two avrgtt passes
        "avrgtt_x2_eq": 8 * 2 * (8 + 5 + 8),    # both post-step passes on x - eq and x = xm + eq
        "post_add_eq": 8 * (8 + 5 + 8),
        "setdt": 8 * 8,
    }


def traffic_from_profiles(kernel: str, workload: str):
    """HBM bytes per launch of `kernel` measured with rocprofv3 --pmc
    (FETCH_SIZE, WRITE_SIZE scaled by the copy-kernel factors), when a
    committed profile summary for this workload exists"""
    path = os.path.join(ROOT, "profiles", f"traffic_{workload}.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        d = json.load(f)
    k = d.get("kernels", {}).get(kernel)
    return None if k is None else k["bytes_per_launch"]


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(config: str, steps: int, threads: int):
    """the oracle timed on `steps` setdt + stepon pairs of the same workload
    (its host-built arrays and residual capture untimed)"""
    from dasha_amd import harris
    from oracle import oracle
    case = harris.CASES[config]()
    oracle.build()
    oracle.set_threads(threads)
    t_setup = time.perf_counter()
    o = oracle.Oracle(case)
    o.init_residual()
    print(f"[bench] cpu baseline: {case.name} oracle set up in {time.perf_counter() - t_setup:.1f}s "
          f"on {threads} threads", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    tt = 0.0
    for _ in range(steps):
        dt = o.compute_dt()
        o.step(tt, dt)
        tt += dt
    el = time.perf_counter() - t0
    o.close()
    print(f"[bench] cpu baseline: {steps} step(s) in {el:.1f}s", file=sys.stderr, flush=True)
    return {"value": case.cells * steps / el, "unit": "cell-updates/s", "cores": threads,
            "kind": "port", "cpu_model": cpu_model(), "host_cpus": os.cpu_count(),
            "affinity_cpus": len(os.sched_getaffinity(0)), "seconds": el,
            "sample": f"{steps} RK4 step(s) incl. setdt of {case.name} ({case.cells} cells) "
                      f"by the C/OpenMP oracle on {threads} threads, {el:.2f} s"}


def time_config(name: str, steps: int, warmup: int):
    """the same timed loop on another configuration (no profiling)"""
    from dasha_amd import harris
    from dasha_amd.mhd import MHDStep
    case = harris.CASES[name](arrays=False)   # the device builds the fields
    g = MHDStep(case)
    g.init_residual()
    dt0 = g.compute_dt()
    tm, _, _ = g.advance(warmup, 0.0, dt0)
    tm, _, ms = g.advance(steps, tm, dt0)
    g.close()
    value = case.cells * steps / (ms / 1e3)
    b_alg = B_ALG_HALL if case.hall else B_ALG_NOHALL
    print(f"[bench] {case.name}: {ms / steps:.3f} ms/step", file=sys.stderr, flush=True)
    return {"workload": f"{case.name} Harris sheet, RK4 Hall MHD step (stepon+setdt)",
            "grid": [case.mx, case.ny, case.mz], "steps": steps, "warmup": warmup,
            "value": value, "unit": "cell-updates/s", "ms_per_step": ms / steps,
            "step_roofline_frac": value * b_alg / HBM_PEAK}


def time_magnetosphere(steps: int = 50, warmup: int = 3):
    """the reference's own case on its default grid (201 x 51 x 105, dipole,
    inner-sphere mask, driven x-left inflow, mhd03.for's constants, Hall
    off): the reference-structured kernels of magnetosphere mode"""
    from dasha_amd import magneto
    from dasha_amd.mhd import MHDStep
    case = magneto.case_default(face=magneto.synthetic_face((51, 105)))
    g = MHDStep(case)
    g.init_residual()
    dt0 = g.compute_dt()
    tm, _, _ = g.advance(warmup, 0.0, dt0)
    tm, _, ms = g.advance(steps, tm, dt0)
    g.close()
    value = case.cells * steps / (ms / 1e3)
    b_alg = B_ALG_HALL if case.hall else B_ALG_NOHALL
    print(f"[bench] {case.name}: {ms / steps:.3f} ms/step", file=sys.stderr, flush=True)
    return {"workload": f"{case.name} dipole magnetosphere, driven inflow, inner-sphere mask "
                        f"(RK4 MHD step, stepon+setdt)",
            "grid": [case.mx, case.ny, case.mz], "steps": steps, "warmup": warmup, "value": value,
            "unit": "cell-updates/s", "ms_per_step": ms / steps,
            "step_roofline_frac": value * b_alg / HBM_PEAK}


def time_ensemble(steps: int, nruns: int = 48):
    """the batched ensemble (48 Harris runs of uneven size and parameters):
    aggregate cell-updates/s of the runs advanced together against the
    same runs advanced one after another"""
    from dasha_amd import harris
    from dasha_amd.mhd import MHDEnsemble, MHDStep
    t0 = time.perf_counter()
    cases = harris.case_c4(nruns=nruns)
    runs = [MHDStep(c) for c in cases]
    for g in runs:
        g.init_residual()
    dt0 = [g.compute_dt() for g in runs]
    cells = sum(c.cells for c in cases)
    print(f"[bench] ensemble: {nruns} runs, {cells} cells, setup {time.perf_counter() - t0:.1f}s",
          file=sys.stderr, flush=True)
    # one warm-up step each, then the runs one after another
    times = [g.advance(1, 0.0, d)[0] for g, d in zip(runs, dt0)]
    seq_ms = 0.0
    for r, (g, d) in enumerate(zip(runs, dt0)):
        times[r], _, ms = g.advance(steps, times[r], d)
        seq_ms += ms
    ens = MHDEnsemble(runs)
    # the autotuner: 4 trial steps of the whole batch per tile shape
    shape, trials = ens.tune_shapes(4, times, dt0)
    print(f"[bench] ensemble: tuned shape {shape}", file=sys.stderr, flush=True)
    tt, _ = ens.advance(2, times, dt0)           # captures the step graphs
    tt, ens_ms = ens.advance(steps, tt, dt0)
    ens.close()
    for g in runs:
        g.close()
    print(f"[bench] ensemble: sequential {seq_ms:.1f} ms, together {ens_ms:.1f} ms", file=sys.stderr, flush=True)
    return {"workload": f"{nruns} Harris runs, 64x32x64 to 192x96x192, uneven eta, d_i, amplitude, "
                        f"thickness (RK4 Hall MHD step each)",
            "runs": nruns, "cells": cells, "steps": steps, "unit": "cell-updates/s",
            "value": cells * steps / (ens_ms / 1e3), "sequential_value": cells * steps / (seq_ms / 1e3),
            "speedup": seq_ms / ens_ms, "ms": ens_ms, "sequential_ms": seq_ms,
            "tuned_shape": shape, "tuner_trials": trials}


def time_adjoint(steps: int = 500):
    """gradient of the magnetic energy at T after `steps` steps of the
    128x64x128 stretched Harris sheet (eta 0.005, d_i 0.15) with respect to
    the initial dAy and eta: forward pass (states kept every ceil(sqrt(N))
    steps) and reverse sweep (segment replays + transposed kernels), HIP
    events on the library stream"""
    from dasha_amd import adjoint
    t0 = time.perf_counter()
    ac = adjoint.case_c5(steps=steps)
    dts = adjoint.dt_sequence(ac)
    val, gay, geta, (fwd_ms, rev_ms) = adjoint.objective(ac, dts, obj="magnetic_energy")
    print(f"[bench] adjoint: forward {fwd_ms:.1f} ms, reverse {rev_ms:.1f} ms "
          f"(wall {time.perf_counter() - t0:.1f}s)", file=sys.stderr, flush=True)
    # the same gradient with every step's stage states kept in HBM (44 fields
    # per step: about 196 GB at 128x64x128 x 500 steps), nothing recomputed
    from dasha_amd.mhd import MHDError
    try:
        v2, gay2, ge2, (fwd2, rev2) = adjoint.objective(ac, dts, obj="magnetic_energy", every=-1)
        same = bool(v2 == val and ge2 == geta and (gay2 == gay).all())
        store_all = {"forward_ms": fwd2, "reverse_ms": rev2, "ratio_reverse_to_forward": rev2 / fwd2,
                     "gradients_identical": same}
        print(f"[bench] adjoint, every step kept: forward {fwd2:.1f} ms, reverse {rev2:.1f} ms", file=sys.stderr,
              flush=True)
    except MHDError as e:
        store_all = {"error": str(e)}
    return {"workload": f"{ac.case.name} stretched Harris sheet, eta {ac.case.eta0}, d_i {ac.case.di}, "
                        f"{steps} RK4 steps, objective magnetic energy at T, controls dAy (mz x mx) and eta",
            "steps": steps, "checkpoint_every": adjoint.sqrt_checkpoints(steps),
            "forward_ms": fwd_ms, "reverse_ms": rev_ms,
            "forward_ms_per_step": fwd_ms / steps, "reverse_ms_per_step": rev_ms / steps,
            "ratio_reverse_to_forward": rev_ms / fwd_ms, "objective": val, "grad_eta": geta,
            "grad_ay_norm": float((gay ** 2).sum() ** 0.5), "every_step_kept": store_all}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None,
                    help="timed steps (default: the configuration's own count, 200 for C3)")
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--config", default="C3", choices=["C1", "C2", "C3"])
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-steps", type=int, default=1)
    ap.add_argument("--no-profile", action="store_true", help="skip the per-kernel event run")
    ap.add_argument("--copy-probe", action="store_true", help="also measure a plain device copy")
    ap.add_argument("--no-c2", action="store_true",
                    help="skip the 256x128x256 timing added to the line")
    ap.add_argument("--no-ensemble", action="store_true", help="skip the batched ensemble timing")
    ap.add_argument("--no-magnetosphere", action="store_true", help="skip the magnetosphere-mode timing")
    ap.add_argument("--ensemble-only", action="store_true", help="time the batched ensemble alone")
    ap.add_argument("--magnetosphere-only", action="store_true",
                    help="time the magnetosphere-mode step alone (the reference's own case)")
    ap.add_argument("--ensemble-runs", type=int, default=48)
    ap.add_argument("--no-adjoint", action="store_true", help="skip the discrete adjoint timing")
    ap.add_argument("--adjoint-steps", type=int, default=500)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported "
              f"(the step runs on one GPU's domain).", file=sys.stderr)
        sys.exit(2)

    if args.ensemble_only:
        print(json.dumps({"ensemble": time_ensemble(10, args.ensemble_runs)}), flush=True)
        return
    if args.magnetosphere_only:
        print(json.dumps({"magnetosphere": time_magnetosphere(args.steps or 50, args.warmup)}), flush=True)
        return

    from dasha_amd import harris
    from dasha_amd.mhd import MHDStep

    # mesh and parameters on the host, the 3-D fields built on the device
    t_setup = time.perf_counter()
    case = harris.CASES[args.config](arrays=False)
    if args.steps is None:
        args.steps = min(case.steps, 200)
    g = MHDStep(case)
    g.init_residual()
    dt0 = g.compute_dt()
    setup_s = time.perf_counter() - t_setup
    print(f"[bench] {case.name}: setup {setup_s:.1f}s, dt0 {dt0:.4e}", file=sys.stderr, flush=True)
    d_start = g.diagnostics()
    # warm-up steps (untimed)
    tm, _, _ = g.advance(args.warmup, 0.0, dt0)
    # timed region: K steps of setdt+stepon, HIP events on the library stream
    tm, dt_last, ms = g.advance(args.steps, tm, dt0)
    ms_per_step = ms / max(args.steps, 1)
    value = case.cells * args.steps / (ms / 1e3)
    print(f"[bench] timed {args.steps} steps: {ms:.1f} ms ({ms_per_step:.3f} ms/step)",
          file=sys.stderr, flush=True)

    d_end = g.diagnostics()
    divb = {"steps": args.warmup + args.steps, "max_start": d_start["divb_max"], "max_end": d_end["divb_max"],
            "rms_start": d_start["divb_rms"], "rms_end": d_end["divb_rms"]}
    # the state crossing PCIe (restart / output points only, never per step)
    t0 = time.perf_counter()
    xs = g.get_state()
    t1 = time.perf_counter()
    g.set_state(xs)
    t2 = time.perf_counter()
    host_io = {"state_bytes": xs.nbytes, "get_state_ms": (t1 - t0) * 1e3, "set_state_ms": (t2 - t1) * 1e3}
    del xs

    # per-kernel-family times (events around every launch) on a short run
    roof = None
    fam = {}
    if not args.no_profile:
        g.set_profiling(True)
        tm, _, _ = g.advance(min(args.steps, 5), tm, dt0)
        fam = g.kernel_times()
        g.set_profiling(False)
        pcb = per_cell_bytes(case)
        best = max(fam.items(), key=lambda kv: kv[1][0])
        name, (tot_ms, calls) = best
        if name in pcb and calls > 0:
            avg_s = tot_ms / calls / 1e3
            alg = pcb[name] * case.cells
            ach = alg / avg_s
            roof = {"bound": "hbm", "achieved": ach / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                    "frac": ach / HBM_PEAK, "traffic": traffic_from_profiles(name, args.config),
                    "kernel": name, "avg_launch_ms": tot_ms / calls, "alg_bytes_per_launch": alg}
    b_alg = B_ALG_HALL if case.hall else B_ALG_NOHALL
    out = {
        "metric": "FP64 cell-updates/s per RK4 step",
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"{case.name} Harris sheet, RK4 Hall MHD step (stepon+setdt)",
                   "grid": [case.mx, case.ny, case.mz], "stretched": args.config != "C1",
                   "hall": case.hall, "eta": case.eta0, "d_i": case.di, "cfl": case.cfl},
        "roofline": roof,
        "step_roofline": {"bytes_per_cell_step": b_alg,
                          "achieved_GBps": value * b_alg / 1e9,
                          "frac": value * b_alg / HBM_PEAK},
        "kernel_families_ms": {k: {"total_ms": v[0], "calls": v[1]} for k, v in fam.items()},
        "host_io": host_io,
        "setup_s": setup_s,
        "divb_drift": divb,
        "cpu_baseline": None,
    }
    if args.copy_probe:
        from dasha_amd.mhd import copy_rate
        out["copy_probe_GBps"] = copy_rate() / 1e9
    g.close()
    if args.config == "C3" and not args.no_c2:
        out["C2"] = time_config("C2", 200, 3)
    if not args.no_magnetosphere:
        out["magnetosphere"] = time_magnetosphere()
    if not args.no_ensemble:
        out["ensemble"] = time_ensemble(10, args.ensemble_runs)
    if not args.no_adjoint:
        out["adjoint"] = time_adjoint(args.adjoint_steps)
    if not args.no_cpu:
        # the host's CPU share for one GPU (OMP_NUM_THREADS on the GPU hosts)
        threads = min(int(os.environ.get("OMP_NUM_THREADS", "16") or 16), len(os.sched_getaffinity(0)))
        out["cpu_baseline"] = cpu_baseline(args.config, args.cpu_steps, threads)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
