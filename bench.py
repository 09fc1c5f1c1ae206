#!/usr/bin/env python
"""Erosion steps/s of the per-step Stokes solve on one MI355X.

A step is one ``rungekutta2`` call (ErosionPackage/src/TimeStepping.jl:
163-176): two GMRES density solves to a 1e-10 residual, two shear-stress
evaluations and -- with the pressure drop held fixed as the reference's
default ParamSet does (fixpdrop, DensityStress.jl:157-167) -- four target
evaluations at the x = +-2 slices.  The plain FP64 variant calls the C-ABI
entry points from the host stepper with the theta-L updates on the host, as
the Julia package does; the accelerated variant runs the whole step in the
library (erosion_rungekutta2).  Workload: 100 bodies x 256 points, wall 2048
points, seeded perturbed-circle geometry, pipe-flow boundary condition,
dt = 1e-3.

Prints ONE JSON line on stdout (everything else goes to stderr):
  value        erosion steps/s over the timed steps
  roofline     the dominant kernel's achieved rate against the MI355X peak,
               from HIP events recorded around every launch in the timed region
  cpu_baseline the reference's own Fortran (oracle/_ref, flang build of
               stokesSolver.f + the FMM) timed on this host on a bounded sample
"""
from __future__ import annotations

import argparse
import copy
import json
import math
import os
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "erosion steps/s at 1e-10 GMRES residual: learned+mixed vs plain FP64 on GPU"
PEAK_FP64_TFLOPS = 78.6  # MI355X FP64 vector (= FP64 matrix) peak
PEAK_FP32_TFLOPS = 157.3  # MI355X FP32 vector peak (packed)
PEAK_HBM_GBS = 8000.0
FF_FLOP_PER_PAIR = 20.0  # canonical flops per ordered source-target pair
STRESS_FLOP_PER_PAIR = 130.0  # per (target body point, source node) pair: 3 passes x 5 fields
# HBM bytes read per FP64 far-field launch on the 100-body workload, from a
# rocprofv3 FETCH_SIZE pass (x 2.00, the factor a 256 MiB device copy gives);
# the same for every launch shape.  The bytes written equal the per-split
# partial sums (a WRITE_SIZE pass reads exactly splits x points x 16 B), so
# the traffic of the shape in use is this plus that.
FF_FETCH_BYTES = {"bodies100_n256": 10.5e6}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def _reference_timings(n, m, no, msub, xx, yy):
    """Time the reference's Fortran (run in a child process, see cpu_baseline)."""
    import numpy as np
    from oracle import reference as R

    N = 2 * no + 2 * n * m + 3 * m
    eta = np.random.default_rng(3).standard_normal(N)
    R.matvec(n, m, no, xx, yy, eta, ibary=1, ifmm=1)  # warm-up (FMM tables, threads)
    t0 = time.perf_counter()
    reps = 2
    for _ in range(reps):
        R.matvec(n, m, no, xx, yy, eta, ibary=1, ifmm=1)
    t_mv = (time.perf_counter() - t0) / reps
    den = np.random.default_rng(4).standard_normal(2 * no + 2 * n * msub + 3 * msub) * 0.1
    t0 = time.perf_counter()
    R.shear_stress(n, msub, no, xx[: n * msub], yy[: n * msub], den)
    return t_mv, reps, time.perf_counter() - t0


def cpu_baseline(w, xx, yy, iters_per_solve):
    """Reference CPU rate on a bounded sample, scaled to steps/s.

    The reference's Fortran prints through its own buffered unit and writes an
    FMM side file, so it runs in a child process with its output discarded and
    a scratch working directory; the child writes its timings to a file."""
    import subprocess

    import numpy as np
    from oracle import reference as R

    if not R.available():
        return None
    n, m, no = w.npts, w.nbodies, w.nouter
    msub = min(10, m)
    with tempfile.TemporaryDirectory() as td:
        np.save(os.path.join(td, "xx.npy"), xx)
        np.save(os.path.join(td, "yy.npy"), yy)
        code = (f"import sys, json, numpy as np; sys.path.insert(0, {ROOT!r}); import bench; "
                f"xx = np.load('xx.npy'); yy = np.load('yy.npy'); "
                f"r = bench._reference_timings({n}, {m}, {no}, {msub}, xx, yy); "
                f"open('t.json', 'w').write(json.dumps(r))")
        subprocess.run([sys.executable, "-c", code], cwd=td, check=True,
                       stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)
        with open(os.path.join(td, "t.json")) as f:
            t_mv, reps, t_st_sub = json.load(f)

    # reference stress cost in complex multiply-adds: wall->bodies, wall
    # boundary values per body, body->body pairs (stokesSolver.f:2707-2925)
    def cost(mm):
        return 5 * mm * no * no + 15 * mm * n * no + 20 * mm * (mm - 1) * n * n

    t_st = t_st_sub * cost(m) / cost(msub)
    t_step = 2 * iters_per_solve * t_mv + 2 * t_st
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {
        "value": 1.0 / t_step,
        "unit": "erosion steps/s",
        "cores": threads,
        "kind": "reference",
        "sample": (f"flang build of the reference (stokesSolver.f + FMM, OpenMP {threads} threads): "
                   f"{reps} matvec_DLP_fmmbary calls at the full workload ({t_mv:.3f} s each) and one "
                   f"computeShearStress on {msub} of the {m} bodies ({t_st_sub:.2f} s, scaled by its "
                   f"pair count to {t_st:.1f} s); step = 2 solves x {iters_per_solve:.0f} GMRES "
                   f"iterations (the plain GPU path's count) + 2 stresses; DGMRES "
                   f"orthogonalisation not counted"),
    }


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="bodies100_n256")
    ap.add_argument("--tol", type=float, default=1e-10)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--skip-plain", action="store_true",
                    help="time only the accelerated variant (for profiling runs)")
    ap.add_argument("--host-stepper", action="store_true",
                    help="accelerated variant: theta-L updates on the host between solver "
                         "calls, as the plain variant (default: the whole step in the library, "
                         "erosion_rungekutta2)")
    ap.add_argument("--body-modes", type=int, default=None,
                    help="accelerated variant: Fourier modes per body in the preconditioner's "
                         "coarse space (default: the library's)")
    ap.add_argument("--graphs", action="store_true",
                    help="replay GMRES iteration chunks as HIP graphs in the accelerated "
                         "variant; kernel timing then runs on one extra untimed step")
    args = ap.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        log("bench.py: the Stokes solve runs on one GPU only; multi-GPU runs are not "
            "part of this implementation (the solve would shard by target rows).")
        return 2

    import numpy as np

    from erosion_amd import stokes
    from erosion_amd.geometry import WORKLOADS, workload_geometry
    from erosion_amd.theta_len import ParamSet, getnxy
    from erosion_amd.density_stress import getstress_
    from erosion_amd.learned import WarmStart
    from erosion_amd.time_stepping import rungekutta2
    from erosion_amd.device_stepper import rungekutta2_device

    try:
        import torch

        sync = torch.cuda.synchronize if torch.cuda.is_available() else (lambda: None)
    except Exception:  # torch is plumbing only
        def sync():
            return None

    w = WORKLOADS[args.workload]
    params = ParamSet(npts=w.npts, nouter=w.nouter, ibary=1, ifmm=1, ibc=w.ibc, dt=w.dt,
                      maxl=8000, fixpdrop=True)
    stokes.set_gmres_tol(args.tol)
    thld0 = workload_geometry(w)
    _m, xx0, yy0 = getnxy(thld0)

    def configure(accelerated):
        if accelerated and args.body_modes is not None:
            stokes.set_preconditioner(1, args.body_modes)
        else:
            stokes.set_preconditioner(1 if accelerated else 0)
        stokes.set_precision(stokes.PRECISION_MIXED if accelerated else stokes.PRECISION_FP64)
        stokes.set_graphs(bool(accelerated and args.graphs))

    def run_variant(label, accelerated):
        """Warm-up steps, then exactly args.steps timed steps from the initial state."""
        configure(accelerated)
        guess = WarmStart(w.nouter, ibc=w.ibc) if accelerated else None
        if accelerated and not args.host_stepper:
            stokes.stepper_reset()
            net = guess.net if guess is not None else None

            def rungekutta2(th, params, guess=None, timings=None):  # noqa: F811
                return rungekutta2_device(th, params, warm=2, timings=timings, guess=net)
        else:
            from erosion_amd.time_stepping import rungekutta2
        th = copy.deepcopy(thld0)
        for i in range(args.warmup):
            th = rungekutta2(th, params, guess=guess, timings={})
            log(f"[{label}] warmup step {i + 1}/{args.warmup}: bodies={len(th.thlenvec)}")
        sync()
        # kernel timing bypasses graph replay: with graphs it runs on one more step
        graphs = accelerated and args.graphs
        stokes.kernel_timing(not graphs)
        timings, solve_info = {}, []
        t0 = time.perf_counter()
        for i in range(args.steps):
            ts = time.perf_counter()
            th = rungekutta2(th, params, guess=guess, timings=timings)
            solve_info.append(stokes.last_solve_info())
            log(f"[{label}] step {i + 1}/{args.steps}: {time.perf_counter() - ts:.3f} s, "
                f"bodies={len(th.thlenvec)}, gmres iters so far {timings.get('gmres_iters', 0)}, "
                f"last true relres {solve_info[-1]['true_relres']:.2e}")
        sync()
        elapsed = time.perf_counter() - t0
        if graphs:
            log(f"[{label}] graph replay: {stokes.graph_stats()}")
            stokes.kernel_timing(True)
            rungekutta2(copy.deepcopy(th), params, guess=copy.deepcopy(guess), timings={})
        kt = stokes.kernel_timers()
        stokes.kernel_timing(False)
        configure(False)
        log(f"[{label}] host-side timings: " +
            ", ".join(f"{k} {v:.3f}" if isinstance(v, float) else f"{k} {v}"
                      for k, v in timings.items()))
        return {"elapsed": elapsed, "timings": timings, "kt": kt, "solve_info": solve_info,
                "state": th, "graphs": graphs}

    def first_stress(accelerated):
        """Shear stress of the initial state through one variant's solve."""
        configure(accelerated)
        atau = getstress_(copy.deepcopy(thld0), params)
        configure(False)
        return np.asarray(atau)

    stress_err = theta_err = None
    if not args.skip_plain:
        tau_plain, tau_acc = first_stress(False), first_stress(True)
        stress_err = float(np.max(np.abs(tau_acc - tau_plain)) / np.max(np.abs(tau_plain)))
        log(f"shear stress of the initial state, accelerated vs plain FP64: {stress_err:.2e} relative")

    plain = None if args.skip_plain else run_variant("plain-fp64", False)
    acc = run_variant("accelerated", True)
    if plain is not None and len(acc["state"].thlenvec) == len(plain["state"].thlenvec):
        theta_err = max(float(np.max(np.abs(a.theta - b.theta)))
                        for a, b in zip(acc["state"].thlenvec, plain["state"].thlenvec))

    def roofline_of_run(r):
        """The run's canonical work at the chip's peaks: far-field pairs at 20
        FLOP (FP64 launches at the FP64 peak, FP32 ones at the FP32 peak) and
        stress pairs at 130 FLOP (FP64).  Returns (ideal seconds for the timed
        steps, ideal seconds per operator application)."""
        k = r["kt"]
        scale = args.steps if r["graphs"] else 1
        t64 = FF_FLOP_PER_PAIR * k["farfield"]["work"] / (PEAK_FP64_TFLOPS * 1e12)
        t32 = FF_FLOP_PER_PAIR * k["farfield32"]["work"] / (PEAK_FP32_TFLOPS * 1e12)
        tst = STRESS_FLOP_PER_PAIR * k["stress_eval"]["work"] / (PEAK_FP64_TFLOPS * 1e12)
        nmv = k["farfield"]["launches"] + k["farfield32"]["launches"]
        return scale * (t64 + t32 + tst), (t64 + t32) / max(1, nmv)

    def summary(r):
        solves = max(1, r["timings"].get("solves", 1))
        ideal_s, ideal_mv = roofline_of_run(r)
        mvps = ((r["kt"]["farfield"]["launches"] + r["kt"]["farfield32"]["launches"])
                * (args.steps if r["graphs"] else 1) / r["elapsed"])
        return {"steps_per_s": args.steps / r["elapsed"],
                # the same steps with every far-field and stress FLOP at peak
                "steps_per_s_at_peak": args.steps / ideal_s if ideal_s > 0 else None,
                "steps_roofline_frac": ideal_s / r["elapsed"],
                "matvecs_per_s_at_peak": 1.0 / ideal_mv if ideal_mv > 0 else None,
                "matvecs_roofline_frac": mvps * ideal_mv,
                "ms_per_step": 1e3 * r["elapsed"] / args.steps,
                "gmres_iters_per_solve": r["timings"].get("gmres_iters", 0) / solves,
                "true_relres_max": max(si["true_relres"] for si in r["solve_info"]),
                "fp32_iters_last_solve": r["solve_info"][-1]["fp32_iters"],
                # operator applications (FP64 and FP32 far field); with graphs
                # the kernel counts come from one extra step, scaled
                "matvecs_per_s": mvps}

    sp, sa = (summary(plain) if plain else None), summary(acc)
    elapsed, kt = acc["elapsed"], acc["kt"]
    steps_per_s = sa["steps_per_s"]
    iters_per_solve = sp["gmres_iters_per_solve"] if sp else None
    ff = kt["farfield"]
    ff_splits = acc["solve_info"][-1]["ff_splits"] if acc["solve_info"] else 0
    ff_ms = ff["ms"] / max(1, ff["launches"])
    ff_pairs = ff["work"] / max(1, ff["launches"])
    ff_tflops = FF_FLOP_PER_PAIR * ff_pairs / (ff_ms * 1e-3) / 1e12
    f32 = kt["farfield32"]
    f32_ms = f32["ms"] / f32["launches"] if f32["launches"] else None
    f32_tflops = (FF_FLOP_PER_PAIR * f32["work"] / f32["launches"] / (f32_ms * 1e-3) / 1e12
                  if f32["launches"] else None)
    gs = kt["gram_schmidt"]
    gs_gbs = gs["work"] / (gs["ms"] * 1e-3) / 1e9 if gs["ms"] > 0 else 0.0
    se = kt["stress_eval"]
    se_tflops = (STRESS_FLOP_PER_PAIR * se["work"] / (se["ms"] * 1e-3) / 1e12
                 if se["ms"] > 0 else None)
    nr = kt["near"]
    near_gbs = nr["work"] / (nr["ms"] * 1e-3) / 1e9 if nr["ms"] > 0 else None
    share = {k: kt[k]["ms"] / (elapsed * 1e3) for k in kt}
    log(f"plain-fp64: {sp}")
    log(f"accelerated: {sa}")
    log(f"kernel time shares of the accelerated timed region: " +
        ", ".join(f"{k} {100 * v:.1f}%" for k, v in share.items()))
    if f32_ms:
        log(f"FP32 far field: {f32_ms:.3f} ms/launch, {f32_tflops:.2f} TFLOP/s over "
            f"{f32['launches']} launches")
    log(f"far field: {ff_ms:.3f} ms/launch, {ff_tflops:.2f} TFLOP/s; Gram-Schmidt {gs_gbs:.0f} GB/s; "
        f"stress eval {se['ms'] / max(1, se['launches']):.2f} ms/launch"
        + (f", {se_tflops:.2f} TFLOP/s" if se_tflops else "")
        + (f"; near-pair gather {near_gbs:.0f} GB/s over {nr['launches']} passes"
           if near_gbs else "; no near pairs"))

    cpu = None
    if not args.no_cpu_baseline and iters_per_solve:
        try:
            cpu = cpu_baseline(w, xx0, yy0, iters_per_solve)
        except Exception as e:  # the baseline is reported, never required
            log(f"cpu baseline unavailable: {e}")

    out = {
        "metric": METRIC,
        "value": steps_per_s,
        "unit": "erosion steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * elapsed / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic: seeded perturbed-circle porous geometry, theta-L state",
        "config": {
            "workload": w.name, "bodies": w.nbodies, "points_per_body": w.npts,
            "wall_points": w.nouter, "gmres_tol": args.tol,
            "variant": "accelerated: per-curve two-level right preconditioner, FP32 far field "
                       "once the residual is below 1e4 tol (FP64 true residual), GMRES guess "
                       "extrapolated from the last two densities or, where none fits, the "
                       "learned network; "
                       + ("theta-L updates on the host" if args.host_stepper else
                          "the whole RK2 step in the library (theta-L updates on the GPU)"),
            "dt": w.dt, "fixpdrop": True, "ibc": w.ibc,
            "hip_graphs": bool(args.graphs),
            "device_stepper": not args.host_stepper,
            "accelerated": sa,
            "plain_fp64": sp,
            "speedup_vs_plain_fp64": sa["steps_per_s"] / sp["steps_per_s"] if sp else None,
            "stress_relerr_vs_plain_fp64": stress_err,
            "theta_maxdiff_after_steps": theta_err,
            "matvecs_per_s": sa["matvecs_per_s"],
            "farfield_ms": ff_ms,
            "farfield32_ms": f32_ms,
            "farfield32_TFLOPs": f32_tflops,
            "farfield32_launches": f32["launches"],
            "gram_schmidt_GBps": gs_gbs,
            "stress_eval_TFLOPs": se_tflops,
            "near_gather_GBps": near_gbs,
            "near_gather_launches": nr["launches"],
            "kernel_share": share,
        },
        "roofline": {
            "bound": "mfma",
            "pipe": "VALU FP64 (the FP64 vector peak equals the FP64 matrix peak on MI355X)",
            "kernel": "k_farfield (all-pairs double layer)",
            "achieved": ff_tflops,
            "peak": PEAK_FP64_TFLOPS,
            "unit": "TFLOP/s",
            "frac": ff_tflops / PEAK_FP64_TFLOPS,
            "traffic": (FF_FETCH_BYTES[w.name] + 16.0 * ff_splits * (w.nbodies * w.npts + w.nouter)
                        if w.name in FF_FETCH_BYTES and ff_splits else None),
            "traffic_unit": "bytes per launch",
        },
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
