#!/usr/bin/env python
"""Erosion steps/s of the per-step Stokes solve on one MI355X.

A step is one ``rungekutta2`` call (ErosionPackage/src/TimeStepping.jl:
189-202): two GMRES density solves to a 1e-10 true relative residual, two
shear-stress evaluations and -- with the pressure drop held fixed as the
reference's default ParamSet does (fixpdrop, DensityStress.jl:157-167) -- four
target evaluations at the x = +-2 slices.

Default workload: 400 bodies x 256 points (wall 4096), seeded, eroded until
50% of the total body area is lost (386 bodies, 102,912 boundary points; the
committed state erosion_amd/data/bodies400_n256_eroded50.npz), pipe-flow
boundary condition, dt = 1e-3.  Two variants from that same state:

  accelerated  the headline ``value``: the whole step in the library
               (erosion_rungekutta2), per-curve two-level preconditioner,
               FP32 far field once the residual is small (FP64 true residual
               with compensated sums), GMRES guesses extrapolated from the
               last two solutions in the preconditioned variable, and the
               learned network's guess where no solution is kept.
  plain_fp64   the reference's iteration: unpreconditioned FP64 GMRES from
               zeros through the C-ABI entry points, theta-L updates on the
               host as the Julia package does; timed over ``--plain-steps``
               steps (one step is about 100 s here).

``config.ablation`` adds short runs that switch the accelerations on one at a
time.  Every solve's true residual is checked; ``tol_misses`` counts solves
that ended above the tolerance.

Prints ONE JSON line on stdout (everything else goes to stderr):
  value        erosion steps/s of the accelerated variant over the timed steps
  roofline     the dominant kernel's achieved rate against the MI355X peak,
               from HIP events recorded around every launch in the timed region
  cpu_baseline the reference's own Fortran (oracle/_ref, flang build of
               stokesSolver.f + the FMM) timed on this host on a bounded sample
"""
from __future__ import annotations

import argparse
import copy
import json
import os
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "erosion steps/s at 1e-10 GMRES residual: learned+mixed vs plain FP64 on GPU"
PEAK_FP64_TFLOPS = 78.6  # MI355X FP64 vector peak
PEAK_FP32_TFLOPS = 157.3  # MI355X FP32 vector peak (packed)
PEAK_HBM_GBS = 8000.0
FF_FLOP_PER_PAIR = 20.0  # canonical flops per ordered source-target pair
STRESS_FLOP_PER_PAIR = 130.0  # per (target body point, source node) pair: 3 passes x 5 fields
# HBM bytes read per FP64 far-field launch, from a rocprofv3 FETCH_SIZE pass
# (x 2.00, the factor a 256 MiB device copy gives); the same for every launch
# shape.  The bytes written equal the per-split partial sums (a WRITE_SIZE
# pass reads exactly splits x points x 16 B), so the traffic of the shape in
# use is this plus that.
FF_FETCH_BYTES = {"bodies100_n256": 10.5e6, "bodies400_n256_eroded50": 45.9e6}
# HBM bytes of the cluster far field's direct sums per operator application
# (k_fmm_direct for body targets and for wall-segment targets: FETCH_SIZE x
# 2.00, the factor a 256 MiB device copy gives, plus WRITE_SIZE), from
# rocprofv3 counter passes on the workload's starting state (386 bodies)
FMM_DIRECT_BYTES = {"bodies400_n256_eroded50": 56.93e6}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def _reference_timings(n, m, no, xx, yy, m_mv, m_st):
    """Time the reference's Fortran on sub-geometries (run in a child process,
    see cpu_baseline): matvec_DLP_fmmbary on the first m_mv bodies and
    computeShearStress on the first m for m in m_st."""
    import numpy as np
    from oracle import reference as R

    def sub(k):
        return xx[: n * k], yy[: n * k]

    N = 2 * no + 2 * n * m_mv + 3 * m_mv
    eta = np.random.default_rng(3).standard_normal(N)
    R.matvec(n, m_mv, no, *sub(m_mv), eta, ibary=1, ifmm=1)  # warm-up (FMM tables, threads)
    t0 = time.perf_counter()
    reps = 2
    for _ in range(reps):
        R.matvec(n, m_mv, no, *sub(m_mv), eta, ibary=1, ifmm=1)
    t_mv = (time.perf_counter() - t0) / reps
    t_st = []
    for k in m_st:
        den = np.random.default_rng(4).standard_normal(2 * no + 2 * n * k + 3 * k) * 0.1
        t0 = time.perf_counter()
        R.shear_stress(n, k, no, *sub(k), den)
        t_st.append(time.perf_counter() - t0)
    # DORTH's inner loop: one dot and one axpy over N_unk doubles
    Nf = 2 * no + 2 * n * m + 3 * m
    a, b = np.random.default_rng(5).standard_normal((2, Nf))
    t0 = time.perf_counter()
    for _ in range(50):
        h = float(a @ b)
        b -= h * 1e-9 * a
    t_dot_axpy = (time.perf_counter() - t0) / 50
    return t_mv, reps, t_st, t_dot_axpy


def cpu_baseline(w, n, m, no, xx, yy, iters_per_solve):
    """Reference CPU rate on a bounded sample, scaled to steps/s.

    The reference holds at most 2^17 unknowns (stokesSolver.f:11), so its
    FMM matvec is timed on the first bodies of the state that fit and scaled
    by point count (the FMM's cost is linear in the points); its shear stress
    is timed on 2, 4 and 8 bodies and extrapolated by the quadratic cost
    model fitted to them; DORTH's modified Gram-Schmidt is counted at the
    GMRES iteration count given.  The Fortran prints through its own buffered
    unit and writes an FMM side file, so it runs in a child process with its
    output discarded and a scratch working directory."""
    import subprocess

    import numpy as np
    from oracle import reference as R

    if not R.available():
        return None
    m_mv = min(m, (R.NMAX - 2 * no) // (2 * n + 3))
    m_st = [k for k in (2, 4, 8) if k <= m]
    with tempfile.TemporaryDirectory() as td:
        np.save(os.path.join(td, "xx.npy"), xx)
        np.save(os.path.join(td, "yy.npy"), yy)
        code = (f"import sys, json, numpy as np; sys.path.insert(0, {ROOT!r}); import bench; "
                f"xx = np.load('xx.npy'); yy = np.load('yy.npy'); "
                f"r = bench._reference_timings({n}, {m}, {no}, xx, yy, {m_mv}, {m_st}); "
                f"open('t.json', 'w').write(json.dumps(r))")
        # matvec_DLP_fmmbary keeps its work arrays on the stack (automatic
        # arrays, stokesSolver.f:586-618): give the child an unlimited stack
        import resource

        def big_stack():
            resource.setrlimit(resource.RLIMIT_STACK,
                               (resource.RLIM_INFINITY, resource.RLIM_INFINITY))

        env = dict(os.environ, OMP_STACKSIZE="256M")
        subprocess.run([sys.executable, "-c", code], cwd=td, check=True, env=env,
                       preexec_fn=big_stack, stdout=subprocess.DEVNULL,
                       stderr=subprocess.DEVNULL)
        with open(os.path.join(td, "t.json")) as f:
            t_mv_sub, reps, t_st, t_da = json.load(f)
    npts_sub, npts = no + n * m_mv, no + n * m
    t_mv = t_mv_sub * npts / npts_sub
    # stress cost: wall -> bodies and body -> body pairs (stokesSolver.f:2707-2925),
    # t(k) = a k + c k (k - 1), least squares over the timed body counts
    ks = np.array(m_st, dtype=float)
    A = np.stack([ks, ks * (ks - 1)], axis=1)
    coef, *_ = np.linalg.lstsq(A, np.array(t_st), rcond=None)
    t_stress = float(coef[0] * m + coef[1] * m * (m - 1))
    fit_err = float(np.max(np.abs(A @ coef - np.array(t_st)) / np.array(t_st)))
    k = iters_per_solve
    t_orth = t_da * k * (k + 1) / 2.0  # DORTH: k dot/axpy pairs at iteration k
    t_solve = k * t_mv + t_orth
    t_step = 2 * t_solve + 2 * t_stress
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {
        "value": 1.0 / t_step,
        "unit": "erosion steps/s",
        "cores": threads,
        "kind": "reference",
        "sample": (f"flang build of the reference (stokesSolver.f + FMM); the FMM matvec runs "
                   f"OpenMP on {threads} threads, the rest of the reference is serial. "
                   f"matvec_DLP_fmmbary on the first {m_mv} of {m} bodies (the 2^17-unknown "
                   f"limit): {t_mv_sub:.3f} s, scaled by points to {t_mv:.3f} s; "
                   f"computeShearStress on {m_st} bodies: "
                   f"{', '.join(f'{t:.2f}' for t in t_st)} s, fit a k + c k(k-1) "
                   f"(max misfit {100 * fit_err:.0f}%) extrapolated to {t_stress:.0f} s; DORTH "
                   f"modified Gram-Schmidt {t_orth:.0f} s per solve (k(k+1)/2 dot+axpy pairs of "
                   f"{2 * no + (2 * n + 3) * m} doubles, {1e3 * t_da:.2f} ms each, 1 thread); "
                   f"step = 2 solves x {k:.0f} GMRES iterations (the plain GPU path's count) "
                   f"+ 2 stresses = {t_step:.0f} s"),
    }


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="bodies400_n256_eroded50")
    ap.add_argument("--tol", type=float, default=1e-10)
    ap.add_argument("--plain-steps", type=int, default=1,
                    help="steps timed in the plain FP64 variant (no warm-up)")
    ap.add_argument("--ablation-steps", type=int, default=-1,
                    help="timed steps per ablation row; -1 (default): the headline's own "
                         "warm-up and timed steps from the same starting state; 0 skips them")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--skip-plain", action="store_true",
                    help="time only the accelerated variant (for profiling runs)")
    ap.add_argument("--body-modes", type=int, default=None,
                    help="accelerated variant: Fourier modes per body in the preconditioner's "
                         "coarse space (default: the library's)")
    ap.add_argument("--timing-period", type=int, default=32,
                    help="time the kernels of one GMRES iteration in this many (each HIP event "
                         "record costs the stream a few microseconds; 1 times every launch)")
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
    from erosion_amd.device_stepper import rungekutta2_device
    from erosion_amd.geometry import WORKLOADS, workload_geometry
    from erosion_amd.learned import WEIGHTS, LearnedGuess
    from erosion_amd.theta_len import ParamSet, getnxy
    from erosion_amd.time_stepping import rungekutta2

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
    m0, xx0, yy0 = getnxy(thld0)
    n, no = w.npts, w.nouter
    net = LearnedGuess(WEIGHTS, nouter=no, ibc=w.ibc) if os.path.exists(WEIGHTS) else None

    def configure(pc, mixed):
        if pc and args.body_modes is not None:
            stokes.set_preconditioner(1, args.body_modes)
        else:
            stokes.set_preconditioner(1 if pc else 0)
        stokes.set_precision(stokes.PRECISION_MIXED if mixed else stokes.PRECISION_FP64)

    def tol_misses():
        i = stokes.last_solve_info()
        return np.array([i["tol_misses_total"], i["tol_misses_at_floor_total"]])

    def run_device(label, warmup, steps, pc=True, mixed=True, warm=2, learned=True,
                   graphs=False, timed_kernels=True):
        """Warm-up steps, then `steps` timed steps of the device stepper from
        the initial state."""
        configure(pc, mixed)
        stokes.set_graphs(graphs)
        stokes.stepper_reset()
        guess = net if learned else None
        th = copy.deepcopy(thld0)
        timings = {}
        for i in range(warmup):
            th = rungekutta2_device(th, params, warm=warm, timings=timings, guess=guess)
            log(f"[{label}] warmup step {i + 1}/{warmup}: bodies={len(th.thlenvec)}")
        sync()
        stokes.kernel_timing(timed_kernels and not graphs, period=args.timing_period)
        misses0 = tol_misses()
        timings, solve_info, relres = {}, [], []
        t0 = time.perf_counter()
        for i in range(steps):
            ts = time.perf_counter()
            th = rungekutta2_device(th, params, warm=warm, timings=timings, guess=guess)
            relres.append(timings["last_info"]["true_relres_max"])
            solve_info.append(stokes.last_solve_info())
            log(f"[{label}] step {i + 1}/{steps}: {time.perf_counter() - ts:.3f} s, "
                f"bodies={len(th.thlenvec)}, gmres iters so far {timings.get('gmres_iters', 0)}, "
                f"largest true relres {relres[-1]:.2e}")
        sync()
        elapsed = time.perf_counter() - t0
        misses = tol_misses() - misses0
        miss_rec = None
        if misses[0]:
            miss_rec = stokes.last_solve_info()["latest_miss_record"]
            log(f"[{label}] {misses[0]} solve(s) above the tolerance ({misses[1]} at the FP64 "
                f"floor of their density); the latest: {miss_rec}")
        if graphs and timed_kernels:
            log(f"[{label}] graph replay: {stokes.graph_stats()}")
            stokes.set_graphs(False)
            stokes.kernel_timing(True)
            rungekutta2_device(copy.deepcopy(th), params, warm=warm, guess=guess)
        kt = stokes.kernel_timers()
        stokes.kernel_timing(False)
        stokes.set_graphs(False)
        configure(False, False)
        return {"elapsed": elapsed, "steps": steps, "timings": timings, "kt": kt,
                "solve_info": solve_info, "true_relres_max": max(relres), "state": th,
                "graphs": graphs, "tol_misses": misses, "miss_record": miss_rec}

    def run_plain(label, steps):
        """The reference's iteration through the C-ABI, theta-L on the host."""
        configure(False, False)
        stokes.set_guess_space(stokes.GUESS_DENSITY)
        th_init = copy.deepcopy(thld0)
        th = th_init
        stokes.kernel_timing(True)
        misses0 = tol_misses()
        timings, solve_info = {}, []
        relres = []
        t0 = time.perf_counter()
        for i in range(steps):
            ts = time.perf_counter()
            th = rungekutta2(th, params, timings=timings)
            solve_info.append(stokes.last_solve_info())
            relres.append(solve_info[-1]["true_relres"])
            log(f"[{label}] step {i + 1}/{steps}: {time.perf_counter() - ts:.3f} s, "
                f"bodies={len(th.thlenvec)}, gmres iters so far {timings.get('gmres_iters', 0)}, "
                f"last true relres {relres[-1]:.2e}")
        sync()
        elapsed = time.perf_counter() - t0
        kt = stokes.kernel_timers()
        stokes.kernel_timing(False)
        return {"elapsed": elapsed, "steps": steps, "timings": timings, "kt": kt,
                "solve_info": solve_info, "true_relres_max": timings["true_relres_max"],
                "state": th, "graphs": False, "tol_misses": tol_misses() - misses0,
                "miss_record": None,
                "first_density": np.asarray(th_init.density)}

    def roofline_of_run(r):
        """The run's canonical work at the chip's peaks: far-field pairs at 20
        FLOP (FP64 launches at the FP64 peak, FP32 ones at the FP32 peak) and
        stress pairs at 130 FLOP (FP64).  Returns (ideal seconds for the timed
        steps, ideal seconds per operator application)."""
        k = r["kt"]
        scale = r["steps"] if r["graphs"] else 1
        # pairs summed directly: the direct far field, and the direct part
        # of the cluster far field (its expansions are not counted)
        pairs64 = k["farfield"]["work"] + k["fmm_direct"]["work"]
        t64 = FF_FLOP_PER_PAIR * pairs64 / (PEAK_FP64_TFLOPS * 1e12)
        t32 = FF_FLOP_PER_PAIR * k["farfield32"]["work"] / (PEAK_FP32_TFLOPS * 1e12)
        tst = STRESS_FLOP_PER_PAIR * k["stress_eval"]["work"] / (PEAK_FP64_TFLOPS * 1e12)
        return scale * (t64 + t32 + tst), (t64 + t32) / max(1, applications(k))

    def applications(k):
        """Operator applications: direct far-field launches (FP64, FP32) and
        body-cluster far-field applications."""
        return k["farfield"]["launches"] + k["farfield32"]["launches"] + k["fmm"]["launches"]

    def summary(r):
        solves = max(1, r["timings"].get("solves", 1))
        ideal_s, ideal_mv = roofline_of_run(r)
        mvps = applications(r["kt"]) * (r["steps"] if r["graphs"] else 1) / r["elapsed"]
        return {"steps_per_s": r["steps"] / r["elapsed"],
                "steps_timed": r["steps"],
                # the same steps with every directly summed far-field pair
                # and every stress pair at peak
                "steps_per_s_at_peak": r["steps"] / ideal_s if ideal_s > 0 else None,
                "steps_roofline_frac": ideal_s / r["elapsed"],
                "matvecs_per_s_at_peak": 1.0 / ideal_mv if ideal_mv > 0 else None,
                "matvecs_roofline_frac": mvps * ideal_mv,
                "ms_per_step": 1e3 * r["elapsed"] / r["steps"],
                "gmres_iters_per_solve": r["timings"].get("gmres_iters", 0) / solves,
                "true_relres_max": r["true_relres_max"],
                "solves_above_tol": int(r["tol_misses"][0]),
                # of those, solves whose density's FP64 floor (the residual a
                # half-ulp perturbation of it leaves) is at least half the
                # residual reached, and the latest miss's record
                "solves_above_tol_at_fp64_floor": int(r["tol_misses"][1]),
                "latest_miss": r["miss_record"],
                # operator applications (FP64 and FP32 far field); with graphs
                # the kernel counts come from one extra step, scaled
                "matvecs_per_s": mvps}

    # --- the headline: accelerated variant ----------------------------------
    acc = run_device("accelerated", args.warmup, args.steps, graphs=args.graphs)
    sa = summary(acc)
    sa["learned_guesses_used"] = int(acc["timings"].get("given_guesses", 0))

    # --- the plain FP64 variant, on its own sample ---------------------------
    plain = None if args.skip_plain else run_plain("plain-fp64", args.plain_steps)
    sp = summary(plain) if plain else None

    # --- the first solve of the state from zeros and from the learned guess,
    # and the stress of the accelerated solve against the plain solve's -------
    first, stress_err = {}, None
    configure(True, True)
    stokes.set_guess_space(stokes.GUESS_PRECONDITIONED)
    try:
        den_acc, it = stokes.stokessolver(n, m0, no, 1, 1, w.ibc, 8000, xx0, yy0)
        info = stokes.last_solve_info()
        first["zeros"] = {"iterations": it, "seconds": info["seconds"],
                          "true_relres": info["true_relres"]}
        if net is not None:
            g = net.predict(thld0, no, scale=True)
            _, it = stokes.stokessolver(n, m0, no, 1, 1, w.ibc, 8000, xx0, yy0, den=g)
            info = stokes.last_solve_info()
            first["learned"] = {"iterations": it, "seconds": info["seconds"],
                                "r0_relres": info["r0_relres"],
                                "true_relres": info["true_relres"]}
    finally:
        stokes.set_guess_space(stokes.GUESS_DENSITY)
        configure(False, False)
    log(f"first solve of the state, accelerated: {first}")
    if plain is not None and len(plain["first_density"]) == len(den_acc):
        tau_p = stokes.computeshearstress(n, m0, no, xx0, yy0, plain["first_density"])
        tau_a = stokes.computeshearstress(n, m0, no, xx0, yy0, den_acc)
        stress_err = float(np.max(np.abs(tau_a - tau_p)) / np.max(np.abs(tau_p)))
        log(f"shear stress of the state, accelerated vs plain FP64: {stress_err:.2e} relative")

    # --- ablation: the accelerations switched on one at a time ---------------
    ablation = {}
    if args.ablation_steps != 0:
        # the same warm-up and timed steps as the headline, from the same
        # starting state, unless a shorter sample is asked for
        ab_warm, ab_steps = ((args.warmup, args.steps) if args.ablation_steps < 0
                             else (1, args.ablation_steps))
        rows = [("preconditioner", dict(pc=True, mixed=False, warm=0, learned=False)),
                ("preconditioner+mixed", dict(pc=True, mixed=True, warm=0, learned=False)),
                ("preconditioner+mixed+extrapolated_guess",
                 dict(pc=True, mixed=True, warm=2, learned=False))]
        for name, kw in rows:
            r = run_device(name, ab_warm, ab_steps, timed_kernels=False, **kw)
            ablation[name] = {"steps_per_s": r["steps"] / r["elapsed"],
                              "gmres_iters_per_solve": r["timings"]["gmres_iters"] /
                              max(1, r["timings"]["solves"]),
                              "true_relres_max": r["true_relres_max"],
                              "solves_above_tol": int(r["tol_misses"][0])}
        ablation["sample"] = (f"each row: {ab_warm} untimed step(s), then {ab_steps} timed "
                              f"steps from the headline's starting state (device stepper)")
    if sp:
        ablation["plain_fp64"] = {k: sp[k] for k in ("steps_per_s", "gmres_iters_per_solve",
                                                      "true_relres_max", "solves_above_tol")}
    # the headline also carries the learned network's guess, which is asked
    # for only where no extrapolation exists (learned_guesses_used counts it)
    ablation["headline: preconditioner+mixed+extrapolated_guess, learned_guess where no "
             "extrapolation exists"] = {
        k: sa[k] for k in ("steps_per_s", "gmres_iters_per_solve", "true_relres_max",
                           "solves_above_tol", "learned_guesses_used")}

    # --- kernel rates of the headline run ------------------------------------
    kt, elapsed = acc["kt"], acc["elapsed"]
    ff = kt["farfield"]
    ff_splits = acc["solve_info"][-1]["ff_splits"] if acc["solve_info"] else 0
    ff_ms = ff["ms"] / max(1, ff["launches"])
    ff_pairs = ff["work"] / max(1, ff["launches"])
    ff_tflops = FF_FLOP_PER_PAIR * ff_pairs / (ff_ms * 1e-3) / 1e12 if ff_ms > 0 else 0.0
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
    # cluster far field: whole application, its direct sums (with the local
    # expansions' evaluation), and the multipole and local expansions
    fm, fd, fx = kt["fmm"], kt["fmm_direct"], kt["fmm_expand"]
    fmm = None
    if fm["launches"]:
        napp = fm["launches"]
        fmm = {"ms_per_application": fm["ms"] / napp,
               "equivalent_TFLOPs": FF_FLOP_PER_PAIR * fm["work"] / (fm["ms"] * 1e-3) / 1e12,
               "direct_ms": fd["ms"] / napp,
               "direct_pairs": fd["work"] / napp,
               "direct_TFLOPs": FF_FLOP_PER_PAIR * fd["work"] / (fd["ms"] * 1e-3) / 1e12,
               "expansions_ms": fx["ms"] / napp,
               "translations": fx["work"] / napp,
               "direct_fraction_of_pairs": fd["work"] / fm["work"]}
    scale = args.steps if acc["graphs"] else 1
    share = {k: scale * kt[k]["ms"] / (elapsed * 1e3) for k in kt}
    log(f"plain-fp64: {sp}")
    log(f"accelerated: {sa}")
    log(f"ablation: {ablation}")
    log("kernel time shares of the accelerated timed region: " +
        ", ".join(f"{k} {100 * v:.1f}%" for k, v in share.items()))
    if f32_ms:
        log(f"FP32 far field: {f32_ms:.3f} ms/launch, {f32_tflops:.2f} TFLOP/s over "
            f"{f32['launches']} launches")
    if fmm:
        log(f"cluster far field: {fmm['ms_per_application']:.3f} ms/application "
            f"(direct {fmm['direct_ms']:.3f} ms at {fmm['direct_TFLOPs']:.1f} TFLOP/s, "
            f"expansions {fmm['expansions_ms']:.3f} ms for {fmm['translations']:.0f} "
            f"translations), {fmm['equivalent_TFLOPs']:.0f} TFLOP/s equivalent")
    log(f"far field: {ff_ms:.3f} ms/launch, {ff_tflops:.2f} TFLOP/s; Gram-Schmidt {gs_gbs:.0f} GB/s; "
        f"stress eval {se['ms'] / max(1, se['launches']):.2f} ms/launch"
        + (f", {se_tflops:.2f} TFLOP/s" if se_tflops else "")
        + (f"; near-pair gather {near_gbs:.0f} GB/s over {nr['launches']} passes"
           if near_gbs else "; no near pairs"))

    # the barycentric near-pair gather pass: the eroded state has no near
    # pairs, so the pass is also timed on the packed 60 x 128 geometry
    near_packed = None
    if near_gbs is None:
        try:
            pw = WORKLOADS["bodies60_n128_tight"]
            pn, pm, pno, pxx, pyy = pw.npts, pw.nbodies, pw.nouter, *getnxy(workload_geometry(pw))[1:]
            eta = np.random.default_rng(7).standard_normal(stokes.nunk(pn, pm, pno))
            r = stokes.time_matvec(pn, pm, pno, pxx, pyy, eta, reps=10)
            if r["near_entries"] > 0:
                g = r["gather_bytes"] / (r["near_ms"] * 1e-3) / 1e9
                near_packed = {"workload": pw.name, "entries": r["near_entries"],
                               "bytes": r["gather_bytes"], "ms": r["near_ms"], "GBps": g,
                               "hbm_frac": g / PEAK_HBM_GBS}
                log(f"near-pair gather on {pw.name}: {r['near_entries']} entries, {g:.0f} GB/s")
        except Exception as e:  # a reported extra, never required
            log(f"packed near-pair measurement unavailable: {e}")

    # the dominant kernel: the direct sums of the cluster far field where the
    # autotuner chose it, else the direct far field
    if fmm and fd["ms"] >= ff["ms"]:
        ach = fmm["direct_TFLOPs"]
        roofline = {
            "bound": "valu",
            "pipe": "VALU FP64 (vector ALU; no matrix instructions in this kernel)",
            "kernel": "k_fmm_direct (direct sums of the cluster far field, FP64, 20 FLOP per pair)",
            "achieved": ach, "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
            "frac": ach / PEAK_FP64_TFLOPS,
            "traffic": FMM_DIRECT_BYTES.get(w.name),
            "traffic_unit": "bytes per launch",
        }
    else:
        roofline = {
            "bound": "valu",
            "pipe": "VALU FP64 (vector ALU; no matrix instructions in this kernel)",
            "kernel": "k_farfield (all-pairs double layer, FP64)",
            "achieved": ff_tflops, "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
            "frac": ff_tflops / PEAK_FP64_TFLOPS,
            "traffic": (FF_FETCH_BYTES[w.name] + 16.0 * ff_splits * (m0 * n + no)
                        if w.name in FF_FETCH_BYTES and ff_splits else None),
            "traffic_unit": "bytes per launch",
        }

    cpu = None
    if not args.no_cpu_baseline and sp:
        try:
            cpu = cpu_baseline(w, n, m0, no, xx0, yy0, sp["gmres_iters_per_solve"])
        except Exception as e:  # the baseline is reported, never required
            log(f"cpu baseline unavailable: {e}")

    out = {
        "metric": METRIC,
        "value": sa["steps_per_s"],
        "unit": "erosion steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * elapsed / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64+f32 (FP32 far field in part of the iterations; everything else FP64)",
        "data": ("synthetic: seeded perturbed-circle porous geometry eroded to half its body "
                 "area (committed theta-L state)" if w.state else
                 "synthetic: seeded perturbed-circle porous geometry, theta-L state"),
        "config": {
            "workload": w.name, "bodies": m0, "points_per_body": n, "wall_points": no,
            "boundary_points": no + n * m0, "gmres_tol": args.tol,
            "variant": "accelerated: per-curve two-level right preconditioner, FP32 far field "
                       "once the residual is below 1e4 tol, FP64 true residual with "
                       "compensated sums and restarts until it meets the tolerance, GMRES "
                       "guesses extrapolated from the last two solutions in the "
                       "preconditioned variable and, where none is kept, the learned "
                       "network's; the whole RK2 step in the library",
            "dt": w.dt, "fixpdrop": True, "ibc": w.ibc,
            "hip_graphs": bool(args.graphs),
            # kernel durations: HIP events around every launch outside the
            # GMRES iterations and around the launches of one iteration in
            # this many (each standing for that many)
            "kernel_timing_period": args.timing_period,
            "accelerated": sa,
            "plain_fp64": sp,
            "plain_fp64_sample": (f"{args.plain_steps} step(s) from the same state, no warm-up"
                                  if sp else None),
            "speedup_vs_plain_fp64": sa["steps_per_s"] / sp["steps_per_s"] if sp else None,
            "ablation": ablation,
            "first_solve_of_state": first,
            "stress_relerr_vs_plain_fp64": stress_err,
            "matvecs_per_s": sa["matvecs_per_s"],
            "cluster_farfield": fmm,
            "farfield_ms": ff_ms,
            "farfield32_ms": f32_ms,
            "farfield32_TFLOPs": f32_tflops,
            "farfield32_launches": f32["launches"],
            "gram_schmidt_GBps": gs_gbs,
            "stress_eval_TFLOPs": se_tflops,
            "near_gather_GBps": near_gbs,
            "near_gather_launches": nr["launches"],
            "near_gather_packed_60x128": near_packed,
            "kernel_share": share,
        },
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
