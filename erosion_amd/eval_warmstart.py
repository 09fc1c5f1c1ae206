"""Warm starts on an eroded 400-body geometry (python -m erosion_amd.eval_warmstart).

Erodes the 400-body x 256-point workload (wall 4096, seed 1005) with the
accelerated solve (preconditioner, mixed precision, tol 1e-10, fixpdrop) and
the time stepper's warm start, until half of the initial body area is lost
(or ``--max-steps``).  Then, on the eroded state, solves the same boundary
from several initial guesses and reports GMRES iterations and seconds:

* plain:                 the reference iteration (FP64, no preconditioner) from zeros;
* accelerated, zeros:    preconditioner + mixed precision from zeros;
* accelerated, learned:  from the trained network's guess (scaled by the factor
                         that minimises its residual);
* accelerated, previous: from 2 u_k - u_(k-1) of the run's last two solutions
                         (host stepper only);
* with ``--body-modes``: accelerated from zeros with other body preconditioner sizes.

Also records the GMRES iterations of every step of the run.  Prints one JSON
line; progress goes to stderr.
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np

from . import learned as LR
from . import stokes
from .geometry import WORKLOADS, workload_geometry
from .theta_len import ParamSet, getnxy
from .time_stepping import body_areas, rungekutta2


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def _accelerated(on: bool, body_modes=None) -> None:
    # accelerated solves take guesses in the preconditioned variable
    stokes.set_guess_space(stokes.GUESS_PRECONDITIONED if on else stokes.GUESS_DENSITY)
    if on:
        if body_modes is None:
            stokes.set_preconditioner(stokes.PRECOND_CURVE_FOURIER)
        else:
            stokes.set_preconditioner(stokes.PRECOND_CURVE_FOURIER, body_modes)
    else:
        stokes.set_preconditioner(stokes.PRECOND_NONE)
    stokes.set_precision(stokes.PRECISION_MIXED if on else stokes.PRECISION_FP64)


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--workload", default="bodies400_n256", choices=sorted(WORKLOADS))
    ap.add_argument("--area-lost", type=float, default=0.5)
    ap.add_argument("--max-steps", type=int, default=40)
    ap.add_argument("--no-plain", action="store_true", help="skip the unpreconditioned solve")
    ap.add_argument("--device-stepper", action="store_true",
                    help="erode with the whole step in the library (erosion_rungekutta2)")
    ap.add_argument("--body-modes", type=int, nargs="*", default=[],
                    help="also solve the eroded state with these body preconditioner sizes")
    ap.add_argument("--save", default="", help="write the eroded boundary (.npz)")
    ap.add_argument("--run-body-modes", type=int, default=None,
                    help="body preconditioner size during the run (default: the library's)")
    a = ap.parse_args(argv)
    w = WORKLOADS[a.workload]
    params = ParamSet(npts=w.npts, nouter=w.nouter, ibary=1, ifmm=1, ibc=w.ibc, dt=w.dt,
                      maxl=8000, fixpdrop=True)
    stokes.set_gmres_tol(1e-10)
    _accelerated(True, a.run_body_modes)
    modes_run = stokes.get_preconditioner()[1]
    th = workload_geometry(w)
    area0 = float(body_areas(th).sum())
    guess = LR.WarmStart(w.nouter, npts=w.npts, ibc=w.ibc)
    steps, t_run, lost = 0, 0.0, 0.0
    timings = {}
    per_step = []
    if a.device_stepper:
        from .device_stepper import rungekutta2_device
        stokes.stepper_reset()
    while lost < a.area_lost and steps < a.max_steps:
        t0 = time.perf_counter()
        before = timings.get("gmres_iters", 0)
        if a.device_stepper:
            th = rungekutta2_device(th, params, warm=2, timings=timings, guess=guess.net)
        else:
            th = rungekutta2(th, params, guess=guess, timings=timings)
        t_run += time.perf_counter() - t0
        steps += 1
        per_step.append(int(timings.get("gmres_iters", 0) - before))
        lost = 1.0 - float(body_areas(th).sum()) / area0
        log(f"step {steps}: {len(th.thlenvec)} bodies, area lost {100 * lost:.1f}%, "
            f"{per_step[-1]} GMRES iterations this step")
    m, xx, yy = getnxy(th)
    n, no = w.npts, w.nouter
    if a.save:
        np.savez(a.save, xx=xx, yy=yy, n=n, m=m, no=no)
    prev = guess.prev(th) if not a.device_stepper else None

    def solve(label, accel, den0=None, body_modes=None):
        _accelerated(accel, body_modes if body_modes is not None else a.run_body_modes)
        t0 = time.perf_counter()
        den, it = stokes.stokessolver(n, m, no, 1, 1, w.ibc, params.maxl, xx, yy, den0)
        dt = time.perf_counter() - t0
        info = stokes.last_solve_info()
        log(f"{label}: {it} iterations, {dt:.2f} s, true relres {info['true_relres']:.2e}")
        return den, {"iterations": it, "seconds": dt, "true_relres": info["true_relres"],
                     "r0_relres": info["r0_relres"]}

    res = {}
    d0, res["accelerated_zeros"] = solve("accelerated, zeros", True)
    if guess.net is not None:
        g = guess.net.predict(th, no, scale=True)
        _, res["accelerated_learned"] = solve("accelerated, learned", True, g)
    if prev is not None:
        _, res["accelerated_previous"] = solve("accelerated, previous solutions", True, prev)
    for k in a.body_modes:
        _, res[f"accelerated_zeros_body_modes_{k}"] = solve(f"accelerated, zeros, body modes {k}",
                                                              True, body_modes=k)
    if not a.no_plain:
        _, res["plain_zeros"] = solve("plain FP64, zeros", False)
    _accelerated(False)
    print(json.dumps({
        "workload": w.name, "bodies_initial": w.nbodies, "bodies_eroded": m,
        "points_per_body": n, "wall_points": no, "gmres_tol": 1e-10,
        "area_lost": lost, "steps": steps, "steps_per_s_accelerated_warm": steps / t_run,
        "gmres_iters_per_solve_run": timings.get("gmres_iters", 0) / max(1, timings.get("solves", 1)),
        "gmres_iters_per_step": per_step,
        "body_modes_run": modes_run,
        "learned_guesses_used_in_run": guess.learned_used,
        "solves": res,
    }), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
