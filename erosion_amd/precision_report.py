"""Per-step error of the mixed-precision solve against the FP64 solve.

Runs ``--steps`` erosion steps (``rungekutta2``, ErosionPackage/src/
TimeStepping.jl:189-202) of a workload twice from the same seeded state:

* the FP64 trajectory: every far field in FP64;
* the mixed trajectory: the FP32 far field once GMRES's implicit residual is
  below ``mixed_switch`` x tol, FP64 true residual and restarts.

Both use the per-curve preconditioner (the only way GMRES reaches 1e-10 in a
few hundred iterations at 1e5 boundary points; the preconditioner changes the
iteration, not the solution beyond the tolerance).  At every state of the
FP64 trajectory the same boundary is also solved in mixed precision, so the
per-step solve error is measured on identical inputs:

* density:      max |x_mixed - x_fp64| / max |x_fp64|;
* shear stress: max |tau_mixed - tau_fp64| / max |tau_fp64| (computeshearstress_
                on each density);
* per-body area after the step: max_b |A_mixed - A_fp64| / A_fp64 between the
  two trajectories (the erosion model's own output);
* the FP64 answer's own spread: a second FP64 solve of the same state along
  another Krylov path (a body preconditioner with half the modes), compared
  with the first in density and shear stress.

The shear stress is that of the solver's density (computeshearstress_), before
the fixed-pressure-drop scaling the time stepper applies.

Prints one JSON line with the per-step records on stdout; progress goes to
stderr.  Usage:

    python -m erosion_amd.precision_report --workload bodies200_n512 --steps 5
"""
from __future__ import annotations

import argparse
import copy
import json
import sys
import time

import numpy as np

from . import stokes
from .geometry import WORKLOADS, workload_geometry
from .theta_len import ParamSet, getnxy
from .time_stepping import PreviousDensity, body_areas, rungekutta2


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def _configure(mixed: bool, body_modes: int = 32) -> None:
    stokes.set_preconditioner(stokes.PRECOND_CURVE_FOURIER, body_modes)
    stokes.set_precision(stokes.PRECISION_MIXED if mixed else stokes.PRECISION_FP64)


def _solve(params: ParamSet, thld, mixed: bool, den0=None, body_modes: int = 32):
    """Density, shear stress, iterations, solver info and seconds for one state."""
    _configure(mixed, body_modes)
    m, xx, yy = getnxy(thld)
    t0 = time.perf_counter()
    den, nits = stokes.stokessolver(params.npts, m, params.nouter, params.ifmm, params.ibary,
                                    params.ibc, params.maxl, xx, yy, den0)
    t1 = time.perf_counter()
    info = stokes.last_solve_info()
    tau = stokes.computeshearstress(params.npts, m, params.nouter, xx, yy, den, params.ibary)
    t2 = time.perf_counter()
    return den, tau, nits, info, t1 - t0, t2 - t1


def _relmax(a, b) -> float:
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


def run(workload: str, steps: int, tol: float, switch: float, save: str = ""):
    w = WORKLOADS[workload]
    # fixpdrop as in the reference's default ParamSet (DensityStress.jl:157-167):
    # with the inflow speed held fixed instead, the stress in a dense array is
    # so large that most bodies are deleted in the first step
    params = ParamSet(npts=w.npts, nouter=w.nouter, ibary=1, ifmm=1, ibc=w.ibc, dt=w.dt,
                      maxl=4000, fixpdrop=True)
    stokes.set_gmres_tol(tol)
    stokes.set_mixed_switch(switch)
    th64 = workload_geometry(w)
    th32 = copy.deepcopy(th64)
    g64, g32 = PreviousDensity(2, w.nouter), PreviousDensity(2, w.nouter)
    area0 = body_areas(th64)
    records = []
    try:
        for k in range(steps + 1):
            # same-state solves: FP64 and mixed on the FP64 trajectory's state
            probe = copy.deepcopy(th64)
            d64, t64, n64, i64, s64, ss64 = _solve(params, probe, mixed=False)
            d32, t32, n32, i32, s32, ss32 = _solve(params, probe, mixed=True)
            # a second FP64 solve along another Krylov path (coarser body
            # preconditioner): how far two FP64 solutions that both meet the
            # stopping test lie apart -- the FP64 answer's own spread
            d2, t2, n2, *_ = _solve(params, probe, mixed=False, body_modes=16)
            if save and k == 0:
                m0, xx0, yy0 = getnxy(probe)
                np.savez(save, xx=xx0, yy=yy0, den_fp64=d64, den_mixed=d32, den_fp64_b=d2,
                         tau_fp64=t64, tau_mixed=t32, tau_fp64_b=t2, n=w.npts, m=m0, no=w.nouter)
            rec = {
                "step": k,
                "bodies": len(th64.thlenvec),
                "density_relerr": _relmax(d32, d64),
                "stress_relerr": _relmax(t32, t64),
                "fp64_spread_density": _relmax(d2, d64),
                "fp64_spread_stress": _relmax(t2, t64),
                "iters_fp64": int(n64), "iters_mixed": int(n32),
                "fp32_iters": int(i32["fp32_iters"]),
                "true_relres_fp64": i64["true_relres"], "true_relres_mixed": i32["true_relres"],
                "solve_s_fp64": s64, "solve_s_mixed": s32,
                "stress_s": ss64,
            }
            if len(th64.thlenvec) == len(th32.thlenvec):
                a64, a32 = body_areas(th64), body_areas(th32)
                rec["area_relerr"] = float(np.max(np.abs(a32 - a64) / a64))
                rec["theta_maxdiff"] = max(float(np.max(np.abs(a.theta - b.theta)))
                                           for a, b in zip(th32.thlenvec, th64.thlenvec))
                if len(a64) == len(area0):
                    rec["area_lost_frac"] = float(1.0 - a64.sum() / area0.sum())
            records.append(rec)
            log(f"[{workload}] state {k}: " + ", ".join(
                f"{key} {v:.3g}" if isinstance(v, float) else f"{key} {v}" for key, v in rec.items()))
            if k == steps:
                break
            _configure(False)
            th64 = rungekutta2(th64, params, guess=g64, timings={})
            _configure(True)
            th32 = rungekutta2(th32, params, guess=g32, timings={})
    finally:
        stokes.set_preconditioner(stokes.PRECOND_NONE)
        stokes.set_precision(stokes.PRECISION_FP64)
    return {
        "workload": w.name, "bodies": w.nbodies, "points_per_body": w.npts,
        "wall_points": w.nouter, "boundary_points": w.nbodies * w.npts + w.nouter,
        "gmres_tol": tol, "mixed_switch": switch, "dt": w.dt, "steps": steps,
        "records": records,
        "max_density_relerr": max(r["density_relerr"] for r in records),
        "max_stress_relerr": max(r["stress_relerr"] for r in records),
        "max_area_relerr": max((r.get("area_relerr", 0.0) for r in records), default=None),
        "max_fp64_spread_stress": max(r["fp64_spread_stress"] for r in records),
    }


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--workload", default="bodies200_n512", choices=sorted(WORKLOADS))
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--tol", type=float, default=1e-10)
    ap.add_argument("--mixed-switch", type=float, default=1e4)
    ap.add_argument("--save", default="", help="write the first state's inputs and densities (.npz)")
    args = ap.parse_args(argv)
    out = run(args.workload, args.steps, args.tol, args.mixed_switch, args.save)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
