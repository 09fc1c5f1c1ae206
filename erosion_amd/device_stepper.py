"""The erosion step with the theta-L update on the GPU.

``rungekutta2_device`` has the meaning of ``time_stepping.rungekutta2``
(ErosionPackage/src/TimeStepping.jl:189-202): one midpoint RK2 step of every
body, two density solves, two shear-stress evaluations, the fixed-pressure-drop
rescaling and the deletion of vanishing bodies.  The whole step runs in the
library (erosion_rungekutta2): the boundary points, densities and stresses of
both stages stay on the GPU, and only the theta-L state crosses to the host,
once in and once out.  GMRES guesses come from the solutions the library keeps
between solves, in the preconditioned variable u = P x (``warm`` 2: linear
extrapolation of the last two, 1: the last, 0: zeros as the reference); when
bodies are deleted the kept solutions drop their unknowns.  Unlike the host
stepper, ``thld0`` is left unchanged (the host version deletes bodies from it
in place).
"""
from __future__ import annotations

import math

import numpy as np

from . import stokes
from .theta_len import ParamSet, ThetaLenType, ThLenDenType, new_thlenden


def rungekutta2_device(thld0: ThLenDenType, params: ParamSet, warm: int = 2,
                       timings: dict = None, guess=None) -> ThLenDenType:
    """``guess(thld0)`` (for example the trained network, ``LearnedGuess``) is
    asked for a first-stage GMRES guess, in the preconditioned variable u, only
    when the library keeps no solution that fits the state (the first step of
    a run)."""
    tv = thld0.thlenvec
    n = params.npts
    if not tv:
        out = new_thlenden([])
        out.tt = thld0.tt + params.dt
        return out
    if any(len(t.theta) != n for t in tv):
        raise ValueError("every body needs params.npts points")
    prm = [params.dt, params.epsilon, params.sigma, float(params.fixarea), float(params.fixpdrop),
           params.ibc, params.ibary, params.maxl, warm]
    g0 = None
    if guess is not None and warm > 0:
        h = stokes.stepper_history()
        if h["count"] == 0 or h["bodies"] != len(tv) or \
                h["length"] != stokes.nunk(n, len(tv), params.nouter):
            g0 = guess(thld0)
    theta, ln, xs, ys, info = stokes.rungekutta2(
        n, params.nouter, prm, np.stack([t.theta for t in tv]), [t.len for t in tv],
        [t.xsm for t in tv], [t.ysm for t in tv], guess=g0)
    out = new_thlenden([ThetaLenType(theta[b].copy(), float(ln[b]), float(xs[b]), float(ys[b]),
                                     math.nan) for b in range(len(ln))])
    out.tt = thld0.tt + params.dt
    if timings is not None:
        timings["gmres_iters"] = timings.get("gmres_iters", 0) + int(info["iters_stage1"] +
                                                                     info["iters_stage2"])
        timings["solves"] = timings.get("solves", 0) + 2
        timings["stresses"] = timings.get("stresses", 0) + 2
        if info["deleted"]:
            timings["deleted"] = timings.get("deleted", 0) + int(info["deleted"])
        for k in ("warm_starts", "given_guesses"):
            timings[k] = timings.get(k, 0) + int(info[k])
        timings["true_relres_max"] = max(timings.get("true_relres_max", 0.0),
                                         info["true_relres_max"])
        timings["last_info"] = info
    return out
