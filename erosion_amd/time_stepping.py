"""Erosion time stepping in the theta-L variables (host side).

Restates ErosionPackage/src/TimeStepping.jl: ``tangvel`` (:42-52),
``getderivs`` (:63-84, :86-95), ``delete_indices`` (:103-126), ``timestep!``
(:130-185, here ``timestep_``) and ``rungekutta2`` (:189-202).  One erosion
step is one ``rungekutta2`` call: two density solves and two shear-stress
evaluations on the GPU, with O(n log n) spectral updates per body on the host
in between.

``den0`` hooks: every solve can start GMRES from an initial guess instead of
zeros.  The reference always starts from zeros (DensityStress.jl:45); the
guess changes only the iteration count, never the stopping test.
"""
from __future__ import annotations

import math
import warnings
from typing import Callable, List, Optional

import numpy as np

from .density_stress import getstress_
from .spectral import expsmooth, specdiff, specint
from .theta_len import ParamSet, ThetaLenType, ThLenDenType, getalpha, new_thlenden


class DerivsType:
    """dL/dt (mterm), the nonlinear theta term, and d(xsm, ysm)/dt."""

    __slots__ = ("mterm", "nterm", "xsmdot", "ysmdot")

    def __init__(self, mterm, nterm, xsmdot, ysmdot):
        self.mterm, self.nterm, self.xsmdot, self.ysmdot = mterm, nterm, xsmdot, ysmdot


def tangvel(dtheta: np.ndarray, vnorm: np.ndarray):
    """Tangential velocity and mterm = dL/dt (TimeStepping.jl:42-52)."""
    dthvn = dtheta * vnorm
    mdthvn = float(np.mean(dthvn))
    mterm = -2 * math.pi * mdthvn
    vtang = specint(dthvn - mdthvn)
    return vtang, mterm


def getderivs_body(thlen: ThetaLenType, params: ParamSet, atau: np.ndarray) -> DerivsType:
    """Time derivatives of one body (TimeStepping.jl:63-84)."""
    theta, ln, matau = thlen.theta, thlen.len, thlen.matau
    ataueff = atau - float(params.fixarea) * matau
    alpha = getalpha(len(theta))
    dtheta = specdiff(theta - alpha) + 1.0
    vnorm = ataueff + params.epsilon * matau * (dtheta - 1.0)
    vtang, mterm = tangvel(dtheta, vnorm)
    datau = specdiff(atau)
    nterm = 2 * math.pi / ln * (datau + dtheta * vtang)
    xsmdot = float(np.mean(-vnorm * np.sin(theta) + vtang * np.cos(theta)))
    ysmdot = float(np.mean(vnorm * np.cos(theta) + vtang * np.sin(theta)))
    return DerivsType(mterm, nterm, xsmdot, ysmdot)


def getderivs_bodies(thlenvec, params: ParamSet, atau: np.ndarray) -> List[DerivsType]:
    """getderivs_body for every body at once (bodies with equal point counts):
    the same arithmetic, batched along the body axis."""
    theta = np.stack([t.theta for t in thlenvec])
    ln = np.array([t.len for t in thlenvec])
    matau = np.array([t.matau for t in thlenvec])[:, None]
    at = np.asarray(atau, dtype=np.float64).T  # (bodies, points)
    ataueff = at - float(params.fixarea) * matau
    alpha = getalpha(theta.shape[1])
    dtheta = specdiff(theta - alpha) + 1.0
    vnorm = ataueff + params.epsilon * matau * (dtheta - 1.0)
    dthvn = dtheta * vnorm
    mdthvn = np.mean(dthvn, axis=1, keepdims=True)
    mterm = -2 * math.pi * mdthvn[:, 0]
    vtang = specint(dthvn - mdthvn)
    datau = specdiff(at)
    nterm = 2 * math.pi / ln[:, None] * (datau + dtheta * vtang)
    xsmdot = np.mean(-vnorm * np.sin(theta) + vtang * np.cos(theta), axis=1)
    ysmdot = np.mean(vnorm * np.cos(theta) + vtang * np.sin(theta), axis=1)
    return [DerivsType(float(mterm[b]), nterm[b], float(xsmdot[b]), float(ysmdot[b]))
            for b in range(len(thlenvec))]


def getderivs(thlenden: ThLenDenType, params: ParamSet, den0=None, timings=None) -> List[DerivsType]:
    """Solve, stress, then per-body derivatives (TimeStepping.jl:86-95)."""
    atau = getstress_(thlenden, params, den0=den0, timings=timings)
    tv = thlenden.thlenvec
    if tv and len({len(t.theta) for t in tv}) == 1:
        return getderivs_bodies(tv, params, atau)
    return [getderivs_body(t, params, atau[:, b]) for b, t in enumerate(tv)]


def delete_indices(thld0: ThLenDenType, dvec: List[DerivsType], dt: float) -> List[int]:
    """Bodies to remove: len <= -sqrt(2) mterm dt or mterm > 10 (TimeStepping.jl:103-126)."""
    ndts, mthresh = 1.0, 10.0
    assert len(dvec) == len(thld0.thlenvec)
    out = []
    for bod, (thl, d) in enumerate(zip(thld0.thlenvec, dvec)):
        minlen = -math.sqrt(2) * d.mterm * ndts * dt
        if d.mterm > 0.0:
            warnings.warn("mterm is positive.")
        if thl.len <= minlen or d.mterm > mthresh:
            out.append(bod)
    return out


def timestep_(thld0: ThLenDenType, thld_derivs: ThLenDenType, dt1: float, dt2: float,
              params: ParamSet, den0=None, timings=None) -> ThLenDenType:
    """Advance every body by dt1 with derivatives taken at thld_derivs (TimeStepping.jl:130-185)."""
    dvec = getderivs(thld_derivs, params, den0=den0, timings=timings)
    deletevec = delete_indices(thld0, dvec, dt1)
    for bod in sorted(deletevec, reverse=True):
        del thld0.thlenvec[bod]
        del dvec[bod]
        if thld0 is not thld_derivs:
            del thld_derivs.thlenvec[bod]
    if timings is not None and deletevec:
        timings["deleted"] = timings.get("deleted", 0) + len(deletevec)
    nbods = len(thld0.thlenvec)
    assert len(dvec) == len(thld_derivs.thlenvec) == nbods
    eps = params.epsilon
    alpha = getalpha(params.npts)

    def zetafun(ln, matau):
        return 2 * math.pi / ln * matau

    if nbods == 0:
        return new_thlenden([])
    # every body at once; per body this is exactly the loop of TimeStepping.jl:150-182
    t0v, tdv = thld0.thlenvec, thld_derivs.thlenvec
    len0 = np.array([t.len for t in t0v])
    mterm = np.array([d.mterm for d in dvec])
    len1 = len0 + dt1 * mterm
    if not np.all(len1 > 0.0):
        raise FloatingPointError("body length became non-positive")
    zeta0 = np.array([zetafun(t.len, t.matau) for t in t0v])
    zetad = np.array([zetafun(t.len, t.matau) for t in tdv])
    fac1 = eps * dt1 * zetad
    fac2 = eps * dt2 * (2 * zetad - zeta0)
    theta0 = np.stack([t.theta for t in t0v])
    nterm = np.stack([d.nterm for d in dvec])
    th1 = expsmooth(theta0 - alpha, fac1) + alpha
    th1 = th1 + dt1 * expsmooth(nterm, fac2)
    thlv1 = [ThetaLenType(th1[b], float(len1[b]), t0v[b].xsm + dt1 * dvec[b].xsmdot,
                          t0v[b].ysm + dt1 * dvec[b].ysmdot, float("nan"))
             for b in range(nbods)]
    return new_thlenden(thlv1)


class PreviousDensity:
    """GMRES initial guess from the most recently computed densities, while
    the body count is unchanged (after a deletion the guess falls back to
    zeros until new densities exist).

    Consecutive solves of an erosion run are half a time step apart.  With
    ``order`` 1 the guess is the previous density; with ``order`` 2 it is the
    linear extrapolation 2 x_k - x_(k-1) of the last two (error O(dt^2)
    instead of O(dt) -- which matters because the solution is about 10^3
    times larger than the right-hand side, so a guess only lowers the
    starting residual when its relative error is well below 1e-3).

    A state whose bodies were deleted after its solve (timestep! deletes from
    the state it just solved, TimeStepping.jl:135-141) holds a density of the
    old body list: with ``nouter`` known such a density is recognised by its
    length and empties the history instead of entering it."""

    def __init__(self, order: int = 1, nouter: Optional[int] = None):
        if order not in (1, 2):
            raise ValueError("order must be 1 or 2")
        self.order = order
        self.nouter = nouter
        self.hist: List[np.ndarray] = []
        self.hist_m = -1

    def _fits(self, den: np.ndarray, thld: ThLenDenType) -> bool:
        m = len(thld.thlenvec)
        n = len(thld.thlenvec[0].theta) if m else 0
        if self.nouter is not None:
            return len(den) == 2 * self.nouter + (2 * n + 3) * m
        # without nouter: the wall part 2*nouter must come out positive and even
        no2 = len(den) - 2 * n * m - 3 * m
        return no2 > 0 and no2 % 2 == 0

    def __call__(self, thld: ThLenDenType):
        m = len(thld.thlenvec)
        if not self.hist or self.hist_m != m or not self._fits(self.hist[-1], thld):
            return None
        if self.order == 2 and len(self.hist) == 2:
            return 2.0 * self.hist[-1] - self.hist[-2]
        return self.hist[-1]

    @property
    def last(self):
        return self.hist[-1] if self.hist else None

    def seen(self, thld: ThLenDenType) -> None:
        if len(thld.density) == 0:
            return
        m = len(thld.thlenvec)
        den = np.asarray(thld.density)
        if self.nouter is not None and not self._fits(den, thld):
            self.hist, self.hist_m = [], -1  # bodies deleted since this solve
            return
        if m != self.hist_m:
            self.hist = []
        self.hist = (self.hist + [den])[-2:]
        self.hist_m = m


def rungekutta2(thld0: ThLenDenType, params: ParamSet,
                guess: Optional[Callable[[ThLenDenType], Optional[np.ndarray]]] = None,
                timings=None) -> ThLenDenType:
    """One erosion step: midpoint RK2 (TimeStepping.jl:189-202).

    ``guess(thlenden)`` may return a GMRES initial guess for that state; if it
    has a ``seen`` method it is told about every state whose density was just
    computed."""
    dt = params.dt
    g0 = guess(thld0) if guess else None
    thld05 = timestep_(thld0, thld0, 0.5 * dt, 0.5 * dt, params, den0=g0, timings=timings)
    if guess is not None and hasattr(guess, "seen"):
        guess.seen(thld0)
    g1 = guess(thld05) if guess else None
    thld1 = timestep_(thld0, thld05, dt, 0.5 * dt, params, den0=g1, timings=timings)
    if guess is not None and hasattr(guess, "seen"):
        guess.seen(thld05)
    thld1.tt = thld0.tt + dt
    return thld1


def body_areas(thlenden: ThLenDenType) -> np.ndarray:
    """Area of every body, (1/2) closed integral of (x dy - y dx) with spectral
    derivatives (the quantity of ErosionPackage/src/PostProcess.jl:36-58)."""
    from .theta_len import getxy

    out = []
    for t in thlenden.thlenvec:
        x, y = getxy(t)
        out.append(0.5 * float(np.mean(x * specdiff(y) - y * specdiff(x))) * 2 * math.pi)
    return np.array(out)
