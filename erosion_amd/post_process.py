"""Post-processing that uses the solver: surface pressure, drag, resistivity.

Restates ErosionPackage/src/PostProcess.jl: ``getns`` (:23-33),
``compute_pressure`` (:103-121), ``drag`` / ``DragData`` (:124-160) and
``resistivity`` (:62-71).  The pressure comes from the GPU's
``computepressure_``; everything here is host-side bookkeeping.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import stokes
from .density_stress import compute_stress, getnxyden, getpdrop, getumax
from .spectral import gaussfilter
from .theta_len import ParamSet, ThLenDenType


def getns(theta: np.ndarray, rotation: bool = False):
    """CCW tangent (sx, sy) and inward normal (nx, ny) (PostProcess.jl:23-33)."""
    if not rotation:
        sx, sy = np.cos(theta), np.sin(theta)
    else:
        sx, sy = -np.sin(theta), np.cos(theta)
    return sx, sy, -sy, sx


def compute_pressure_xy(xx, yy, density, npts: int, nbods: int, nouter: int) -> np.ndarray:
    """Fortran-signature wrapper (PostProcess.jl:103-112)."""
    if nbods == 0:
        return np.zeros(0)
    return stokes.computepressure(npts, nbods, nouter, xx, yy, density)


def compute_pressure(thlenden: ThLenDenType, params: ParamSet, fixpdrop: bool = False,
                     rotation: bool = False) -> np.ndarray:
    """Pressure on every body, shape (npts, nbods) (PostProcess.jl:114-121)."""
    nbods, xv, yv, density = getnxyden(thlenden, params, fixpdrop, rotation)
    p = compute_pressure_xy(xv, yv, density, params.npts, nbods, params.nouter)
    return p.reshape(nbods, params.npts).T


@dataclass
class DragData:
    pdragx: float
    pdragy: float
    vdragx: float
    vdragy: float
    umax: float
    tau_all: np.ndarray
    atau_all: np.ndarray
    press_all: np.ndarray


def drag(thlenden: ThLenDenType, params: ParamSet, rotation: bool = False) -> DragData:
    """Pressure and viscous drag summed over the bodies (PostProcess.jl:130-160)."""
    tv = thlenden.thlenvec
    nbods, npts = len(tv), params.npts
    tau_all = compute_stress(thlenden, params, fixpdrop=False, rotation=rotation)
    press_all = compute_pressure(thlenden, params, fixpdrop=False, rotation=rotation)
    atau_all = np.zeros((npts, nbods))
    pdx = pdy = vdx = vdy = 0.0
    for b in range(nbods):
        press, tau = press_all[:, b], tau_all[:, b]
        atau_all[:, b] = gaussfilter(np.abs(tau), params.sigma)
        sx, sy, nx, ny = getns(tv[b].theta, rotation)
        ds = tv[b].len / npts
        pdx += float(press @ nx) * ds
        pdy += float(press @ ny) * ds
        vdx += float(tau @ sx) * ds
        vdy += float(tau @ sy) * ds
    umax = getumax(thlenden, params, params.fixpdrop)
    return DragData(pdx, pdy, vdx, vdy, umax, tau_all, atau_all, press_all)


def resistivity(thlenden: ThLenDenType, params: ParamSet, x0: float = 2.0,
                rotation: bool = False) -> float:
    """0.5 pdrop / qavg, less the wall part 3 x0 for pipe flow (PostProcess.jl:62-71)."""
    pdrop, qavg = getpdrop(thlenden, params, x0, rotation=rotation)
    resist = 0.5 * pdrop / qavg
    if params.ibc == 0:
        resist += -3 * x0
    return resist
