"""Host-side callers of the Stokes library.

Restates the Julia wrappers of ErosionPackage/src/DensityStress.jl with the
same names, argument meaning and call sequence: ``compute_density`` (:43-54),
``compute_density!`` (:58-66, here ``compute_density_``),
``compute_qoi_targets`` (:76-88, :91-98), ``regulargrid`` (:102-115),
``getpdrop`` (:129-154), ``getumax`` (:157-167), ``getnxyden`` (:173-185),
``compute_stress`` (:191-201, :204-210) and ``getstress!`` (:23-36, here
``getstress_``).
"""
from __future__ import annotations

import dataclasses
import time
import warnings

import numpy as np

from . import stokes
from .spectral import gaussfilter
from .theta_len import ParamSet, ThLenDenType, getnxy


def xyrot(xv: np.ndarray, yv: np.ndarray):
    """Rotate coordinates by 90 degrees CCW (DensityStress.jl:40)."""
    return -yv, xv


def compute_density(xx, yy, nbods: int, params: ParamSet, den0=None, verbose: bool = False):
    """Density solve through stokessolver_ (DensityStress.jl:43-54).

    ``den0`` is the GMRES initial guess; the reference passes zeros."""
    density, nits = stokes.stokessolver(params.npts, nbods, params.nouter, params.ifmm,
                                        params.ibary, params.ibc, params.maxl, xx, yy, den0)
    if verbose:
        print(f"The total number of GMRES iterations is {nits}")
    return density, nits


def compute_density_(thlenden: ThLenDenType, params: ParamSet, rotation: bool = False,
                     den0=None):
    """Fill thlenden.density (or denrot) unless already computed (DensityStress.jl:58-66).

    Returns the GMRES iteration count, or None when the cached density was used."""
    cached = thlenden.denrot if rotation else thlenden.density
    if len(cached) > 0:
        return None
    nbods, xv, yv = getnxy(thlenden)
    if rotation:
        xv, yv = xyrot(xv, yv)
    density, nits = compute_density(xv, yv, nbods, params, den0)
    if rotation:
        thlenden.denrot = density
    else:
        thlenden.density = density
    return nits


@dataclasses.dataclass
class TargetsType:
    """Target points and the quantities computed there (ThetaLen.jl:56-60)."""

    xtar: np.ndarray
    ytar: np.ndarray
    utar: np.ndarray = None
    vtar: np.ndarray = None
    ptar: np.ndarray = None
    vortar: np.ndarray = None


def compute_qoi_targets(xx, yy, density, xtar, ytar, npts: int, nbods: int, nouter: int,
                        ibary: int):
    """Velocity, pressure, vorticity at targets via computeqoitargets_ (DensityStress.jl:76-88)."""
    return stokes.computeqoitargets(npts, nbods, nouter, ibary, xx, yy, density, xtar, ytar)


def compute_qoi_targets_(thlenden: ThLenDenType, targets: TargetsType, params: ParamSet,
                         fixpdrop: bool = False, rotation: bool = False) -> None:
    """Fill targets.{utar,vtar,ptar,vortar} (DensityStress.jl:91-98)."""
    nbods, xv, yv, density = getnxyden(thlenden, params, fixpdrop, rotation)
    targets.utar, targets.vtar, targets.ptar, targets.vortar = compute_qoi_targets(
        xv, yv, density, targets.xtar, targets.ytar, params.npts, nbods, params.nouter,
        params.ibary)


def regulargrid(xlocs, ylocs):
    """Targets on the tensor grid xlocs x ylocs, x-major (DensityStress.jl:102-115)."""
    xlocs, ylocs = np.asarray(xlocs, float), np.asarray(ylocs, float)
    return np.repeat(xlocs, len(ylocs)), np.tile(ylocs, len(xlocs))


def regulargridtargs(xlocs, ylocs) -> TargetsType:
    xtar, ytar = regulargrid(xlocs, ylocs)
    return TargetsType(xtar, ytar)


def getpdrop(thlenden: ThLenDenType, params: ParamSet, x0: float = 2.0, rotation: bool = False):
    """Pressure drop from -x0 to x0 and the mean flux (DensityStress.jl:129-154)."""
    nypts = 13
    dy = 2 / nypts
    ylocs = -1 + 0.5 * dy + dy * np.arange(nypts)
    tarp = regulargridtargs([x0], ylocs)
    tarm = regulargridtargs([-x0], ylocs)
    compute_qoi_targets_(thlenden, tarp, params, fixpdrop=False, rotation=rotation)
    compute_qoi_targets_(thlenden, tarm, params, fixpdrop=False, rotation=rotation)
    pdrop = float(np.mean(tarm.ptar) - np.mean(tarp.ptar))
    qplus, qminus = float(np.mean(tarp.utar)), float(np.mean(tarm.utar))
    qavg = 0.5 * (qplus + qminus)
    qreldiff = (qplus - qminus) / qavg
    if qreldiff > 1e-3:
        warnings.warn(f"The flux does not match at x0 and -x0: qreldiff = {qreldiff:.3g}")
    return pdrop, qavg


def getumax(thlenden: ThLenDenType, params: ParamSet, fixpdrop: bool, x0: float = 2.0) -> float:
    """Density rescaling that holds the pressure drop at 4 x0 (DensityStress.jl:157-167)."""
    umax = 1.0
    if fixpdrop:
        pdrop = getpdrop(thlenden, params)[0]
        umax = 4 * x0 / (pdrop + 1e-4)
    return umax


def getnxyden(thlenden: ThLenDenType, params: ParamSet, fixpdrop: bool, rotation: bool):
    """Bodies, coordinates and the (rescaled) density (DensityStress.jl:173-185)."""
    nbods, xv, yv = getnxy(thlenden)
    rescale = getumax(thlenden, params, fixpdrop)
    if rotation:
        xv, yv = xyrot(xv, yv)
        density = rescale * thlenden.denrot
    else:
        density = rescale * thlenden.density
    return nbods, xv, yv, density


def compute_stress(thlenden: ThLenDenType, params: ParamSet, fixpdrop: bool = False,
                   rotation: bool = False) -> np.ndarray:
    """Shear stress as an (npts, nbods) array (DensityStress.jl:191-210)."""
    nbods, xv, yv, density = getnxyden(thlenden, params, fixpdrop, rotation)
    if nbods == 0:
        return np.zeros((params.npts, 0))
    tau = stokes.computeshearstress(params.npts, nbods, params.nouter, xv, yv, density, params.ibary)
    return tau.reshape(nbods, params.npts).T


def getstress_(thlenden: ThLenDenType, params: ParamSet, den0=None, timings: dict = None) -> np.ndarray:
    """Smoothed |tau| per body and its mean matau (DensityStress.jl:23-36)."""
    t0 = time.perf_counter()
    nits = compute_density_(thlenden, params, den0=den0)
    t1 = time.perf_counter()
    stress = compute_stress(thlenden, params, fixpdrop=params.fixpdrop)
    t2 = time.perf_counter()
    if timings is not None:
        timings["density_s"] = timings.get("density_s", 0.0) + (t1 - t0)
        timings["stress_s"] = timings.get("stress_s", 0.0) + (t2 - t1)
        if nits is not None:
            timings["gmres_iters"] = timings.get("gmres_iters", 0) + nits
            timings["solves"] = timings.get("solves", 0) + 1
            timings["true_relres_max"] = max(timings.get("true_relres_max", 0.0),
                                             stokes.last_solve_info()["true_relres"])
        timings["stresses"] = timings.get("stresses", 0) + 1
    for bod in range(len(thlenden.thlenvec)):
        stress[:, bod] = gaussfilter(np.abs(stress[:, bod]), params.sigma)
        thlenden.thlenvec[bod].matau = float(np.mean(stress[:, bod]))
    return stress
