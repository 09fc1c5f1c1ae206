"""Host-side callers of the Stokes library.

Restates the Julia wrappers of ErosionPackage/src/DensityStress.jl with the
same names, argument meaning and call sequence: ``compute_density`` (:43-54),
``compute_density!`` (:58-66, here ``compute_density_``), ``compute_stress``
(:191-201, :204-210) and ``getstress!`` (:23-36, here ``getstress_``).  The
pressure-drop rescaling (``fixpdrop``, :129-185) needs computeqoitargets_,
which this library does not provide; callers run with fixpdrop = False.
"""
from __future__ import annotations

import time

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
                     den0=None) -> int:
    """Fill thlenden.density (or denrot) unless already computed (DensityStress.jl:58-66).

    Returns the GMRES iteration count (0 when the cached density was used)."""
    cached = thlenden.denrot if rotation else thlenden.density
    if len(cached) > 0:
        return 0
    nbods, xv, yv = getnxy(thlenden)
    if rotation:
        xv, yv = xyrot(xv, yv)
    density, nits = compute_density(xv, yv, nbods, params, den0)
    if rotation:
        thlenden.denrot = density
    else:
        thlenden.density = density
    return nits


def compute_stress(thlenden: ThLenDenType, params: ParamSet, fixpdrop: bool = False,
                   rotation: bool = False) -> np.ndarray:
    """Shear stress as an (npts, nbods) array (DensityStress.jl:191-210)."""
    if fixpdrop:
        raise NotImplementedError("fixpdrop needs computeqoitargets_, which is not provided")
    nbods, xv, yv = getnxy(thlenden)
    density = thlenden.denrot if rotation else thlenden.density
    if rotation:
        xv, yv = xyrot(xv, yv)
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
        timings["gmres_iters"] = timings.get("gmres_iters", 0) + nits
        timings["solves"] = timings.get("solves", 0) + (1 if nits else 0)
    for bod in range(len(thlenden.thlenvec)):
        stress[:, bod] = gaussfilter(np.abs(stress[:, bod]), params.sigma)
        thlenden.thlenvec[bod].matau = float(np.mean(stress[:, bod]))
    return stress
