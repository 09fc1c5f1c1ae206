"""Theta-L state of the bodies (host side).

Restates ErosionPackage/src/ThetaLen.jl: ``ParamSet`` (:13-37),
``ThetaLenType`` (:42-44), ``ThLenDenType`` (:46-51), ``getxy`` (:92-101),
``getnxy`` (:103-111) and ``getalpha`` (:113-116).  Each body is a closed CCW
curve held as its tangent angle theta on the offset equal-arclength grid
alpha_j = (j - 1/2) 2 pi / n, its length L and the surface means of x and y.
"""
from __future__ import annotations

import dataclasses
import warnings
from typing import List

import numpy as np

from .spectral import specint


@dataclasses.dataclass
class ParamSet:
    """Run parameters with the reference defaults (ThetaLen.jl:13-37)."""

    npts: int = 128
    ibary: int = 1
    ifmm: int = 1
    ibc: int = 1
    epsfac: float = 15.0
    sigfac: float = 10.0
    dt: float = 1e-3
    outstride: int = 4
    fixpdrop: bool = True
    fixarea: bool = False
    tfin: float = 1.0
    maxl: int = 8000
    nouter: int = 1024

    @property
    def epsilon(self) -> float:
        return self.epsfac / self.npts

    @property
    def sigma(self) -> float:
        return self.sigfac / self.npts


@dataclasses.dataclass
class ThetaLenType:
    theta: np.ndarray
    len: float
    xsm: float
    ysm: float
    matau: float = float("nan")


@dataclasses.dataclass
class ThLenDenType:
    thlenvec: List[ThetaLenType]
    tt: float = 0.0
    density: np.ndarray = dataclasses.field(default_factory=lambda: np.zeros(0))
    denrot: np.ndarray = dataclasses.field(default_factory=lambda: np.zeros(0))


def new_thlenden(thlenvec: List[ThetaLenType]) -> ThLenDenType:
    return ThLenDenType(list(thlenvec), 0.0, np.zeros(0), np.zeros(0))


def getalpha(npts: int) -> np.ndarray:
    """Offset grid alpha_j = (j - 1/2) dalpha (ThetaLen.jl:113-116)."""
    dalpha = 2 * np.pi / npts
    return 0.5 * dalpha + dalpha * np.arange(npts, dtype=np.float64)


def test_theta_means(theta: np.ndarray) -> None:
    """Closure check of ThetaLen.jl:65-73: warns if mean(cos), mean(sin) are large."""
    maxmean = max(abs(np.mean(np.cos(theta))), abs(np.mean(np.sin(theta))))
    thresh = 20.0 / len(theta)
    if maxmean > thresh:
        warnings.warn(f"theta means: {maxmean:.3g} > {thresh:.3g}")


def getxy(thlen: ThetaLenType):
    """x, y of one body from theta, L and the surface means (ThetaLen.jl:92-101)."""
    theta, ln = thlen.theta, thlen.len
    test_theta_means(theta)
    if not ln > 0.0:
        raise ValueError("getxy: body length must be positive")
    dx = ln / (2 * np.pi) * (np.cos(theta) - np.mean(np.cos(theta)))
    dy = ln / (2 * np.pi) * (np.sin(theta) - np.mean(np.sin(theta)))
    return specint(dx) + thlen.xsm, specint(dy) + thlen.ysm


def getnxy(thlenden: ThLenDenType):
    """Body count and concatenated body-major x, y (ThetaLen.jl:103-111)."""
    nb = len(thlenden.thlenvec)
    if nb == 0:
        return 0, np.zeros(0), np.zeros(0)
    tv = thlenden.thlenvec
    if len({len(t.theta) for t in tv}) != 1:
        xs, ys = zip(*(getxy(t) for t in tv))
        return nb, np.concatenate(xs), np.concatenate(ys)
    # every body at once: the same arithmetic as getxy, batched over bodies
    theta = np.stack([t.theta for t in tv])
    ln = np.array([t.len for t in tv])
    if not np.all(ln > 0.0):
        raise ValueError("getxy: body length must be positive")
    c, s = np.cos(theta), np.sin(theta)
    mc, ms = np.mean(c, axis=1, keepdims=True), np.mean(s, axis=1, keepdims=True)
    maxmean = np.maximum(np.abs(mc), np.abs(ms))[:, 0]
    thresh = 20.0 / theta.shape[1]
    if np.any(maxmean > thresh):
        warnings.warn(f"theta means: {maxmean.max():.3g} > {thresh:.3g}")
    dx = ln[:, None] / (2 * np.pi) * (c - mc)
    dy = ln[:, None] / (2 * np.pi) * (s - ms)
    x = specint(dx) + np.array([t.xsm for t in tv])[:, None]
    y = specint(dy) + np.array([t.ysm for t in tv])[:, None]
    return nb, x.ravel(), y.ravel()


def circ2thlen(npts: int, rad: float, xc: float, yc: float) -> ThetaLenType:
    """A circle in theta-L form (Main.jl:29-34)."""
    alpha = getalpha(npts)
    return ThetaLenType(0.5 * np.pi + alpha, 2 * np.pi * rad, xc, yc, float("nan"))
