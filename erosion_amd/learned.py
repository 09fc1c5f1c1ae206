"""Learned GMRES initial guess for the density solve (no counterpart in the
reference, whose solves start from zeros, ErosionPackage/src/DensityStress.jl:45).

A multilayer perceptron (3 hidden layers x 256 tanh units, BF16 products with
FP32 accumulation on the matrix cores, libstokes_hip.so's erosion_mlp_*) maps
each body's features to that body's density in Fourier coefficients:

  features (128): the body's theta-L descriptor -- Fourier modes 1..16 of
    theta - alpha (32 values), cos and sin of its mean (the orientation), the
    length L and the centre (xsm, ysm) -- and, for each of its 8 nearest
    neighbours by centre distance, the relative centre (2), L (1) and modes
    1..4 of theta - alpha (8); zero padded to 128.
  output (69): Fourier coefficients 0..16 (cosine, sine) of eta_x and eta_y
    along the body's points, times the body's speed L / 2 pi (66; the density
    of a small body is large, its rigid-motion part scaling like the
    stokeslet over the speed), and its stokeslet x, y and rotlet (3).

The wall part of the guess is the density of the same wall with no bodies
(one solve, kept per wall size).  The guess only changes the GMRES iteration
count: every solve still meets the same tolerance, and zeros remain the
fallback (for bodies whose point count the network was not trained on).
"""
from __future__ import annotations

import ctypes
import math
import os
from typing import Dict, Optional

import numpy as np

from . import stokes
from .theta_len import ThLenDenType, getalpha, getnxy

NIN, NHID, NLAYERS, NOUT = 128, 256, 3, 69
NMODES = 16            # own descriptor and output modes
NNEIGH, NMODES_NB = 8, 4
WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "warmstart_mlp.npz")

_fp = ctypes.POINTER(ctypes.c_float)
_dpp = ctypes.POINTER(ctypes.c_double)


def _bind():
    L = stokes.lib()
    if not getattr(L, "_mlp_bound", False):
        L.erosion_mlp_param_count.argtypes = [ctypes.c_int] * 4 + [ctypes.POINTER(ctypes.c_int64)]
        L.erosion_mlp_forward.argtypes = [ctypes.c_int] * 4 + [_fp, ctypes.c_int64, _fp, _fp]
        L.erosion_mlp_gradients.argtypes = ([ctypes.c_int] * 4 + [_fp, ctypes.c_int64, _fp, _fp, _fp,
                                                                   _dpp])
        L.erosion_mlp_train.argtypes = ([ctypes.c_int] * 4 + [_fp, ctypes.c_int64, _fp, _fp,
                                                               ctypes.c_int, ctypes.c_int,
                                                               ctypes.c_float, ctypes.c_uint64,
                                                               _dpp])
        for f in ("erosion_mlp_param_count", "erosion_mlp_forward", "erosion_mlp_gradients",
                  "erosion_mlp_train"):
            getattr(L, f).restype = ctypes.c_int
        L._mlp_bound = True
    return L


def _f32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float32)


def _fptr(a: np.ndarray):
    return a.ctypes.data_as(_fp)


def param_count(shape=(NIN, NHID, NLAYERS, NOUT)) -> int:
    c = ctypes.c_int64(0)
    if _bind().erosion_mlp_param_count(*shape, ctypes.byref(c)) != 0:
        raise stokes.StokesError(stokes.last_error())
    return int(c.value)


def init_params(shape=(NIN, NHID, NLAYERS, NOUT), seed: int = 0) -> np.ndarray:
    """Glorot-uniform weights, zero biases, in the flat layout of mlp.hip."""
    nin, nhid, nlayers, nout = shape
    rng = np.random.default_rng(seed)
    parts, prev = [], nin
    for l in range(nlayers + 1):
        o = nhid if l < nlayers else nout
        lim = math.sqrt(6.0 / (prev + o))
        parts.append(rng.uniform(-lim, lim, size=o * prev))
        parts.append(np.zeros(o))
        prev = o
    return _f32(np.concatenate(parts))


def mlp_forward(params, x, shape=(NIN, NHID, NLAYERS, NOUT)) -> np.ndarray:
    x = _f32(x)
    params = _f32(params)
    y = np.zeros((x.shape[0], shape[3]), dtype=np.float32)
    if _bind().erosion_mlp_forward(*shape, _fptr(params), x.shape[0], _fptr(x), _fptr(y)) != 0:
        raise stokes.StokesError(stokes.last_error())
    return y


def mlp_gradients(params, x, t, shape=(NIN, NHID, NLAYERS, NOUT)):
    x, t, params = _f32(x), _f32(t), _f32(params)
    g = np.zeros_like(params)
    loss = ctypes.c_double(0.0)
    if _bind().erosion_mlp_gradients(*shape, _fptr(params), x.shape[0], _fptr(x), _fptr(t),
                                     _fptr(g), ctypes.byref(loss)) != 0:
        raise stokes.StokesError(stokes.last_error())
    return float(loss.value), g


def mlp_train(params, x, t, epochs: int, batch: int, lr: float, seed: int = 0,
              shape=(NIN, NHID, NLAYERS, NOUT)):
    """Adam on the GPU; returns (trained params, per-epoch mean loss)."""
    x, t = _f32(x), _f32(t)
    p = _f32(params).copy()
    hist = np.zeros(epochs)
    if _bind().erosion_mlp_train(*shape, _fptr(p), x.shape[0], _fptr(x), _fptr(t), int(epochs),
                                 int(batch), float(lr), int(seed),
                                 hist.ctypes.data_as(_dpp)) != 0:
        raise stokes.StokesError(stokes.last_error())
    return p, hist


# ------------------------------------------------------------ features
def _modes(theta: np.ndarray, kmax: int) -> np.ndarray:
    """Fourier modes 1..kmax (real, imaginary) of theta - alpha, scaled by 1/n."""
    n = len(theta)
    c = np.fft.fft(theta - getalpha(n)) / n
    return np.concatenate([[c[k].real, c[k].imag] for k in range(1, kmax + 1)])


def body_features(thlenden: ThLenDenType) -> np.ndarray:
    """[bodies, NIN] features of every body of a state."""
    tv = thlenden.thlenvec
    m = len(tv)
    F = np.zeros((m, NIN))
    if m == 0:
        return F
    n = len(tv[0].theta)
    alpha = getalpha(n)
    own = np.array([_modes(t.theta, NMODES) for t in tv])
    nb_desc = own[:, :2 * NMODES_NB]
    mean = np.array([np.mean(t.theta - alpha) for t in tv])
    ln = np.array([t.len for t in tv])
    ctr = np.array([[t.xsm, t.ysm] for t in tv])
    F[:, :2 * NMODES] = own
    F[:, 2 * NMODES] = np.cos(mean)
    F[:, 2 * NMODES + 1] = np.sin(mean)
    F[:, 2 * NMODES + 2] = ln
    F[:, 2 * NMODES + 3:2 * NMODES + 5] = ctr
    d2 = np.sum((ctr[:, None, :] - ctr[None, :, :]) ** 2, axis=2)
    np.fill_diagonal(d2, np.inf)
    order = np.argsort(d2, axis=1)[:, :NNEIGH]
    base = 2 * NMODES + 5
    w = 3 + 2 * NMODES_NB
    for b in range(m):
        for j in range(min(NNEIGH, m - 1)):
            o = order[b, j]
            s = base + j * w
            F[b, s:s + 2] = ctr[o] - ctr[b]
            F[b, s + 2] = ln[o]
            F[b, s + 3:s + w] = nb_desc[o]
    return F


# ------------------------------------------------------------ targets
def _basis(n: int) -> np.ndarray:
    """[n, 2 NMODES + 1] real Fourier basis: 1, cos k j 2pi/n, sin k j 2pi/n."""
    j = 2 * np.pi * np.arange(n) / n
    cols = [np.ones(n)]
    for k in range(1, NMODES + 1):
        cols += [np.cos(k * j), np.sin(k * j)]
    return np.array(cols).T


def density_targets(den: np.ndarray, n: int, m: int, no: int, speed: np.ndarray) -> np.ndarray:
    """[bodies, NOUT]: least-squares Fourier coefficients of each body's
    density (eta_x, eta_y) times its speed, and its stokeslet/rotlet."""
    B = _basis(n)
    P = np.linalg.pinv(B)
    T = np.zeros((m, NOUT))
    nf = B.shape[1]
    for b in range(m):
        o = 2 * no + 2 * b * n
        T[b, :nf] = P @ den[o:o + n] * speed[b]
        T[b, nf:2 * nf] = P @ den[o + n:o + 2 * n] * speed[b]
        T[b, 2 * nf:] = den[2 * no + 2 * n * m + 3 * b:2 * no + 2 * n * m + 3 * b + 3]
    return T


def assemble_density(T: np.ndarray, n: int, m: int, no: int, wall: np.ndarray,
                     speed: np.ndarray) -> np.ndarray:
    """Density vector from per-body coefficients and a wall density."""
    B = _basis(n)
    nf = B.shape[1]
    den = np.zeros(2 * no + 2 * n * m + 3 * m)
    den[:2 * no] = wall
    for b in range(m):
        o = 2 * no + 2 * b * n
        den[o:o + n] = B @ T[b, :nf] / speed[b]
        den[o + n:o + 2 * n] = B @ T[b, nf:2 * nf] / speed[b]
        den[2 * no + 2 * n * m + 3 * b:2 * no + 2 * n * m + 3 * b + 3] = T[b, 2 * nf:]
    return den


def body_speeds(thlenden: ThLenDenType) -> np.ndarray:
    """|z'(alpha)| = L / 2 pi of every body (equal-arclength points)."""
    return np.array([t.len / (2 * np.pi) for t in thlenden.thlenvec])


def wall_rhs(no: int, m: int, n: int, ibc: int = 0) -> np.ndarray:
    """Right-hand side of the solve: wall rows (1 - y^2, 0) for ibc = 0 or
    (1, 0) for ibc = 1, zero elsewhere (bd_condition, stokesSolver.f:208-314)."""
    th = 2 * np.pi * np.arange(no) / no
    y = (np.cos(th) ** 8 + np.sin(th) ** 8) ** (-1 / 8) * np.sin(th)
    b = np.zeros(2 * no + 2 * n * m + 3 * m)
    b[:no] = 1.0 - y * y if ibc == 0 else 1.0
    return b


# ------------------------------------------------------------ the guess
class LearnedGuess:
    """GMRES initial guess from the trained network, in the ``guess(thlenden)``
    form that ``time_stepping.rungekutta2`` accepts."""

    def __init__(self, path: str = WEIGHTS, nouter: Optional[int] = None, ibc: int = 0):
        z = np.load(path, allow_pickle=False)
        self.params = _f32(z["params"])
        self.xmean, self.xstd = z["xmean"], z["xstd"]
        self.ymean, self.ystd = z["ymean"], z["ystd"]
        self.npts = int(z["npts"])
        self.ibc = ibc
        self._wall: Dict[int, np.ndarray] = {}
        self.nouter = nouter

    def wall_density(self, no: int) -> np.ndarray:
        """Density of the wall alone (no bodies), one solve per wall size."""
        if no not in self._wall:
            den, _ = stokes.stokessolver(self.npts, 0, no, 1, 1, self.ibc, 4000,
                                         np.zeros(0), np.zeros(0))
            self._wall[no] = den[:2 * no].copy()
        return self._wall[no]

    def predict(self, thlenden: ThLenDenType, no: int, scale: bool = True) -> Optional[np.ndarray]:
        """The network's density; with ``scale`` multiplied by the factor
        alpha minimising ||b - alpha A guess|| (one operator application), so
        the starting residual is never above that of zeros."""
        tv = thlenden.thlenvec
        if not tv or len(tv[0].theta) != self.npts:
            return None
        F = body_features(thlenden)
        Y = mlp_forward(self.params, (F - self.xmean) / self.xstd)
        T = Y.astype(np.float64) * self.ystd + self.ymean
        n, m = self.npts, len(tv)
        den = assemble_density(T, n, m, no, self.wall_density(no), body_speeds(thlenden))
        if scale:
            _, xx, yy = getnxy(thlenden)
            b = wall_rhs(no, m, n, self.ibc)
            ag = stokes.matvec(n, m, no, xx, yy, den)
            aa = float(ag @ ag)
            den = den * (float(ag @ b) / aa if aa > 0.0 else 0.0)
        return den

    def __call__(self, thlenden: ThLenDenType) -> Optional[np.ndarray]:
        if self.nouter is None:
            return None
        return self.predict(thlenden, self.nouter)


class WarmStart:
    """The initial guess a time stepper uses: extrapolated from the most
    recent densities while the body count is unchanged, the trained network's
    guess otherwise (the first solve of a run, after a deletion), zeros
    without weights."""

    def __init__(self, nouter: int, path: str = WEIGHTS, ibc: int = 0, order: int = 2):
        from .time_stepping import PreviousDensity
        self.prev = PreviousDensity(order, nouter)
        self.nouter = nouter
        self.net = LearnedGuess(path, nouter=nouter, ibc=ibc) if os.path.exists(path) else None
        self.learned_used = 0

    def __call__(self, thlenden: ThLenDenType) -> Optional[np.ndarray]:
        g = self.prev(thlenden)
        if g is None and self.net is not None:
            g = self.net(thlenden)
            if g is not None:
                self.learned_used += 1
        return g

    def seen(self, thlenden: ThLenDenType) -> None:
        self.prev.seen(thlenden)
