"""Learned GMRES initial guess for the density solve (no counterpart in the
reference, whose solves start from zeros, ErosionPackage/src/DensityStress.jl:45).

The guess is made in the preconditioned variable u = P x of the right
preconditioner (precond.hip), not in the density x.  The density of a small
body is huge and almost all rigid rotation (about 10^3 to 10^5 times the
right-hand side), so a density guess with a relative error of 1e-3 still
leaves a starting residual near ||b||.  u is O(||b||) instead: on a body's
rows A x = b gives P x = -(the velocity the other curves induce there) up to
the high-mode part of the body's own block, a smooth flow along the body.  A
guess u0 enters as x0 = P^{-1} u0, where P^{-1} supplies the large rigid part
exactly.

A multilayer perceptron (3 hidden layers x 256 tanh units, BF16 products with
FP32 accumulation on the matrix cores, libstokes_hip.so's erosion_mlp_*) maps
each body's features to that body's part of u in Fourier coefficients:

  features (128): the body's theta-L descriptor -- Fourier modes 1..16 of
    theta - alpha (32 values), cos and sin of its mean (the orientation), the
    length L and the centre (xsm, ysm) -- and, for each of its 8 nearest
    neighbours by centre distance, the relative centre (2), L (1) and modes
    1..4 of theta - alpha (8); zero padded to 128.
  output (69): Fourier coefficients 0..16 (cosine, sine) of u_x and u_y
    along the body's points (66) and u's three constraint entries (3).

The wall part of the guess is u of the same wall with no bodies (one solve,
kept per wall size).  The guess only changes the GMRES iteration count: every
solve still meets the same tolerance, and zeros remain the fallback (for
bodies whose point count the network was not trained on).
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


def solution_targets(u: np.ndarray, n: int, m: int, no: int) -> np.ndarray:
    """[bodies, NOUT]: least-squares Fourier coefficients of each body's part
    of u (u_x, u_y along its points) and its three constraint entries."""
    B = _basis(n)
    P = np.linalg.pinv(B)
    T = np.zeros((m, NOUT))
    nf = B.shape[1]
    for b in range(m):
        o = 2 * no + 2 * b * n
        T[b, :nf] = P @ u[o:o + n]
        T[b, nf:2 * nf] = P @ u[o + n:o + 2 * n]
        T[b, 2 * nf:] = u[2 * no + 2 * n * m + 3 * b:2 * no + 2 * n * m + 3 * b + 3]
    return T


def assemble_solution(T: np.ndarray, n: int, m: int, no: int, wall: np.ndarray) -> np.ndarray:
    """u vector from per-body coefficients and a wall part."""
    B = _basis(n)
    nf = B.shape[1]
    u = np.zeros(2 * no + 2 * n * m + 3 * m)
    u[:2 * no] = wall
    for b in range(m):
        o = 2 * no + 2 * b * n
        u[o:o + n] = B @ T[b, :nf]
        u[o + n:o + 2 * n] = B @ T[b, nf:2 * nf]
        u[2 * no + 2 * n * m + 3 * b:2 * no + 2 * n * m + 3 * b + 3] = T[b, 2 * nf:]
    return u


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
    """GMRES initial guess u0 from the trained network, in the ``guess(thlenden)``
    form that ``time_stepping.rungekutta2`` and ``rungekutta2_device`` accept.
    The solve must take it in the preconditioned variable
    (``stokes.set_guess_space(stokes.GUESS_PRECONDITIONED)``; the device
    stepper always does) with the preconditioner on."""

    def __init__(self, path: str = WEIGHTS, nouter: Optional[int] = None, ibc: int = 0):
        z = np.load(path, allow_pickle=False)
        self.params = _f32(z["params"])
        self.xmean, self.xstd = z["xmean"], z["xstd"]
        self.ymean, self.ystd = z["ymean"], z["ystd"]
        self.npts = int(z["npts"])
        self.ibc = ibc
        self._wall: Dict[int, np.ndarray] = {}
        self.nouter = nouter

    def wall_solution(self, no: int) -> np.ndarray:
        """u of the wall alone (no bodies), one solve per wall size."""
        if no not in self._wall:
            old = stokes.get_guess_space()
            stokes.set_guess_space(stokes.GUESS_PRECONDITIONED)
            try:
                stokes.stokessolver(self.npts, 0, no, 1, 1, self.ibc, 4000, np.zeros(0), np.zeros(0))
                self._wall[no] = stokes.last_preconditioned(self.npts, 0, no)[:2 * no].copy()
            finally:
                stokes.set_guess_space(old)
        return self._wall[no]

    def predict(self, thlenden: ThLenDenType, no: int, scale: bool = True) -> Optional[np.ndarray]:
        """The network's u0 for every body of the state, with the wall part;
        with ``scale`` multiplied by the factor alpha that minimises
        ||b - alpha A P^{-1} u0|| (one preconditioner and one operator
        application), so the starting residual is never above that of zeros.

        Each body's Fourier modes 0 and 1 of u (its mean and linear ambient
        flow) set its large rigid-motion density through P^{-1}; they have to
        be right to about 1e-3 of their spread over the bodies before the
        starting residual falls well below ||b||, an accuracy this network does
        not reach."""
        tv = thlenden.thlenvec
        if not tv or len(tv[0].theta) != self.npts:
            return None
        F = body_features(thlenden)
        Y = mlp_forward(self.params, (F - self.xmean) / self.xstd)
        T = Y.astype(np.float64) * self.ystd + self.ymean
        n, m = self.npts, len(tv)
        u0 = assemble_solution(T, n, m, no, self.wall_solution(no))
        if scale:
            _, xx, yy = getnxy(thlenden)
            ag = stokes.matvec(n, m, no, xx, yy, stokes.precond_apply(n, m, no, xx, yy, u0))
            b = wall_rhs(no, m, n, self.ibc)
            aa = float(ag @ ag)
            u0 = u0 * (float(ag @ b) / aa if aa > 0.0 else 0.0)
        return u0

    def __call__(self, thlenden: ThLenDenType) -> Optional[np.ndarray]:
        if self.nouter is None:
            return None
        return self.predict(thlenden, self.nouter)


class PreviousSolution:
    """Guess from the solutions of the most recent solves, in the
    preconditioned variable: u_k (``order`` 1) or 2 u_k - u_(k-1) (``order``
    2), while the body count is unchanged.  ``seen`` reads u of the solve
    just made (stokes.last_preconditioned), so the solves must run with the
    preconditioned guess space."""

    def __init__(self, npts: int, nouter: int, order: int = 2):
        if order not in (1, 2):
            raise ValueError("order must be 1 or 2")
        self.npts, self.nouter, self.order = npts, nouter, order
        self.hist, self.hist_m = [], -1

    def __call__(self, thld: ThLenDenType) -> Optional[np.ndarray]:
        if not self.hist or self.hist_m != len(thld.thlenvec):
            return None
        if self.order == 2 and len(self.hist) == 2:
            return 2.0 * self.hist[-1] - self.hist[-2]
        return self.hist[-1]

    def seen(self, thld: ThLenDenType) -> None:
        den = np.asarray(thld.density)
        if len(den) == 0:
            return
        m = (len(den) - 2 * self.nouter) // (2 * self.npts + 3)
        u = stokes.last_preconditioned(self.npts, m, self.nouter)
        if m != len(thld.thlenvec):  # bodies deleted since this solve
            self.hist, self.hist_m = [], -1
            return
        if m != self.hist_m:
            self.hist = []
        self.hist = (self.hist + [u])[-2:]
        self.hist_m = m


class WarmStart:
    """The initial guess a host time stepper uses, in the preconditioned
    variable: extrapolated from the most recent solutions while the body count
    is unchanged, the trained network's guess otherwise (the first solve of a
    run, after a deletion), zeros without weights."""

    def __init__(self, nouter: int, npts: int = 256, path: str = WEIGHTS, ibc: int = 0,
                 order: int = 2):
        self.prev = PreviousSolution(npts, nouter, order)
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
