"""numpy restatement of the per-curve two-level preconditioner (test helper).

The preconditioner has no counterpart in the reference (solveBIE runs GMRES
without one, ErosionPackage/fortran_code/stokesSolver.f:490).  Its blocks are
the diagonal blocks of the completed operator of matvec_DLP_fmmbary
(stokesSolver.f:570-1124) that couple one curve's own unknowns; the tests pin
``body_block`` and ``wall_block`` against columns of the oracle's operator.
"""
from __future__ import annotations

import numpy as np

TWOPI = 2.0 * np.pi


def fourier_basis(N: int, K: int) -> np.ndarray:
    """[N, 2K+1] orthonormal: constant, then cos k theta, sin k theta."""
    th = TWOPI * np.arange(N) / N
    cols = [np.full(N, 1.0 / np.sqrt(N))]
    for k in range(1, K + 1):
        cols.append(np.sqrt(2.0 / N) * np.cos(k * th))
        cols.append(np.sqrt(2.0 / N) * np.sin(k * th))
    return np.array(cols).T


def _self_dlp(x, y, px, py, cur, sp):
    """Trapezoid self-interaction + curvature diagonal + jump of one curve, [2N, 2N]."""
    N = len(x)
    rx = x[:, None] - x[None, :]
    ry = y[:, None] - y[None, :]
    rho2 = rx * rx + ry * ry
    np.fill_diagonal(rho2, 1.0)
    w = (rx * px[None, :] + ry * py[None, :]) / rho2**2 * sp[None, :] * (2.0 / N)
    np.fill_diagonal(w, 0.0)
    B = np.zeros((2 * N, 2 * N))
    B[:N, :N] = w * rx * rx
    B[:N, N:] = w * rx * ry
    B[N:, :N] = w * ry * rx
    B[N:, N:] = w * ry * ry
    k = np.arange(N)
    f = cur * sp / N
    B[k, k] += -f * py * py - 0.5
    B[k, N + k] += f * py * px
    B[N + k, k] += f * px * py
    B[N + k, N + k] += -f * px * px - 0.5
    return B


def body_block(geo, xx, yy, b: int) -> np.ndarray:
    """Rows/columns (eta_x, eta_y of body b, then its stokeslet x, y and rotlet)."""
    n = geo.n
    s = slice(b * n, (b + 1) * n)
    x, y = np.asarray(xx)[s], np.asarray(yy)[s]
    px, py = geo.array("px")[s], geo.array("py")[s]
    cur, sp = geo.array("cur")[s], geo.array("speed")[s]
    cx, cy = geo.array("cx")[b], geo.array("cy")[b]
    B = np.zeros((2 * n + 3, 2 * n + 3))
    B[:2 * n, :2 * n] = _self_dlp(x, y, px, py, cur, sp)
    qx, qy = x - cx, y - cy
    r2 = qx * qx + qy * qy
    lg = -0.5 * np.log(r2)
    c = 0.5 / TWOPI
    B[:n, 2 * n] = c * (lg + qx * qx / r2)
    B[n:2 * n, 2 * n] = c * qx * qy / r2
    B[:n, 2 * n + 1] = c * qx * qy / r2
    B[n:2 * n, 2 * n + 1] = c * (lg + qy * qy / r2)
    B[:n, 2 * n + 2] = qy / r2
    B[n:2 * n, 2 * n + 2] = -qx / r2
    B[2 * n, :n] = -sp / n
    B[2 * n + 1, n:2 * n] = -sp / n
    B[2 * n + 2, :n] = -y * sp / n
    B[2 * n + 2, n:2 * n] = x * sp / n
    B[2 * n:, 2 * n:] += np.eye(3)
    return B


def wall_block(geo) -> np.ndarray:
    """Rows/columns (eta_x, eta_y of the wall), including the rank-one term."""
    no = geo.no
    x, y = geo.array("xo"), geo.array("yo")
    px, py = geo.array("px0"), geo.array("py0")
    B = _self_dlp(x, y, px, py, geo.array("cur0"), geo.array("speed0"))
    sp = geo.array("speed0")
    nv = np.concatenate([px, py])
    wv = np.concatenate([px * sp, py * sp]) * TWOPI / no
    return B + np.outer(nv, wv)


def body_indices(n: int, m: int, no: int, b: int) -> np.ndarray:
    d = 2 * no + 2 * b * n
    c = 2 * no + 2 * n * m + 3 * b
    return np.concatenate([np.arange(d, d + 2 * n), np.arange(c, c + 3)])


def coarse_q(N: int, K: int, constraints: bool) -> np.ndarray:
    F = fourier_basis(N, K)
    nf = F.shape[1]
    extra = 3 if constraints else 0
    Q = np.zeros((2 * N + extra, 2 * nf + extra))
    Q[:N, :nf] = F
    Q[N:2 * N, nf:2 * nf] = F
    if constraints:
        Q[2 * N:, 2 * nf:] = np.eye(3)
    return Q


class TwoLevel:
    """P^{-1} r = Q (Q^T B Q)^{-1} Q^T r - 2 (r - Q Q^T r) on every curve."""

    def __init__(self, geo, xx, yy, body_modes: int = 8, wall_modes: int = 16):
        n, m, no = geo.n, geo.m, geo.no
        self.parts = []
        Qw = coarse_q(no, wall_modes, False)
        self.parts.append((np.arange(2 * no), Qw, np.linalg.inv(Qw.T @ wall_block(geo) @ Qw)))
        if m:
            Qb = coarse_q(n, body_modes, True)
            for b in range(m):
                G = Qb.T @ body_block(geo, xx, yy, b) @ Qb
                self.parts.append((body_indices(n, m, no, b), Qb, np.linalg.inv(G)))

    def __call__(self, r: np.ndarray) -> np.ndarray:
        z = np.array(r, dtype=np.float64, copy=True)
        for idx, Q, Gi in self.parts:
            rb = r[idx]
            c = Q.T @ rb
            z[idx] = Q @ (Gi @ c) - 2.0 * (rb - Q @ c)
        return z
