"""Seeded porous-geometry generator and the workload table.

The reference builds its initial bodies from circle files made by
ErosionPackage/src/MakeGeos.jl; the workloads measured here instead use one
seeded generator of perturbed circles:

* radius r0 ~ U[rmin, rmax], perturbed radially as
  r(phi) = r0 (1 + sum_{k=1..8} (amp/k^2)(a_k cos k phi + b_k sin k phi)),
  a_k, b_k ~ N(0, 1), amp = 0.02 (the perturbation is relative to r0);
* centres uniform in the unit box |x|, |y| < 1, accepted only when the
  bounding circles (radius max_phi r) stay ``gap`` apart from each other and
  from the box edge (rejection sampling);
* each curve resampled at the offset equal-arclength grid and stored in the
  theta-L form of ErosionPackage/src/ThetaLen.jl (tangent angle, length,
  surface means).

Random draws happen in a fixed order (per body: r0, a_1..a_8, b_1..b_8, then
centre proposals), so a seed fixes the geometry bit for bit.
"""
from __future__ import annotations

import dataclasses
import os
from typing import Dict, List, Optional, Tuple

import numpy as np

from .theta_len import ThetaLenType, getalpha, getnxy, new_thlenden

NMODES = 8
AMP = 0.02


@dataclasses.dataclass(frozen=True)
class Workload:
    """One measured workload: bodies, points, wall points, radii and gap."""

    name: str
    nbodies: int
    npts: int
    nouter: int
    rmin: float
    rmax: float
    gap: float
    seed: int
    ibc: int = 0
    steps: int = 1
    dt: float = 1e-3
    # a committed theta-L state (erosion_amd/data) the workload starts from
    # instead of the seeded initial geometry
    state: Optional[str] = None


WORKLOADS: Dict[str, Workload] = {
    # 10 smooth bodies x 128 points, wall 512, 20 RK2 steps.
    "bodies10_n128": Workload("bodies10_n128", 10, 128, 512, 0.05, 0.10, 0.05, 1001, steps=20),
    # 100 bodies x 256 points, wall 2048, 500-step erosion run.
    "bodies100_n256": Workload("bodies100_n256", 100, 256, 2048, 0.02, 0.05, 0.02, 1002, steps=500),
    # 200 bodies x 512 points (~1e5 boundary points), wall 4096, 100 steps.
    "bodies200_n512": Workload("bodies200_n512", 200, 512, 4096, 0.01, 0.03, 0.01, 1003, steps=100),
    # 400 bodies x 256 points, wall 4096 (learned-warm-start workload).
    "bodies400_n256": Workload("bodies400_n256", 400, 256, 4096, 0.015, 0.035, 0.015, 1005, steps=1),
    # The 400-body workload eroded until 50% of the total body area is lost
    # (46 RK2 steps, 386 bodies left; python -m erosion_amd.eroded_state):
    # the learned-warm-start workload, 102,912 boundary points.
    "bodies400_n256_eroded50": Workload("bodies400_n256_eroded50", 400, 256, 4096, 0.015, 0.035,
                                        0.015, 1005, steps=20,
                                        state="bodies400_n256_eroded50.npz"),
    # 60 bodies x 128 points packed 0.004 apart, wall 1024: most bodies have
    # near neighbours, so the barycentric near-pair pass does real work.
    "bodies60_n128_tight": Workload("bodies60_n128_tight", 60, 128, 1024, 0.05, 0.10, 0.004, 1006),
}


def _radial(r0: float, a: np.ndarray, b: np.ndarray, phi: np.ndarray):
    """r(phi) and r'(phi) of the perturbed circle."""
    k = np.arange(1, NMODES + 1, dtype=np.float64)[:, None]
    c = (AMP / k[:, 0] ** 2)[:, None]
    cs, sn = np.cos(k * phi[None, :]), np.sin(k * phi[None, :])
    pert = np.sum(c * (a[:, None] * cs + b[:, None] * sn), axis=0)
    dpert = np.sum(c * k * (-a[:, None] * sn + b[:, None] * cs), axis=0)
    return r0 * (1.0 + pert), r0 * dpert


def perturbed_circle_thetalen(npts: int, xc: float, yc: float, r0: float, a: np.ndarray,
                              b: np.ndarray, nfine: int = 2048) -> ThetaLenType:
    """Resample c + r(phi) e^{i phi} at equal arclength; return its theta-L form."""
    phi_f = 2 * np.pi * np.arange(nfine) / nfine
    r, dr = _radial(r0, a, b, phi_f)
    speed = np.sqrt(r * r + dr * dr)
    ghat = np.fft.fft(speed) / nfine
    mean_speed = ghat[0].real
    length = 2 * np.pi * mean_speed
    kk = np.fft.fftfreq(nfine, 1.0 / nfine)
    # modes of the (smooth) speed below 1e-17 of its mean change the
    # arclength by less than its rounding: leave them out of the evaluations
    keep = (kk != 0) & (np.abs(kk) < nfine // 2) & (np.abs(ghat) > 1e-17 * mean_speed)
    kk, gk = kk[keep], ghat[keep]

    def arclen(ph: np.ndarray) -> np.ndarray:
        # s(phi) = mean*phi + G(phi) - G(0), G the mean-free antiderivative
        e = np.exp(1j * kk[None, :] * ph[:, None])
        G = np.real(e @ (gk / (1j * kk)))
        G0 = np.real(np.sum(gk / (1j * kk)))
        return mean_speed * ph + G - G0

    alpha = getalpha(npts)
    s_target = alpha * length / (2 * np.pi)
    ph = alpha.copy()  # initial guess: arclength ~ angle for near-circles
    for _ in range(30):
        rr, drr = _radial(r0, a, b, ph)
        sp = np.sqrt(rr * rr + drr * drr)
        step = (arclen(ph) - s_target) / sp
        ph = ph - step
        if np.max(np.abs(step)) < 1e-15:
            break
    rr, drr = _radial(r0, a, b, ph)
    theta = ph + 0.5 * np.pi - np.arctan(drr / rr)
    x = xc + rr * np.cos(ph)
    y = yc + rr * np.sin(ph)
    return ThetaLenType(theta, float(length), float(np.mean(x)), float(np.mean(y)), float("nan"))


def make_bodies(nbodies: int, npts: int, rmin: float, rmax: float, gap: float, seed: int,
                box: float = 1.0, max_tries: int = 200000) -> List[ThetaLenType]:
    """Seeded rejection-sampled perturbed circles in theta-L form."""
    rng = np.random.default_rng(seed)
    phi_f = 2 * np.pi * np.arange(512) / 512
    bodies: List[Tuple[float, float, float, float, np.ndarray, np.ndarray]] = []
    for _ in range(nbodies):
        r0 = rng.uniform(rmin, rmax)
        a = rng.standard_normal(NMODES)
        b = rng.standard_normal(NMODES)
        rbound = float(np.max(_radial(r0, a, b, phi_f)[0]))
        lo, hi = -box + rbound + gap, box - rbound - gap
        for _t in range(max_tries):
            cx, cy = rng.uniform(lo, hi), rng.uniform(lo, hi)
            if all((cx - q[0]) ** 2 + (cy - q[1]) ** 2 >= (rbound + q[2] + gap) ** 2 for q in bodies):
                break
        else:
            raise RuntimeError(f"make_bodies: could not place body {len(bodies)} after {max_tries} tries")
        bodies.append((cx, cy, rbound, r0, a, b))
    return [perturbed_circle_thetalen(npts, cx, cy, r0, a, b) for (cx, cy, _rb, r0, a, b) in bodies]


def workload_geometry(w: Workload, nbodies: Optional[int] = None):
    """theta-L state of a workload (optionally truncated to fewer bodies)."""
    if w.state is not None:
        from .eroded_state import DATA, load_state
        th = load_state(os.path.join(DATA, w.state))
        if nbodies is not None:
            th.thlenvec = th.thlenvec[:nbodies]
        return th
    thl = make_bodies(w.nbodies, w.npts, w.rmin, w.rmax, w.gap, w.seed)
    if nbodies is not None:
        thl = thl[:nbodies]
    return new_thlenden(thl)


def workload_xy(w: Workload, nbodies: Optional[int] = None):
    """Solver input (n, m, nouter, xx, yy) of a workload."""
    thld = workload_geometry(w, nbodies)
    m, xx, yy = getnxy(thld)
    return w.npts, m, w.nouter, xx, yy
