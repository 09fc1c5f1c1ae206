"""Train the learned GMRES warm start on the GPU (python -m erosion_amd.train_warmstart).

1. Snapshots: erosion runs of seeded 100-body x 256-point geometries (the
   100-body workload's generator with other seeds), every density solve at
   1e-10 recorded with its theta-L state: features and Fourier targets per
   body (erosion_amd.learned).  The solves use the preconditioned path, which
   returns the plain solver's density to the tolerance.
2. Training: standardised inputs and outputs, Adam on the GPU with the
   library's BF16 forward/backward kernels (erosion_mlp_train).
3. Writes erosion_amd/data/warmstart_mlp.npz (parameters, normalisation,
   loss history, the seeds used) and prints held-out iteration counts.
"""
from __future__ import annotations

import argparse
import copy
import os
import sys
import time

import numpy as np

from . import learned as LR
from . import stokes
from .geometry import make_bodies
from .theta_len import ParamSet, getnxy, new_thlenden
from .time_stepping import PreviousDensity, rungekutta2


class Collector(PreviousDensity):
    """Previous-density guess that also records (features, targets) of every solve."""

    def __init__(self, n: int, no: int):
        super().__init__()
        self.n, self.no = n, no
        self.X, self.T = [], []

    def seen(self, thld) -> None:
        super().seen(thld)
        m = len(thld.thlenvec)
        # a deletion after the solve leaves fewer bodies than the density has
        if m and len(thld.density) == 2 * self.no + 2 * self.n * m + 3 * m:
            self.X.append(LR.body_features(thld))
            self.T.append(LR.density_targets(np.asarray(thld.density), self.n, m, self.no,
                                             LR.body_speeds(thld)))


def generate(seeds, steps, n=256, no=2048, log=print):
    col = Collector(n, no)
    params = ParamSet(npts=n, nouter=no, ibary=1, ifmm=1, ibc=0, dt=1e-3, maxl=8000, fixpdrop=True)
    for sd in seeds:
        th = new_thlenden(make_bodies(100, n, 0.02, 0.05, 0.02, sd))
        t0 = time.time()
        for _ in range(steps):
            th = rungekutta2(th, params, guess=col)
        log(f"seed {sd}: {steps} steps in {time.time() - t0:.1f} s, samples so far "
            f"{sum(len(x) for x in col.X)}")
    return np.concatenate(col.X), np.concatenate(col.T)


def held_out_iterations(weights, seeds, n=256, no=2048, pc=1, nbodies=100,
                        radii=(0.02, 0.05, 0.02), log=print):
    """GMRES iterations at 1e-10 from zeros, from the learned guess, and from
    the learned guess scaled by the factor that minimises its residual; plus
    the relative error and starting residual of the guess."""
    g = LR.LearnedGuess(weights)
    out = []
    stokes.set_preconditioner(pc)
    for sd in seeds:
        th = new_thlenden(make_bodies(nbodies, n, radii[0], radii[1], radii[2], sd))
        m, xx, yy = getnxy(th)
        d0, it0 = stokes.stokessolver(n, m, no, 1, 1, 0, 8000, xx, yy)
        guess = g.predict(th, no, scale=False)
        _, it1 = stokes.stokessolver(n, m, no, 1, 1, 0, 8000, xx, yy, den=guess)
        b = LR.wall_rhs(no, m, n)
        ag = stokes.matvec(n, m, no, xx, yy, guess)
        alpha = float(ag @ b / (ag @ ag))
        _, it2 = stokes.stokessolver(n, m, no, 1, 1, 0, 8000, xx, yy, den=alpha * guess)
        err = float(np.linalg.norm(guess - d0) / np.linalg.norm(d0))
        r0 = float(np.linalg.norm(b - ag) / np.linalg.norm(b))
        out.append(dict(seed=sd, bodies=m, preconditioner=pc, zeros=it0, learned=it1,
                        learned_scaled=it2, guess_relerr=err, guess_relres=r0, alpha=alpha))
        log(f"held-out seed {sd}, {m} bodies (preconditioner {pc}): zeros {it0}, learned {it1}, "
            f"scaled learned {it2} iterations; guess error {err:.3f}, residual {r0:.3f}, "
            f"alpha {alpha:.3f}")
    stokes.set_preconditioner(0)
    return out


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--seeds", type=int, nargs="+", default=list(range(2100, 2110)))
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--epochs", type=int, default=60)
    ap.add_argument("--batch", type=int, default=512)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--out", default=LR.WEIGHTS)
    ap.add_argument("--held-out", type=int, nargs="+", default=[2200, 2201])
    ap.add_argument("--bodies400", action="store_true",
                    help="also evaluate on a 400-body x 256-point geometry")
    a = ap.parse_args(argv)

    def log(*s):
        print(*s, file=sys.stderr, flush=True)

    stokes.set_gmres_tol(1e-10)
    stokes.set_preconditioner(1)
    stokes.set_precision(stokes.PRECISION_MIXED)
    t0 = time.time()
    X, T = generate(a.seeds, a.steps, log=log)
    stokes.set_precision(stokes.PRECISION_FP64)
    log(f"{len(X)} body samples from {len(a.seeds)} runs in {time.time() - t0:.0f} s")
    xmean, xstd = X.mean(0), X.std(0)
    xstd[xstd < 1e-12] = 1.0
    ymean, ystd = T.mean(0), T.std(0)
    ystd[ystd < 1e-12] = 1.0
    Xn, Tn = (X - xmean) / xstd, (T - ymean) / ystd
    p0 = LR.init_params(seed=7)
    t0 = time.time()
    p, hist = LR.mlp_train(p0, Xn, Tn, a.epochs, a.batch, a.lr, seed=11)
    log(f"trained {a.epochs} epochs in {time.time() - t0:.1f} s; loss {hist[0]:.4f} -> {hist[-1]:.4f}")
    os.makedirs(os.path.dirname(os.path.abspath(a.out)), exist_ok=True)
    np.savez(a.out, params=p, xmean=xmean, xstd=xstd, ymean=ymean, ystd=ystd, npts=256,
             loss=hist, nsamples=len(X), seeds=np.array(a.seeds), steps=a.steps)
    res = held_out_iterations(a.out, a.held_out, pc=1, log=log)
    res += held_out_iterations(a.out, a.held_out, pc=0, log=log)
    if a.bodies400:
        # 400 bodies x 256 points, wall 4096 (the larger learned-guess workload)
        res += held_out_iterations(a.out, [1005], no=4096, pc=1, nbodies=400,
                                   radii=(0.015, 0.035, 0.015), log=log)
    keys = ["seed", "bodies", "preconditioner", "zeros", "learned", "learned_scaled",
            "guess_relerr", "guess_relres", "alpha"]
    np.savez(a.out, params=p, xmean=xmean, xstd=xstd, ymean=ymean, ystd=ystd, npts=256,
             loss=hist, nsamples=len(X), seeds=np.array(a.seeds), steps=a.steps,
             eval_keys=np.array(keys), eval=np.array([[r[k] for k in keys] for r in res], dtype=float))
    return 0


if __name__ == "__main__":
    sys.exit(main())
