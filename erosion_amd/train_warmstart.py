"""Train the learned GMRES warm start on the GPU (python -m erosion_amd.train_warmstart).

1. Snapshots: erosion runs of seeded 100-body x 256-point geometries (the
   100-body workload's generator with other seeds).  Every density solve at
   1e-10 (preconditioned, mixed precision, guesses in the preconditioned
   variable) is recorded with its theta-L state: features per body and the
   Fourier coefficients of the body's part of u = P x (erosion_amd.learned).
2. Training: standardised inputs and outputs, Adam on the GPU with the
   library's BF16 forward/backward kernels (erosion_mlp_train).
3. Writes erosion_amd/data/warmstart_mlp.npz (parameters, normalisation,
   loss history, the seeds used) and records held-out iteration counts: from
   zeros and from the network's guess, preconditioned.
"""
from __future__ import annotations

import argparse
import os
import sys
import time

import numpy as np

from . import learned as LR
from . import stokes
from .geometry import make_bodies
from .theta_len import ParamSet, getnxy, new_thlenden
from .time_stepping import rungekutta2


class Collector(LR.PreviousSolution):
    """Extrapolated guess that also records (features, targets) of every solve."""

    def __init__(self, n: int, no: int):
        super().__init__(n, no, order=2)
        self.X, self.T = [], []

    def seen(self, thld) -> None:
        super().seen(thld)
        m = len(thld.thlenvec)
        # a deletion after the solve leaves fewer bodies than the solution has
        if m and self.hist and self.hist_m == m:
            self.X.append(LR.body_features(thld))
            self.T.append(LR.solution_targets(self.hist[-1], self.npts, m, self.nouter))


def generate(seeds, steps, n=256, no=2048, log=print):
    col = Collector(n, no)
    params = ParamSet(npts=n, nouter=no, ibary=1, ifmm=1, ibc=0, dt=1e-3, maxl=8000, fixpdrop=True)
    for sd in seeds:
        th = new_thlenden(make_bodies(100, n, 0.02, 0.05, 0.02, sd))
        col.hist, col.hist_m = [], -1
        t0 = time.time()
        for _ in range(steps):
            th = rungekutta2(th, params, guess=col)
        log(f"seed {sd}: {steps} steps in {time.time() - t0:.1f} s, samples so far "
            f"{sum(len(x) for x in col.X)}")
    return np.concatenate(col.X), np.concatenate(col.T)


def held_out_iterations(weights, seeds, n=256, no=2048, nbodies=100, radii=(0.02, 0.05, 0.02),
                        log=print):
    """Preconditioned GMRES at 1e-10 from zeros and from the network's guess
    (both in the preconditioned variable): iterations and the relative
    residual of the starting guess."""
    g = LR.LearnedGuess(weights)
    out = []
    for sd in seeds:
        th = new_thlenden(make_bodies(nbodies, n, radii[0], radii[1], radii[2], sd))
        m, xx, yy = getnxy(th)
        _, it0 = stokes.stokessolver(n, m, no, 1, 1, 0, 8000, xx, yy)
        r00 = stokes.last_solve_info()["r0_relres"]
        _, it1 = stokes.stokessolver(n, m, no, 1, 1, 0, 8000, xx, yy, den=g.predict(th, no))
        r01 = stokes.last_solve_info()["r0_relres"]
        out.append(dict(seed=sd, bodies=m, zeros=it0, learned=it1, r0_zeros=r00, r0_learned=r01))
        log(f"held-out seed {sd}, {m} bodies: zeros {it0} iterations (r0 {r00:.3f}), learned "
            f"{it1} (r0 {r01:.3e})")
    return out


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--seeds", type=int, nargs="+", default=list(range(2100, 2110)))
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--epochs", type=int, default=200)
    ap.add_argument("--batch", type=int, default=512)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--out", default=LR.WEIGHTS)
    ap.add_argument("--held-out", type=int, nargs="+", default=[2200, 2201])
    ap.add_argument("--data", default=None,
                    help="npz of (X, T) snapshots to reuse instead of generating them")
    ap.add_argument("--save-data", default=None, help="write the snapshots to this npz")
    a = ap.parse_args(argv)

    def log(*s):
        print(*s, file=sys.stderr, flush=True)

    stokes.set_gmres_tol(1e-10)
    stokes.set_preconditioner(stokes.PRECOND_CURVE_FOURIER)
    stokes.set_precision(stokes.PRECISION_MIXED)
    stokes.set_guess_space(stokes.GUESS_PRECONDITIONED)
    try:
        if a.data:
            with np.load(a.data, allow_pickle=False) as z:
                X, T = z["X"], z["T"]
        else:
            t0 = time.time()
            X, T = generate(a.seeds, a.steps, log=log)
            log(f"{len(X)} body samples from {len(a.seeds)} runs in {time.time() - t0:.0f} s")
            if a.save_data:
                np.savez(a.save_data, X=X, T=T)
        xmean, xstd = X.mean(0), X.std(0)
        xstd[xstd < 1e-12] = 1.0
        ymean, ystd = T.mean(0), T.std(0)
        ystd[ystd < 1e-12] = 1.0
        Xn, Tn = (X - xmean) / xstd, (T - ymean) / ystd
        p0 = LR.init_params(seed=7)
        t0 = time.time()
        p, hist = LR.mlp_train(p0, Xn, Tn, a.epochs, a.batch, a.lr, seed=11)
        log(f"trained {a.epochs} epochs in {time.time() - t0:.1f} s; loss {hist[0]:.4f} -> "
            f"{hist[-1]:.4f}")
        os.makedirs(os.path.dirname(os.path.abspath(a.out)), exist_ok=True)
        meta = dict(params=p, xmean=xmean, xstd=xstd, ymean=ymean, ystd=ystd, npts=256,
                    loss=hist, nsamples=len(X), seeds=np.array(a.seeds), steps=a.steps,
                    target=np.array("preconditioned"))
        np.savez(a.out, **meta)
        res = held_out_iterations(a.out, a.held_out, log=log)
        keys = ["seed", "bodies", "zeros", "learned", "r0_zeros", "r0_learned"]
        np.savez(a.out, **meta, eval_keys=np.array(keys),
                 eval=np.array([[r[k] for k in keys] for r in res], dtype=float))
    finally:
        stokes.set_guess_space(stokes.GUESS_DENSITY)
        stokes.set_precision(stokes.PRECISION_FP64)
        stokes.set_preconditioner(stokes.PRECOND_NONE)
    return 0


if __name__ == "__main__":
    sys.exit(main())
