"""Throughput of many independent geometries (python -m erosion_amd.bench_batch).

Geometries of uneven size from the seeded generator -- bodies U{20..80},
points per body from {64, 128, 256}, wall points U{1024..2048}, seed 4000 + g
-- each solved to 1e-10 with the accelerated settings (preconditioner, mixed
precision) and its shear stress evaluated: once one geometry after another
on a single context, once through erosion_solve_batch with worker contexts.
Prints one JSON line (geometries/s both ways).
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np

from . import stokes
from .geometry import make_bodies
from .theta_len import getnxy, new_thlenden


def geometries(count: int):
    rng = np.random.default_rng(4000)
    out = []
    for g in range(count):
        m = int(rng.integers(20, 81))
        n = int(rng.choice([64, 128, 256]))
        no = int(rng.integers(1024, 2049))
        _, xx, yy = getnxy(new_thlenden(make_bodies(m, n, 0.02, 0.05, 0.02, 4000 + g)))
        out.append((n, m, no, xx, yy))
    return out


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--count", type=int, default=16)
    ap.add_argument("--workers", type=int, default=4)
    a = ap.parse_args(argv)
    geoms = geometries(a.count)
    stokes.set_gmres_tol(1e-10)
    stokes.set_preconditioner(stokes.PRECOND_CURVE_FOURIER)
    stokes.set_precision(stokes.PRECISION_MIXED)
    stokes.solve_batch(geoms[:2], workers=1, stress=True)  # warm-up: tuning, allocation
    t0 = time.perf_counter()
    seq = stokes.solve_batch(geoms, workers=1, stress=True)
    t_seq = time.perf_counter() - t0
    stokes.solve_batch(geoms[:a.workers], workers=a.workers, stress=True)  # warm the workers
    t0 = time.perf_counter()
    par = stokes.solve_batch(geoms, workers=a.workers, stress=True)
    t_par = time.perf_counter() - t0
    diff = max(float(np.max(np.abs(p["den"] - s["den"])) / np.max(np.abs(s["den"])))
               for p, s in zip(par, seq))
    print(json.dumps({
        "workload": f"{a.count} geometries, bodies U{{20..80}}, points {{64,128,256}}, wall U{{1024..2048}}",
        "points_total": int(sum(g[2] + g[0] * g[1] for g in geoms)),
        "gmres_iters_total": int(sum(r["nits"] for r in par)),
        "sequential_geometries_per_s": a.count / t_seq,
        "batched_geometries_per_s": a.count / t_par,
        "workers": a.workers,
        "speedup": t_seq / t_par,
        "density_maxdiff_batched_vs_sequential": diff,
        "status": [r["status"] for r in par],
    }), flush=True)
    stokes.set_preconditioner(stokes.PRECOND_NONE)
    stokes.set_precision(stokes.PRECISION_FP64)
    return 0


if __name__ == "__main__":
    sys.exit(main())
