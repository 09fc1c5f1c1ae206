"""Throughput of many independent geometries (python -m erosion_amd.bench_batch).

Geometries of uneven size from the seeded generator, seed 4000 + g for
geometry g, each solved to 1e-10 with the accelerated settings
(preconditioner, mixed precision) and its shear stress evaluated: once one
geometry after another on a single context, once through erosion_solve_batch
with worker contexts.  Two size mixes:

  small (default): 16 geometries, bodies U{20..80}, points per body from
                   {64, 128, 256}, wall points U{1024..2048}, radii U[0.02, 0.05],
                   gap 0.02;
  full:            64 geometries, bodies U{20..200}, points per body from
                   {64, 128, 256, 512}, wall points U{1024..4096}, radii
                   U[0.01, 0.03], gap 0.01 (the radii of the 200-body workload, so
                   that 200 bodies fit the unit box).

``--graphs`` replays GMRES iteration chunks as HIP graphs in both runs.
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


MIXES = {
    # bodies, points per body, wall points, radii, gap
    "small": dict(count=16, bodies=(20, 80), points=(64, 128, 256), wall=(1024, 2048),
                  radii=(0.02, 0.05), gap=0.02),
    "full": dict(count=64, bodies=(20, 200), points=(64, 128, 256, 512), wall=(1024, 4096),
                 radii=(0.01, 0.03), gap=0.01),
}


def geometries(count: int, mix: str = "small"):
    spec = MIXES[mix]
    rng = np.random.default_rng(4000)
    out = []
    for g in range(count):
        m = int(rng.integers(spec["bodies"][0], spec["bodies"][1] + 1))
        n = int(rng.choice(spec["points"]))
        no = int(rng.integers(spec["wall"][0], spec["wall"][1] + 1))
        thl = make_bodies(m, n, spec["radii"][0], spec["radii"][1], spec["gap"], 4000 + g)
        _, xx, yy = getnxy(new_thlenden(thl))
        out.append((n, m, no, xx, yy))
    return out


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--mix", default="small", choices=sorted(MIXES))
    ap.add_argument("--count", type=int, default=None, help="geometries (default: the mix's)")
    ap.add_argument("--workers", type=int, default=4)
    ap.add_argument("--maxl", type=int, default=2000)
    ap.add_argument("--graphs", action="store_true")
    a = ap.parse_args(argv)
    count = a.count or MIXES[a.mix]["count"]
    t0 = time.perf_counter()
    geoms = geometries(count, a.mix)
    print(f"{count} geometries generated in {time.perf_counter() - t0:.1f} s", file=sys.stderr,
          flush=True)
    stokes.set_gmres_tol(1e-10)
    stokes.set_preconditioner(stokes.PRECOND_CURVE_FOURIER)
    stokes.set_precision(stokes.PRECISION_MIXED)
    stokes.set_graphs(a.graphs)
    stokes.solve_batch(geoms[:2], maxl=a.maxl, workers=1, stress=True)  # warm-up
    t0 = time.perf_counter()
    seq = stokes.solve_batch(geoms, maxl=a.maxl, workers=1, stress=True)
    t_seq = time.perf_counter() - t0
    print(f"sequential: {t_seq:.2f} s", file=sys.stderr, flush=True)
    stokes.solve_batch(geoms[:a.workers], maxl=a.maxl, workers=a.workers, stress=True)  # warm-up
    t0 = time.perf_counter()
    par = stokes.solve_batch(geoms, maxl=a.maxl, workers=a.workers, stress=True)
    t_par = time.perf_counter() - t0
    print(f"batched: {t_par:.2f} s", file=sys.stderr, flush=True)
    diff = max(float(np.max(np.abs(p["den"] - s["den"])) / np.max(np.abs(s["den"])))
               for p, s in zip(par, seq))
    spec = MIXES[a.mix]
    print(json.dumps({
        "workload": f"{count} geometries, bodies U{{{spec['bodies'][0]}..{spec['bodies'][1]}}}, "
                    f"points {set(spec['points'])}, wall U{{{spec['wall'][0]}..{spec['wall'][1]}}}, "
                    f"radii U{list(spec['radii'])}, gap {spec['gap']}",
        "hip_graphs": a.graphs,
        "points_total": int(sum(g[2] + g[0] * g[1] for g in geoms)),
        "gmres_iters_total": int(sum(r["nits"] for r in par)),
        "sequential_geometries_per_s": count / t_seq,
        "batched_geometries_per_s": count / t_par,
        "workers": a.workers,
        "speedup": t_seq / t_par,
        "density_maxdiff_batched_vs_sequential": diff,
        "status": [r["status"] for r in par],
    }), flush=True)
    stokes.set_preconditioner(stokes.PRECOND_NONE)
    stokes.set_precision(stokes.PRECISION_FP64)
    stokes.set_graphs(False)
    return 0


if __name__ == "__main__":
    sys.exit(main())
