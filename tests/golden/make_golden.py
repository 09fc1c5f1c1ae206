"""Write golden vectors from the reference's own Fortran.

Runs only in a container that holds /root/reference: ``make -C oracle ref``
builds oracle/_ref/libstokes_ref.so from the reference sources with flang,
and this script calls its routines through oracle/reference.py on seeded
inputs, saving inputs and outputs as small .npz fixtures next to this file.

    python tests/golden/make_golden.py                  # every fixture
    python tests/golden/make_golden.py --pressure-only  # add computePressure

Fixtures (all FP64):
  bodies10_n128.npz  10 bodies x 128, wall 512 (bodies far apart: the
                     near-pair branch of the operator stays inactive)
  tight8_n64.npz     8 bodies x 64, wall 256, gaps 0.004 and bodies near the
                     wall: wall->body, body->wall and body->body barycentric
                     replacements are all active
  dlp_n64.npz        StokesExteriorDLP / StokesInteriorDLP with nder=2 at
                     targets from 1e-3 to 0.3 off the curve, plus one target
                     1e-11 from a node (the coincident-node branch)
"""
from __future__ import annotations

import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from erosion_amd.geometry import WORKLOADS, make_bodies, workload_xy  # noqa: E402
from erosion_amd.theta_len import getnxy, new_thlenden  # noqa: E402
from oracle import reference as R  # noqa: E402


def solver_case(name, n, m, no, xx, yy, ibc, maxl, seed):
    rng = np.random.default_rng(seed)
    N = 2 * no + 2 * n * m + 3 * m
    eta = rng.standard_normal(N)
    out = dict(n=n, m=m, no=no, ibc=ibc, xx=xx, yy=yy, eta=eta)
    geo = R.geometry_arrays(n, m, no, xx, yy)
    out.update({"geo_" + k: v for k, v in geo.items()})
    out["rhs"] = R.rhs(n, m, no, xx, yy, ibc)
    out["mv_trap"] = R.matvec(n, m, no, xx, yy, eta, ibary=0, ifmm=0)  # matvec_DLP
    out["mv_fmmbary"] = R.matvec(n, m, no, xx, yy, eta, ibary=1, ifmm=1)  # matvec_DLP_fmmbary
    # shipped entry point: tol 1e-8 hard-coded
    den8, it8 = R.stokessolver(n, m, no, xx, yy, ifmm=1, ibary=1, ibc=ibc, maxl=maxl)
    out["den_tol8"], out["iter_tol8"] = den8, it8
    # DGMRES with solveBIE's settings and a tight tolerance
    den12, it12, err12, ierr12 = R.solve(n, m, no, xx, yy, ibary=1, ifmm=1, ibc=ibc, maxl=maxl,
                                         tol=1e-12)
    out["den_tol12"], out["iter_tol12"], out["err_tol12"] = den12, it12, err12
    out["tau_tol12"] = R.shear_stress(n, m, no, xx, yy, den12)
    # ibary=0 solve (trapezoid operator) to a tight tolerance
    den0, it0, _e, _ie = R.solve(n, m, no, xx, yy, ibary=0, ifmm=0, ibc=ibc, maxl=maxl, tol=1e-12)
    out["den_trap_tol12"], out["iter_trap_tol12"] = den0, it0
    # computeQoiTargets at the pressure-drop slices x = +-2 (DensityStress.jl:131-136),
    # at every body centre (inside: zeroed) and just outside two bodies
    dy = 2.0 / 13
    ys = -1 + 0.5 * dy + dy * np.arange(13)
    geo_cx, geo_cy = out["geo_cx"], out["geo_cy"]
    xt = np.concatenate([np.full(13, 2.0), np.full(13, -2.0), geo_cx,
                         xx[[0, n]] - 0.01 * out["geo_px"][[0, n]]])
    yt = np.concatenate([ys, ys, geo_cy, yy[[0, n]] - 0.01 * out["geo_py"][[0, n]]])
    out["qoi_xt"], out["qoi_yt"] = xt, yt
    out["qoi_bary"] = np.array(R.qoi_targets(n, m, no, 1, xx, yy, den12, xt, yt))
    out["qoi_trap"] = np.array(R.qoi_targets(n, m, no, 0, xx, yy, den12, xt, yt))
    # computePressure on the body points (PostProcess.jl:107-110)
    out["pressure_tol12"] = R.pressure(n, m, no, xx, yy, den12)
    np.savez_compressed(os.path.join(HERE, name + ".npz"), **out)
    print(f"{name}: N={N} iters tol8={it8} tol12={it12} trap={it0}")


def dlp_case(name, seed):
    thl = make_bodies(1, 64, 0.2, 0.2, 0.05, seed)
    m, xs, ys = getnxy(new_thlenden(thl))
    g = R.geometry_arrays(64, 1, 256, xs, ys)
    px, py = g["px"], g["py"]
    rng = np.random.default_rng(seed)
    dx, dy = rng.standard_normal(64), rng.standard_normal(64)
    # targets along the normal at several distances, both sides of the curve;
    # px, py is the body's inward normal (stokesSolver.f:449-450)
    j = np.arange(0, 64, 5)
    offs = np.array([1e-3, 1e-2, 5e-2, 0.3])
    xe = np.concatenate([xs[j] - o * px[j] for o in offs])
    ye = np.concatenate([ys[j] - o * py[j] for o in offs])
    xi = np.concatenate([xs[j] + o * 0.5 * px[j] for o in offs[:3]])
    yi = np.concatenate([ys[j] + o * 0.5 * py[j] for o in offs[:3]])
    # one exterior target 1e-11 from node 7: exercises :5492-5506
    xc = np.array([xs[7] - 1e-11 * px[7]])
    yc = np.array([ys[7] - 1e-11 * py[7]])
    ext = R.exterior_dlp(xs, ys, dx, dy, px, py, xe, ye, nder=2)
    inn = R.interior_dlp(xs, ys, dx, dy, px, py, xi, yi, nder=2)
    coin = R.exterior_dlp(xs, ys, dx, dy, px, py, xc, yc, nder=2)
    np.savez_compressed(os.path.join(HERE, name + ".npz"), xs=xs, ys=ys, dx=dx, dy=dy, px=px, py=py,
                        xe=xe, ye=ye, xi=xi, yi=yi, xc=xc, yc=yc,
                        ext=np.array(ext), inn=np.array(inn), coin=np.array(coin))
    print(f"{name}: {len(xe)} exterior, {len(xi)} interior targets")


def add_pressure(name):
    """Add computePressure of the stored tol-1e-12 density to an existing fixture."""
    path = os.path.join(HERE, name + ".npz")
    with np.load(path, allow_pickle=False) as z:
        out = {k: z[k] for k in z.files}
    n, m, no = int(out["n"]), int(out["m"]), int(out["no"])
    out["pressure_tol12"] = R.pressure(n, m, no, out["xx"], out["yy"], out["den_tol12"])
    np.savez_compressed(path, **out)
    print(f"{name}: pressure added")


def main():
    if not R.available():
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "ref"], check=True)
    if "--pressure-only" in sys.argv:
        for name in ("bodies10_n128", "tight8_n64"):
            add_pressure(name)
        return
    w = WORKLOADS["bodies10_n128"]
    n, m, no, xx, yy = workload_xy(w)
    solver_case("bodies10_n128", n, m, no, xx, yy, ibc=0, maxl=2000, seed=11)
    thl = make_bodies(8, 64, 0.08, 0.16, 0.004, 77)
    m, xx, yy = getnxy(new_thlenden(thl))
    solver_case("tight8_n64", 64, m, 256, xx, yy, ibc=0, maxl=2000, seed=12)
    dlp_case("dlp_n64", 13)


if __name__ == "__main__":
    main()
