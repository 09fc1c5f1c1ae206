"""Parity of the HIP library against the oracle and the reference goldens.

Every call goes through the C-ABI of libstokes_hip.so.  Tolerances:
  operator (matvec)  1e-12 relative (max-norm) to the FP64 oracle / reference
  solve              density 1e-10 relative to the reference at tol 1e-12
  shear stress       1e-10 relative to the reference on the same density
"""
import numpy as np
import pytest

from conftest import relerr
from erosion_amd import stokes
from erosion_amd.geometry import WORKLOADS, make_bodies, workload_xy
from erosion_amd.theta_len import getnxy, new_thlenden
from oracle import oracle as O

pytestmark = pytest.mark.gpu
FIXTURES = ["tight8_n64", "bodies10_n128"]


def _case(g):
    return int(g["n"]), int(g["m"]), int(g["no"]), g["xx"], g["yy"]


@pytest.fixture(autouse=True)
def _tol_reset():
    stokes.set_gmres_tol(1e-8)
    yield
    stokes.set_gmres_tol(1e-8)


@pytest.mark.parametrize("name", FIXTURES)
@pytest.mark.parametrize("ibary", [0, 1])
def test_matvec_matches_reference(golden, name, ibary):
    g = golden(name)
    n, m, no, xx, yy = _case(g)
    v = stokes.matvec(n, m, no, xx, yy, g["eta"], ibary=ibary)
    ref = g["mv_fmmbary"] if ibary else g["mv_trap"]
    assert relerr(v, ref) < 1e-12
    assert relerr(v, O.Geometry(n, m, no, xx, yy).matvec(g["eta"], ibary)) < 1e-12


def test_matvec_full_size_vs_oracle():
    """100 bodies x 256 points, wall 2048 (27,648 boundary points).

    Every row is a 27,647-term sum and the wall curvature (a second spectral
    derivative of a 2048-point superellipse) already differs by ~1e-11 between
    any two FP64 derivative routines, so the bar here is 1e-11 max-norm and
    1e-12 RMS."""
    n, m, no, xx, yy = workload_xy(WORKLOADS["bodies100_n256"])
    N = 2 * no + 2 * n * m + 3 * m
    eta = np.random.default_rng(7).standard_normal(N)
    v = stokes.matvec(n, m, no, xx, yy, eta, ibary=1)
    vo = O.Geometry(n, m, no, xx, yy).matvec(eta, 1)
    assert relerr(v, vo) < 1e-11
    assert np.sqrt(np.mean((v - vo) ** 2)) / np.sqrt(np.mean(vo ** 2)) < 1e-12


def test_matvec_every_farfield_shape_vs_oracle():
    """Each far-field launch shape the autotuner may pick (targets per lane,
    grid, shared reciprocal), in FP64, on 20 bodies x 256 points."""
    n, m, no, xx, yy = workload_xy(WORKLOADS["bodies100_n256"], nbodies=20)
    N = 2 * no + 2 * n * m + 3 * m
    eta = np.random.default_rng(11).standard_normal(N)
    vo = O.Geometry(n, m, no, xx, yy).matvec(eta, 1)
    try:
        for k in range(stokes.farfield_shape_count()):
            stokes.set_farfield_shape(k)
            assert relerr(stokes.matvec(n, m, no, xx, yy, eta, ibary=1), vo) < 1e-12, k
    finally:
        stokes.set_farfield_shape(-1)
    with pytest.raises(stokes.StokesError):
        stokes.set_farfield_shape(stokes.farfield_shape_count())


def test_matvec_tight_geometry_many_near_pairs():
    """Bodies 0.002 apart and against the wall: every near branch with many entries."""
    thl = make_bodies(30, 128, 0.05, 0.12, 0.002, 31)
    m, xx, yy = getnxy(new_thlenden(thl))
    n, no = 128, 1024
    N = 2 * no + 2 * n * m + 3 * m
    eta = np.random.default_rng(8).standard_normal(N)
    geo = O.Geometry(n, m, no, xx, yy)
    vo = geo.matvec(eta, 1)
    assert relerr(vo, geo.matvec(eta, 0)) > 1e-6  # near corrections are active
    assert relerr(stokes.matvec(n, m, no, xx, yy, eta, ibary=1), vo) < 1e-12


@pytest.mark.parametrize("name", FIXTURES)
def test_solve_matches_reference(golden, name):
    g = golden(name)
    n, m, no, xx, yy = _case(g)
    den, it = stokes.stokessolver(n, m, no, 1, 1, 0, 2000, xx, yy)
    assert stokes.last_status() == 0
    it_ref = int(g["iter_tol8"])
    assert abs(it - it_ref) <= max(2, it_ref // 50)
    assert relerr(den, g["den_tol8"]) < 1e-7
    stokes.set_gmres_tol(1e-12)
    den12, it12 = stokes.stokessolver(n, m, no, 1, 1, 0, 2000, xx, yy)
    info = stokes.last_solve_info()
    assert relerr(den12, g["den_tol12"]) < 1e-10
    assert info["true_relres"] < 1e-10


@pytest.mark.parametrize("name", FIXTURES)
def test_shear_stress_matches_reference(golden, name):
    g = golden(name)
    n, m, no, xx, yy = _case(g)
    tau = stokes.computeshearstress(n, m, no, xx, yy, g["den_tol12"], 1)
    assert relerr(tau, g["tau_tol12"]) < 1e-10


def test_shear_stress_vs_oracle_20_bodies():
    n, m, no, xx, yy = workload_xy(WORKLOADS["bodies100_n256"], nbodies=20)
    N = 2 * no + 2 * n * m + 3 * m
    den = np.random.default_rng(9).standard_normal(N) * 0.1
    tau = stokes.computeshearstress(n, m, no, xx, yy, den, 1)
    assert relerr(tau, O.shear_stress(n, m, no, xx, yy, den)) < 1e-10


def test_shear_stress_tight_geometry_vs_oracle():
    """Bodies 0.002 apart, 128 points (two bodies' targets per 256-lane block)."""
    thl = make_bodies(30, 128, 0.05, 0.12, 0.002, 31)
    m, xx, yy = getnxy(new_thlenden(thl))
    n, no = 128, 1024
    N = 2 * no + 2 * n * m + 3 * m
    den = np.random.default_rng(10).standard_normal(N) * 0.1
    tau = stokes.computeshearstress(n, m, no, xx, yy, den, 1)
    assert relerr(tau, O.shear_stress(n, m, no, xx, yy, den)) < 1e-10


def test_trapezoid_solve_vs_reference(golden):
    g = golden("tight8_n64")
    n, m, no, xx, yy = _case(g)
    stokes.set_gmres_tol(1e-12)
    den, it = stokes.stokessolver(n, m, no, 0, 0, 0, 2000, xx, yy)
    assert relerr(den, g["den_trap_tol12"]) < 1e-10


def test_wall_only():
    """No bodies: the wall alone (rows of the stokeslet/constraint blocks vanish)."""
    n, m, no = 32, 0, 256
    xx = yy = np.zeros(0)
    stokes.set_gmres_tol(1e-12)
    den, it = stokes.stokessolver(n, m, no, 1, 1, 0, 500, xx, yy)
    ref, it_o = O.stokessolver(n, m, no, xx, yy, ibary=1, ibc=0, maxl=500, tol=1e-12)
    assert relerr(den, ref) < 1e-10
    assert stokes.computeshearstress(n, m, no, xx, yy, den, 1).size == 0


def test_initial_guess_is_used(golden):
    g = golden("tight8_n64")
    n, m, no, xx, yy = _case(g)
    den, it = stokes.stokessolver(n, m, no, 1, 1, 0, 2000, xx, yy, den=g["den_tol12"])
    assert it == 0  # the guess already meets 1e-8
    assert np.array_equal(den, g["den_tol12"])


def test_maxl_exhaustion(golden):
    """maxl reached: status 1, nits = maxl, the partial solution returned.

    The 7th iterate sits in GMRES's plateau (it moves by ~1e-2 per iteration
    and its least-squares problem amplifies ~1e-13 geometry rounding to
    ~1e-8), so it is compared at 1e-6; converged solves are compared at 1e-10
    elsewhere."""
    g = golden("tight8_n64")
    n, m, no, xx, yy = _case(g)
    den, it = stokes.stokessolver(n, m, no, 1, 1, 0, 7, xx, yy)
    assert it == 7 and stokes.last_status() == 1
    geo = O.Geometry(n, m, no, xx, yy)
    ref, it_o, _err, ierr = geo.gmres(geo.rhs(0), ibary=1, maxl=7, tol=1e-8)
    ref8, _, _, _ = geo.gmres(geo.rhs(0), ibary=1, maxl=8, tol=1e-8)
    assert it_o == 7 and ierr == 1
    assert relerr(den, ref) < 1e-6 < relerr(ref8, ref)
    implicit = stokes.last_solve_info()["implicit_relres"]
    assert abs(implicit - _err) < 1e-8


def test_solve_record_and_fp64_floor(golden):
    """erosion_last_solve_info's cycle record.  A solve that meets its
    tolerance says so and measures no floor.  A tolerance of 1e-17 cannot be
    met by any FP64 density: the solve restarts until a restart stops halving
    the true residual, is counted above the tolerance, and measures the FP64
    floor of its density (the residual of a half-ulp perturbation of it),
    which must be positive, far below 1e-12 for this O(1) density, and within
    a factor 1000 of the residual reached."""
    g = golden("tight8_n64")
    n, m, no, xx, yy = _case(g)
    stokes.set_preconditioner(stokes.PRECOND_CURVE_FOURIER)
    try:
        stokes.set_gmres_tol(1e-10)
        stokes.stokessolver(n, m, no, 1, 1, 0, 2000, xx, yy)
        ok = stokes.last_solve_info()
        stokes.set_gmres_tol(1e-17)
        stokes.stokessolver(n, m, no, 1, 1, 0, 2000, xx, yy)
        miss = stokes.last_solve_info()
    finally:
        stokes.set_gmres_tol(1e-8)
        stokes.set_preconditioner(stokes.PRECOND_NONE)
    rec = ok["cycle_record"]
    assert rec["stop"] == "met tolerance" and rec["fp64_floor_relres"] == 0.0
    assert rec["true_relres"][ok["cycles"] - 1] == ok["true_relres"] <= 1e-10
    assert sum(rec["iters"]) == ok["iters"]
    rec = miss["cycle_record"]
    print(f"tol 1e-17: {miss['cycles']} cycles, record {rec}")
    assert miss["tol_miss"] == 1 and miss["latest_miss_record"] == rec
    assert miss["tol_misses_total"] >= ok["tol_misses_total"] + 1
    assert rec["stop"] != "met tolerance"
    assert sum(rec["iters"]) == miss["iters"]
    fl = rec["fp64_floor_relres"]
    assert 0.0 < fl < 1e-12 and miss["true_relres"] < 1000 * fl


def test_sampled_kernel_timing(golden):
    """Kernel timing with a period of 4 records one GMRES iteration's launches
    in four, each standing for four.  The estimated totals match those of
    timing every launch: launch counts within the last incomplete period (3),
    work within what the Gram-Schmidt passes' growth over a period allows, and
    the measured per-launch means agree."""
    g = golden("bodies10_n128")
    n, m, no, xx, yy = _case(g)
    stokes.set_preconditioner(stokes.PRECOND_CURVE_FOURIER)
    kt = {}
    try:
        for period in (1, 4):
            stokes.kernel_timing(True, period=period)
            den, its = stokes.stokessolver(n, m, no, 1, 1, 0, 4000, xx, yy)
            kt[period] = stokes.kernel_timers()
            stokes.kernel_timing(False)
        with pytest.raises(stokes.StokesError):
            stokes.kernel_timing(True, period=0)
    finally:
        stokes.kernel_timing(False)
        stokes.set_preconditioner(stokes.PRECOND_NONE)
    fams = [f for f, v in kt[1].items() if v["launches"] > 0]
    print(f"{its} iterations; families {fams}")
    assert "gram_schmidt" in fams and "precond" in fams and len(fams) >= 3
    for f in fams:
        a, b = kt[1][f], kt[4][f]
        assert abs(b["launches"] - a["launches"]) <= 3, (f, a, b)
        assert abs(b["work"] - a["work"]) <= 0.1 * a["work"], (f, a, b)
        ma, mb = a["ms"] / a["launches"], b["ms"] / b["launches"]
        print(f"{f}: {a['launches']} launches, mean {ma * 1e3:.1f} us; sampled "
              f"{b['launches']}, mean {mb * 1e3:.1f} us")
        assert 0.6 < mb / ma < 1.6, (f, ma, mb)


def test_cluster_far_field_one_stream_identical():
    """The counter-pass hook runs the cluster far field's expansions on the
    library stream ahead of the direct sums: the same kernels on the same
    data, so the operator's output is identical bit for bit."""
    n, m, no, xx, yy = workload_xy(WORKLOADS["bodies100_n256"], nbodies=40)
    eta = np.random.default_rng(41).standard_normal(2 * no + 2 * n * m + 3 * m)
    stokes.set_farfield_shape(stokes.farfield_shape_count() - 1)
    try:
        v2 = stokes.matvec(n, m, no, xx, yy, eta)
        stokes.set_fmm_one_stream(True)
        v1 = stokes.matvec(n, m, no, xx, yy, eta)
        stokes.set_fmm_one_stream(False)
        v2b = stokes.matvec(n, m, no, xx, yy, eta)
    finally:
        stokes.set_fmm_one_stream(False)
        stokes.set_farfield_shape(-1)
    assert np.array_equal(v1, v2) and np.array_equal(v2b, v2)
    assert stokes.lib().erosion_set_fmm_one_stream(2) == -1


def test_unsupported_combinations_raise(golden):
    g = golden("tight8_n64")
    n, m, no, xx, yy = _case(g)
    with pytest.raises(stokes.StokesError):
        stokes.stokessolver(n, m, no, 0, 1, 0, 100, xx, yy)
    with pytest.raises(stokes.StokesError):
        stokes.computeshearstress(n, m, no, xx, yy, g["den_tol12"], 0)
    with pytest.raises(stokes.StokesError):
        stokes.matvec(0, m, no, xx, yy, g["eta"], 1)
    # the library stays usable after an error
    v = stokes.matvec(n, m, no, xx, yy, g["eta"], ibary=1)
    assert relerr(v, g["mv_fmmbary"]) < 1e-12


@pytest.mark.parametrize("name", FIXTURES)
@pytest.mark.parametrize("ibary,key", [(1, "qoi_bary"), (0, "qoi_trap")])
def test_qoi_targets_match_reference(golden, name, ibary, key):
    g = golden(name)
    n, m, no, xx, yy = _case(g)
    out = stokes.computeqoitargets(n, m, no, ibary, xx, yy, g["den_tol12"], g["qoi_xt"], g["qoi_yt"])
    for i in range(4):
        assert relerr(out[i], g[key][i]) < 1e-10, i


def test_qoi_few_targets_path_matches_many_targets_path(golden):
    """Up to 64 targets take one wave per (target, curve); more take the
    tiled kernel.  The same targets through both agree to 1e-11 of the
    largest value (the sums run in different orders; near a boundary their
    terms grow like 1/rho^3)."""
    g = golden("bodies10_n128")
    n, m, no, xx, yy = _case(g)
    rng = np.random.default_rng(12)
    xt = rng.uniform(-1.5, 1.5, 150)
    yt = rng.uniform(-0.9, 0.9, 150)
    many = stokes.computeqoitargets(n, m, no, 1, xx, yy, g["den_tol12"], xt, yt)
    few = stokes.computeqoitargets(n, m, no, 1, xx, yy, g["den_tol12"], xt[:40], yt[:40])
    for i in range(4):
        assert np.max(np.abs(few[i] - many[i][:40])) <= 1e-11 * max(1.0, np.max(np.abs(many[i]))), i


def test_pipe_flow_pressure_drop():
    """No bodies, parabolic inflow u = 1 - y^2: the pressure drop between
    x = -2 and x = 2 is close to the Poiseuille value 8 (DensityStress.jl:158)."""
    from erosion_amd.density_stress import getpdrop
    from erosion_amd.theta_len import ParamSet, new_thlenden
    params = ParamSet(npts=64, nouter=1024, ibc=0, fixpdrop=True)
    thld = new_thlenden([])
    stokes.set_gmres_tol(1e-12)
    from erosion_amd.density_stress import compute_density_
    compute_density_(thld, params)
    pdrop, qavg = getpdrop(thld, params)
    assert abs(pdrop - 8.0) < 0.4
    ys = -1 + (np.arange(13) + 0.5) * 2 / 13
    assert abs(qavg - np.mean(1 - ys ** 2)) < 0.05  # Poiseuille profile sampled at the slice


@pytest.mark.parametrize("name", FIXTURES)
def test_pressure_matches_reference(golden, name):
    """computepressure_ against the reference's computePressure (PostProcess.jl:107-110)."""
    g = golden(name)
    n, m, no, xx, yy = _case(g)
    p = stokes.computepressure(n, m, no, xx, yy, g["den_tol12"])
    assert relerr(p, g["pressure_tol12"]) < 1e-10
    assert relerr(p, O.pressure(n, m, no, xx, yy, g["den_tol12"])) < 1e-10


def test_pressure_full_size_and_drag():
    """100 bodies x 256: GPU pressure against the oracle, and the drag
    bookkeeping of PostProcess.jl:130-160 on top of the GPU stress and pressure."""
    from erosion_amd.post_process import drag
    from erosion_amd.theta_len import ParamSet
    from erosion_amd.geometry import workload_geometry
    w = WORKLOADS["bodies100_n256"]
    n, m, no, xx, yy = workload_xy(w)
    N = 2 * no + 2 * n * m + 3 * m
    den = np.random.default_rng(5).standard_normal(N)
    p = stokes.computepressure(n, m, no, xx, yy, den)
    po = O.pressure(n, m, no, xx, yy, den)
    assert relerr(p, po) < 1e-10
    from erosion_amd.density_stress import compute_density_
    th = workload_geometry(w, nbodies=10)
    params = ParamSet(npts=n, nouter=no, ibary=1, ifmm=1, ibc=0, dt=1e-3, maxl=4000,
                      fixpdrop=False)
    # as in PostProcess.jl, drag reads a density computed beforehand
    with pytest.raises(stokes.StokesError):
        drag(th, params)
    compute_density_(th, params)
    d = drag(th, params)
    assert d.press_all.shape == (n, 10) and np.all(np.isfinite(d.press_all))
    print(f"drag on 10 bodies: pressure ({d.pdragx:.4f}, {d.pdragy:.4f}), "
          f"viscous ({d.vdragx:.4f}, {d.vdragy:.4f}), umax {d.umax:.4f}")
    assert np.isfinite([d.pdragx, d.pdragy, d.vdragx, d.vdragy]).all() and d.pdragx != 0.0
