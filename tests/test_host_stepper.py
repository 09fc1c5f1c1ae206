"""Host side of the erosion step (theta-L state, spectral routines, RK2).

Mirrors ErosionPackage/src/SpectralMethods.jl, ThetaLen.jl and
TimeStepping.jl.  Pinned by closed forms (trigonometric polynomials, circles)
and by the per-body formulas the batched code must reproduce exactly.
"""
import math

import numpy as np
import pytest

from erosion_amd import time_stepping as TS
from erosion_amd.geometry import make_bodies
from erosion_amd.spectral import expsmooth, gaussfilter, specdiff, specint
from erosion_amd.theta_len import (ParamSet, circ2thlen, getalpha, getnxy, getxy,
                                   new_thlenden)


def test_specdiff_trig_polynomial():
    n = 64
    x = 2 * np.pi * np.arange(n) / n
    f = np.sin(3 * x) + 0.5 * np.cos(7 * x)
    assert np.abs(specdiff(f) - (3 * np.cos(3 * x) - 3.5 * np.sin(7 * x))).max() < 1e-12
    # interval length scales the derivative
    assert np.abs(specdiff(f, 1.0) - 2 * np.pi * specdiff(f)).max() < 1e-11


def test_specint_inverts_specdiff_on_mean_free():
    n = 128
    x = 2 * np.pi * np.arange(n) / n
    f = np.cos(2 * x) - 0.3 * np.sin(5 * x)
    F = specint(f)
    assert abs(F.mean()) < 1e-13
    assert np.abs(specdiff(F) - f).max() < 1e-12
    assert np.abs(F - (np.sin(2 * x) / 2 + 0.3 * np.cos(5 * x) / 5)).max() < 1e-13


def test_expsmooth_mode_factors():
    n = 32
    x = 2 * np.pi * np.arange(n) / n
    f = np.cos(4 * x)
    assert np.abs(expsmooth(f, 0.01) - np.exp(-0.16) * f).max() < 1e-13
    assert np.abs(gaussfilter(f, 0.2) - np.exp(-0.5 * 0.04 * 16) * f).max() < 1e-13


def test_batched_spectral_equals_rows():
    rng = np.random.default_rng(2)
    A = rng.standard_normal((5, 48))
    for fn in (specdiff, specint):
        B = fn(A)
        for r in range(5):
            assert np.array_equal(B[r], fn(A[r]))
    fac = rng.random(5) * 0.01
    B = expsmooth(A, fac)
    for r in range(5):
        assert np.array_equal(B[r], expsmooth(A[r], fac[r]))


def test_circle_state():
    t = circ2thlen(128, 0.2, 0.3, -0.1)
    x, y = getxy(t)
    assert np.abs(np.hypot(x - 0.3, y + 0.1) - 0.2).max() < 1e-13
    area = TS.body_areas(new_thlenden([t]))[0]
    assert abs(area - np.pi * 0.04) < 1e-13


def test_getnxy_batched_equals_per_body():
    thl = make_bodies(7, 64, 0.05, 0.1, 0.05, 9)
    nb, x, y = getnxy(new_thlenden(thl))
    xs, ys = zip(*(getxy(t) for t in thl))
    assert nb == 7
    assert np.array_equal(x, np.concatenate(xs)) and np.array_equal(y, np.concatenate(ys))


def test_getderivs_circle_closed_form():
    """A circle under the stress tau0 + tau1 cos(alpha) (alpha the position
    angle): erosion at that inward normal speed keeps it a circle to first
    order, shrinks its length at dL/dt = -2 pi tau0 and moves its centre by
    -tau1 along x (the mode-1 part of the normal speed shifts the circle), so
    getderivs / tangvel (TimeStepping.jl:42-84) must give mterm = -2 pi tau0,
    nterm = 0, xsmdot = -tau1, ysmdot = 0; with fixarea the mean stress is
    removed and the length stays."""
    n, rad = 128, 0.3
    t = circ2thlen(n, rad, 0.1, -0.2)
    t.matau = 0.7
    alpha = getalpha(n)
    for tau0, tau1 in ((0.7, 0.0), (0.7, 0.25), (1.3, -0.4)):
        atau = tau0 + tau1 * np.cos(alpha)
        for fixarea in (0, 1):
            params = ParamSet(npts=n, nouter=256, fixarea=fixarea, epsfac=15)
            d = TS.getderivs_body(t, params, atau)
            mean_tau = tau0 - fixarea * t.matau
            assert abs(d.mterm - (-2 * math.pi * mean_tau)) < 1e-13
            assert np.abs(d.nterm).max() < 1e-12
            assert abs(d.xsmdot - (-tau1)) < 1e-13 and abs(d.ysmdot) < 1e-13


def test_getderivs_batched_equals_per_body():
    thl = make_bodies(6, 64, 0.05, 0.1, 0.05, 4)
    rng = np.random.default_rng(5)
    for t in thl:
        t.matau = 1.0 + rng.random()
    params = ParamSet(npts=64, nouter=256)
    atau = rng.random((64, 6))
    d1 = [TS.getderivs_body(t, params, atau[:, b]) for b, t in enumerate(thl)]
    d2 = TS.getderivs_bodies(thl, params, atau)
    for a, b in zip(d1, d2):
        assert a.mterm == b.mterm and a.xsmdot == b.xsmdot and a.ysmdot == b.ysmdot
        assert np.array_equal(a.nterm, b.nterm)


def test_rk2_step_matches_per_body_formulas(monkeypatch):
    """One RK2 step with a fixed synthetic stress: the batched update equals
    the per-body loop of TimeStepping.jl:150-182 evaluated here directly."""
    thl = make_bodies(4, 64, 0.05, 0.1, 0.05, 12)
    params = ParamSet(npts=64, nouter=256, dt=1e-3)
    stress = {}

    def fake_getstress(thlenden, params, den0=None, timings=None):
        tv = thlenden.thlenvec
        at = np.stack([0.5 + 0.1 * np.cos(t.theta) for t in tv]).T
        for b, t in enumerate(tv):
            t.matau = float(np.mean(np.abs(at[:, b])))
        stress[id(thlenden)] = at
        return at

    monkeypatch.setattr(TS, "getstress_", fake_getstress)
    th0 = new_thlenden(thl)
    th1 = TS.rungekutta2(th0, params)

    # per-body restatement of the two half steps
    def half(t0s, tds, dt1, dt2):
        at = np.stack([0.5 + 0.1 * np.cos(t.theta) for t in tds]).T
        for b, t in enumerate(tds):
            t.matau = float(np.mean(np.abs(at[:, b])))
        out = []
        alpha = getalpha(64)
        for b, (t0, td) in enumerate(zip(t0s, tds)):
            d = TS.getderivs_body(td, params, at[:, b])
            z0 = 2 * math.pi / t0.len * t0.matau
            zd = 2 * math.pi / td.len * td.matau
            th = expsmooth(t0.theta - alpha, params.epsilon * dt1 * zd) + alpha
            th = th + dt1 * expsmooth(d.nterm, params.epsilon * dt2 * (2 * zd - z0))
            out.append((th, t0.len + dt1 * d.mterm, t0.xsm + dt1 * d.xsmdot,
                        t0.ysm + dt1 * d.ysmdot))
        return out

    import copy
    ref0 = copy.deepcopy(make_bodies(4, 64, 0.05, 0.1, 0.05, 12))
    h = half(ref0, ref0, 0.5e-3, 0.5e-3)
    mid = [TS.ThetaLenType(a, b, c, d, float("nan")) for a, b, c, d in h]
    full = half(ref0, mid, 1e-3, 0.5e-3)
    for t, (th, ln, xs, ys) in zip(th1.thlenvec, full):
        assert np.abs(t.theta - th).max() < 1e-14
        assert abs(t.len - ln) < 1e-14 and abs(t.xsm - xs) < 1e-15 and abs(t.ysm - ys) < 1e-15
    assert th1.tt == pytest.approx(1e-3)


def test_previous_density_extrapolation():
    """order 2 returns 2 x_k - x_(k-1) of the last two densities of states
    with the same body count, order 1 the last one; a body-count change
    empties the history."""
    from erosion_amd.time_stepping import PreviousDensity
    thl = make_bodies(3, 16, 0.05, 0.1, 0.05, 1)
    n, m, no = 16, 3, 8
    N = 2 * no + 2 * n * m + 3 * m

    def state(v, bodies=thl):
        s = new_thlenden(list(bodies))
        s.density = np.full(2 * no + 2 * n * len(bodies) + 3 * len(bodies), float(v))
        return s

    g1, g2 = PreviousDensity(1), PreviousDensity(2)
    assert g2(state(0)) is None
    for g in (g1, g2):
        g.seen(state(1.0))
    assert np.all(g2(state(0)) == 1.0)          # one density: no extrapolation yet
    for g in (g1, g2):
        g.seen(state(1.5))
    assert np.all(g1(state(0)) == 1.5)
    assert np.all(g2(state(0)) == 2.0) and len(g2(state(0))) == N
    fewer = thl[:2]
    assert g2(state(0, fewer)) is None          # body count changed
    g2.seen(state(7.0, fewer))
    assert np.all(g2(state(0, fewer)) == 7.0)   # history restarted


def test_previous_density_drops_a_density_of_deleted_bodies():
    """timestep! deletes bodies from the state it just solved: with nouter
    known, that state's density (old body count) empties the history."""
    from erosion_amd.time_stepping import PreviousDensity
    thl = make_bodies(3, 16, 0.05, 0.1, 0.05, 1)
    n, no = 16, 8

    def state(v, bodies, m_solved):
        s = new_thlenden(list(bodies))
        s.density = np.full(2 * no + (2 * n + 3) * m_solved, float(v))
        return s

    g = PreviousDensity(2, nouter=no)
    g.seen(state(1.0, thl, 3))
    g.seen(state(2.0, thl, 3))
    assert np.all(g(state(0, thl, 3)) == 3.0)
    g.seen(state(4.0, thl[:2], 3))         # solved with 3 bodies, one deleted since
    assert g(state(0, thl[:2], 2)) is None
    g.seen(state(5.0, thl[:2], 2))
    assert np.all(g(state(0, thl[:2], 2)) == 5.0)
    assert len(g(state(0, thl[:2], 2))) == 2 * no + (2 * n + 3) * 2
