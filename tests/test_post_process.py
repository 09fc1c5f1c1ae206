"""Post-processing on the solver's results (GPU): drag, resistivity and the
rotated-flow solve of ErosionPackage/src/PostProcess.jl:62-71, 124-160 and
Main.jl:69-75 / DensityStress.jl:40.

Bars:
  drag of each body + its stokeslet strength   <= 1e-7 max|lambda| (force balance,
                                                  quadrature accuracy at 128 points)
  resistivity vs the oracle's target values     <= 1e-10 relative
  rotated-flow density vs the oracle's solve    <= 1e-10 relative (tol 1e-12)
"""
import numpy as np
import pytest

from conftest import relerr
from erosion_amd import stokes
from erosion_amd.density_stress import compute_density_, xyrot
from erosion_amd.post_process import drag, getns, resistivity
from erosion_amd.theta_len import ParamSet, getnxy, new_thlenden
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _state(nb=10, n=128, seed=1001):
    from erosion_amd.geometry import make_bodies
    return new_thlenden(make_bodies(nb, n, 0.05, 0.1, 0.05, seed))


@pytest.fixture(autouse=True)
def _tight():
    stokes.set_gmres_tol(1e-12)
    yield
    stokes.set_gmres_tol(1e-8)


def test_drag_balances_the_stokeslets():
    """Force balance: the 2D stokeslet (1/4 pi)(-log rho I + r r / rho^2)
    lambda of the completed representation is the flow of a point force
    lambda on the fluid (unit viscosity), and the double layer and the rotlet
    carry no net force, so the fluid's force on body b -- its pressure drag
    plus its viscous drag as drag() integrates them (PostProcess.jl:130-160)
    -- is -lambda_b.  A lost pressure or viscous part, or a wrong sign in
    either, breaks the balance."""
    th = _state()
    params = ParamSet(npts=128, nouter=512, ibc=0, fixpdrop=False, maxl=4000)
    compute_density_(th, params)
    d = drag(th, params)
    n, m, no = 128, len(th.thlenvec), 512
    lam = np.asarray(th.density)[2 * no + 2 * n * m:].reshape(m, 3)[:, :2]
    force = np.zeros((m, 2))
    for b, t in enumerate(th.thlenvec):
        sx, sy, nx, ny = getns(t.theta)
        ds = t.len / n
        p, tau = d.press_all[:, b], d.tau_all[:, b]
        force[b] = [(p @ nx + tau @ sx) * ds, (p @ ny + tau @ sy) * ds]
    err = np.abs(force + lam).max() / np.abs(lam).max()
    print(f"drag + stokeslet, max over bodies: {err:.2e} of max |lambda| = {np.abs(lam).max():.3e}")
    assert err <= 1e-7
    assert abs(d.pdragx + d.vdragx + lam[:, 0].sum()) <= 1e-7 * np.abs(lam).sum()


def test_resistivity_matches_the_oracle_targets():
    """resistivity (PostProcess.jl:62-71) from the library's target
    evaluations against the same formula on the oracle's velocity and
    pressure at the same 26 targets, on the library's density."""
    th = _state()
    params = ParamSet(npts=128, nouter=512, ibc=0, fixpdrop=False, maxl=4000)
    compute_density_(th, params)
    r_lib = resistivity(th, params)
    n, no = 128, 512
    m, xx, yy = getnxy(th)
    ny_ = 13
    dy = 2 / ny_
    ylocs = -1 + 0.5 * dy + dy * np.arange(ny_)
    out = {}
    for x0 in (2.0, -2.0):
        u, v, p, vort = O.qoi_targets(n, m, no, 1, xx, yy, np.asarray(th.density),
                                      np.full(ny_, x0), ylocs)
        out[x0] = (np.mean(u), np.mean(p))
    pdrop = out[-2.0][1] - out[2.0][1]
    qavg = 0.5 * (out[2.0][0] + out[-2.0][0])
    r_orc = 0.5 * pdrop / qavg - 3 * 2.0
    print(f"resistivity: library {r_lib:.12f}, oracle {r_orc:.12f}")
    assert abs(r_lib - r_orc) <= 1e-10 * max(1.0, abs(r_orc))


def test_rotated_flow_density_matches_the_oracle():
    """The rotated-flow density (denrot: bodies rotated by 90 degrees inside
    the fixed wall, DensityStress.jl:40 and :58-66) against the oracle's
    GMRES on the same rotated boundary, both at tol 1e-12."""
    th = _state(8, 64, 77)
    params = ParamSet(npts=64, nouter=256, ibc=0, fixpdrop=False, maxl=4000)
    compute_density_(th, params, rotation=True)
    m, xx, yy = getnxy(th)
    xr, yr = xyrot(xx, yy)
    ref, _ = O.stokessolver(64, m, 256, xr, yr, ibary=1, ibc=0, maxl=4000, tol=1e-12)
    e = relerr(th.denrot, ref)
    print(f"rotated-flow density vs oracle: {e:.2e}")
    assert len(th.density) == 0  # only the rotated density was solved
    assert e <= 1e-10
