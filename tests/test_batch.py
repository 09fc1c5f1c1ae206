"""Many independent geometries in one call (erosion_solve_batch, GPU)."""
import numpy as np
import pytest

from conftest import relerr

pytestmark = pytest.mark.gpu

CASES = [(4, 64, 256, 1), (7, 64, 256, 2), (5, 128, 512, 3), (10, 64, 512, 4), (3, 128, 256, 5),
         (8, 64, 256, 6)]


def _geoms():
    from erosion_amd.geometry import make_bodies
    from erosion_amd.theta_len import getnxy, new_thlenden
    out = []
    for m, n, no, sd in CASES:
        _, xx, yy = getnxy(new_thlenden(make_bodies(m, n, 0.06, 0.12, 0.05, 300 + sd)))
        out.append((n, m, no, xx, yy))
    return out


@pytest.fixture
def st():
    from erosion_amd import stokes
    stokes.set_gmres_tol(1e-10)
    yield stokes
    stokes.set_preconditioner(0)
    stokes.set_precision(0)
    stokes.set_gmres_tol(1e-8)


@pytest.mark.parametrize("accel", [False, True])
def test_batch_equals_one_by_one(st, accel):
    if accel:
        st.set_preconditioner(1, 8, 16)
        st.set_precision(1)
    geoms = _geoms()
    res = st.solve_batch(geoms, workers=3, stress=True)
    for (n, m, no, xx, yy), r in zip(geoms, res):
        den, it = st.stokessolver(n, m, no, 1, 1, 0, 8000, xx, yy)
        tau = st.computeshearstress(n, m, no, xx, yy, den)
        assert r["status"] == 0
        assert abs(r["nits"] - it) <= 2
        assert relerr(r["den"], den) < (1e-8 if accel else 1e-11)
        assert relerr(r["tau"], tau) < (1e-8 if accel else 1e-10)


def test_batch_isolates_a_failing_geometry(st):
    """A geometry with too few points per body for the preconditioner (3: no
    Fourier mode besides the mean) fails alone; the others are solved."""
    from erosion_amd.geometry import make_bodies
    from erosion_amd.theta_len import getnxy, new_thlenden
    st.set_preconditioner(1, 8, 16)
    geoms = _geoms()[:3]
    _, xx, yy = getnxy(new_thlenden(make_bodies(2, 3, 0.1, 0.12, 0.1, 9)))
    geoms.insert(1, (3, 2, 256, xx, yy))
    res = st.solve_batch(geoms, workers=2)
    assert [r["status"] for r in res] == [0, -1, 0, 0]
    assert "geometry 1" in st.last_error() and "points per body" in st.last_error()
