"""Per-curve two-level right preconditioner.

CPU: the numpy blocks equal columns of the oracle's operator, and a numpy
right-preconditioned GMRES converges to the reference density in fewer
iterations.  GPU (marked): the library's P^{-1} equals the numpy one, and
preconditioned solves return the reference density with a true residual
below the tolerance in fewer iterations.
"""
import numpy as np
import pytest

from conftest import load_golden, relerr
from oracle import oracle as O
from precond_ref import TwoLevel, body_block, body_indices, fourier_basis, wall_block


def _geo(name):
    g = load_golden(name)
    n, m, no = int(g["n"]), int(g["m"]), int(g["no"])
    return g, n, m, no, O.Geometry(n, m, no, g["xx"], g["yy"])


def gmres_right(A, Minv, b, tol, maxl):
    """Textbook GMRES on A Minv with two-pass classical Gram-Schmidt."""
    N = len(b)
    V = np.zeros((maxl + 1, N))
    H = np.zeros((maxl + 1, maxl))
    beta = np.linalg.norm(b)
    V[0] = b / beta
    for j in range(maxl):
        w = A(Minv(V[j]))
        for _ in range(2):
            h = V[:j + 1] @ w
            w -= V[:j + 1].T @ h
            H[:j + 1, j] += h
        H[j + 1, j] = np.linalg.norm(w)
        V[j + 1] = w / H[j + 1, j]
        e = np.zeros(j + 2)
        e[0] = beta
        y = np.linalg.lstsq(H[:j + 2, :j + 1], e, rcond=None)[0]
        if np.linalg.norm(H[:j + 2, :j + 1] @ y - e) / beta <= tol:
            break
    return Minv(V[:j + 1].T @ y), j + 1


def test_fourier_basis_orthonormal():
    for N, K in [(64, 8), (256, 16), (33, 16)]:
        F = fourier_basis(N, K)
        assert np.abs(F.T @ F - np.eye(2 * K + 1)).max() < 1e-13


@pytest.mark.parametrize("name", ["tight8_n64", "bodies10_n128"])
def test_body_block_is_operator_block(name):
    """Columns of the closed-form body block = oracle A e_j restricted to the body."""
    g, n, m, no, geo = _geo(name)
    b = m // 2
    idx = body_indices(n, m, no, b)
    B = body_block(geo, g["xx"], g["yy"], b)
    rng = np.random.default_rng(3)
    cols = np.concatenate([rng.choice(2 * n, 12, replace=False), 2 * n + np.arange(3)])
    for c in cols:
        e = np.zeros(geo.nunk)
        e[idx[c]] = 1.0
        col = geo.matvec(e, 1)[idx]
        # near-pair corrections couple different curves only
        assert np.abs(col - B[:, c]).max() < 1e-14 * max(1.0, np.abs(B).max())


def test_wall_block_is_operator_block():
    g, n, m, no, geo = _geo("tight8_n64")
    W = wall_block(geo)
    for c in [0, 5, no - 1, no, no + 77, 2 * no - 1]:
        e = np.zeros(geo.nunk)
        e[c] = 1.0
        col = geo.matvec(e, 1)[:2 * no]
        assert np.abs(col - W[:, c]).max() < 1e-14 * max(1.0, np.abs(W).max())


def test_two_level_exact_on_coarse_space():
    """For r = B Q c on a single curve, P^{-1} r restricted to Q equals c
    when B maps the coarse space to itself; in general the coarse part of
    P^{-1} B Q c is c up to the coupling of high modes, so check the defining
    identity Q^T B P^{-1}-coarse directly: G^{-1} (Q^T B Q) = I."""
    g, n, m, no, geo = _geo("tight8_n64")
    P = TwoLevel(geo, g["xx"], g["yy"], 8, 16)
    for idx, Q, Gi in P.parts[:3]:
        blk = wall_block(geo) if len(idx) == 2 * no else None
        if blk is None:
            b = (idx[0] - 2 * no) // (2 * n)
            blk = body_block(geo, g["xx"], g["yy"], b)
        assert np.abs(Gi @ (Q.T @ blk @ Q) - np.eye(Q.shape[1])).max() < 1e-11


def test_preconditioned_gmres_cpu_fewer_iterations():
    """tight8 at 1e-10: same density as the reference solve, fewer iterations."""
    g, n, m, no, geo = _geo("tight8_n64")
    b = geo.rhs(0)
    A = lambda v: geo.matvec(v, 1)  # noqa: E731
    P = TwoLevel(geo, g["xx"], g["yy"], 8, 16)
    x, it = gmres_right(A, P, b, 1e-10, 400)
    assert np.linalg.norm(b - A(x)) / np.linalg.norm(b) < 2e-10
    assert relerr(x, g["den_tol12"]) < 1e-8
    x0, it0 = gmres_right(A, lambda r: r, b, 1e-10, 400)
    assert it < 0.8 * it0


# ------------------------------------------------------------------ GPU
gpu = pytest.mark.gpu


@pytest.fixture
def pc_reset():
    from erosion_amd import stokes
    stokes.set_gmres_tol(1e-8)
    yield stokes
    stokes.set_preconditioner(0)
    stokes.set_gmres_tol(1e-8)


@gpu
@pytest.mark.parametrize("name", ["tight8_n64", "bodies10_n128"])
@pytest.mark.parametrize("modes", [(8, 16), (3, 5), (16, 16), (20, 32)])
def test_gpu_apply_matches_numpy(pc_reset, name, modes):
    stokes = pc_reset
    g, n, m, no, geo = _geo(name)
    stokes.set_preconditioner(1, *modes)
    assert stokes.get_preconditioner() == (1, modes[0], modes[1])
    r = np.random.default_rng(11).standard_normal(geo.nunk)
    z = stokes.precond_apply(n, m, no, g["xx"], g["yy"], r)
    zr = TwoLevel(geo, g["xx"], g["yy"], *modes)(r)
    assert relerr(z, zr) < 1e-11


@gpu
@pytest.mark.parametrize("name", ["tight8_n64", "bodies10_n128"])
def test_gpu_preconditioned_solve(pc_reset, name):
    stokes = pc_reset
    g, n, m, no, geo = _geo(name)
    stokes.set_gmres_tol(1e-12)
    stokes.set_preconditioner(0)
    _, it_plain = stokes.stokessolver(n, m, no, 1, 1, 0, 8000, g["xx"], g["yy"])
    stokes.set_preconditioner(1, 8, 16)
    den, it = stokes.stokessolver(n, m, no, 1, 1, 0, 8000, g["xx"], g["yy"])
    info = stokes.last_solve_info()
    assert stokes.last_status() == 0 and info["preconditioner"] == 1
    assert info["true_relres"] < 1e-10
    assert relerr(den, g["den_tol12"]) < 1e-9
    assert it < 0.6 * it_plain, (it, it_plain)


@gpu
def test_gpu_preconditioned_solve_full_size(pc_reset):
    """100 bodies x 256 points, wall 2048 at 1e-10: true residual below the
    tolerance, density equal to the plain solve's within 1e-9 (both at
    1e-10), the oracle's operator on the library geometry giving a residual
    within 2e-10 -- which checks the operator applications that reuse the
    preconditioner's constraint rows and rank-one parts inside GMRES -- in
    far fewer iterations."""
    import numpy as np
    from erosion_amd.geometry import WORKLOADS, workload_xy
    from test_configs import oracle_on_library_geometry
    stokes = pc_reset
    n, m, no, xx, yy = workload_xy(WORKLOADS["bodies100_n256"])
    stokes.set_gmres_tol(1e-10)
    d0, it0 = stokes.stokessolver(n, m, no, 1, 1, 0, 8000, xx, yy)
    stokes.set_preconditioner(1, 8, 16)
    d1, it1 = stokes.stokessolver(n, m, no, 1, 1, 0, 8000, xx, yy)
    info = stokes.last_solve_info()
    g = oracle_on_library_geometry(n, m, no, xx, yy)
    b = g.rhs(0)
    r1 = float(np.linalg.norm(b - g.matvec(d1)) / np.linalg.norm(b))
    print(f"plain {it0} iterations, preconditioned {it1}, true relres {info['true_relres']:.2e}, "
          f"oracle residual {r1:.2e}, density vs plain {relerr(d1, d0):.2e}")
    assert info["true_relres"] < 1e-10
    assert r1 <= 2e-10
    assert relerr(d1, d0) < 1e-9
    assert it1 < 0.5 * it0


@gpu
def test_gpu_preconditioner_default_modes(pc_reset):
    """32 body modes and 16 wall modes unless the caller asks for others."""
    stokes = pc_reset
    stokes.set_preconditioner(stokes.PRECOND_CURVE_FOURIER)
    assert stokes.get_preconditioner() == (1, 32, 16)


@gpu
def test_gpu_preconditioner_mode_limits(pc_reset):
    from erosion_amd.geometry import make_bodies
    from erosion_amd.theta_len import getnxy, new_thlenden
    stokes = pc_reset
    with pytest.raises(stokes.StokesError):
        stokes.set_preconditioner(1, 0, 16)
    with pytest.raises(stokes.StokesError):
        stokes.set_preconditioner(1, 8, 33)
    with pytest.raises(stokes.StokesError):
        stokes.set_preconditioner(2)
    # 32 points per body hold modes 0..15: 16 body modes are lowered to 15
    thl = make_bodies(2, 32, 0.08, 0.12, 0.05, 5)
    _, xx, yy = getnxy(new_thlenden(thl))
    stokes.set_preconditioner(1, 16, 16)
    stokes.set_gmres_tol(1e-10)
    den16, it16 = stokes.stokessolver(32, 2, 128, 1, 1, 0, 500, xx, yy)
    assert stokes.last_status() == 0 and stokes.get_preconditioner() == (1, 16, 16)
    stokes.set_preconditioner(1, 15, 16)
    den15, it15 = stokes.stokessolver(32, 2, 128, 1, 1, 0, 500, xx, yy)
    assert it15 == it16 and np.array_equal(den15, den16)
    stokes.set_gmres_tol(1e-8)
