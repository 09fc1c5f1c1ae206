"""Pin the C restatement (oracle/) before trusting it as the checker.

Each test names the reference routine it pins and how: against golden vectors
written by the flang build of the reference (tests/golden/make_golden.py), or
against a closed form / identity that a plausible mistake (dropped term, wrong
sign, transposed operand) would break.
"""
import numpy as np
import pytest

from conftest import relerr
from oracle import oracle as O

FIXTURES = ["bodies10_n128", "tight8_n64"]


def _geom(g):
    return O.Geometry(int(g["n"]), int(g["m"]), int(g["no"]), g["xx"], g["yy"])


# --------------------------------------------------------------- geometry
@pytest.mark.parametrize("name", FIXTURES)
def test_geometry_matches_reference(golden, name):
    """inner_geometry / outer_geometry / compute_derivs (stokesSolver.f:136-204, 417-457)."""
    g = golden(name)
    geo = _geom(g)
    for k in O.Geometry.NAMES:
        tol = 1e-10 if k.startswith("cur") else 1e-12
        assert relerr(geo.array(k), g["geo_" + k]) < tol, k


def test_circle_closed_form():
    """A circle: speed = r, inward normal -(cos, sin), curvature -1/r."""
    n, r, cx, cy = 64, 0.3, 0.1, -0.2
    th = 2 * np.pi * np.arange(n) / n
    geo = O.Geometry(n, 1, 128, cx + r * np.cos(th), cy + r * np.sin(th))
    assert np.allclose(geo.array("speed"), r, rtol=1e-13)
    assert np.allclose(geo.array("px"), -np.cos(th), atol=1e-13)
    assert np.allclose(geo.array("py"), -np.sin(th), atol=1e-13)
    assert np.allclose(geo.array("cur"), -1.0 / r, rtol=1e-11)
    assert np.allclose([geo.array("cx")[0], geo.array("cy")[0]], [cx, cy], atol=1e-15)


def test_fourier_diff_trig_polynomial():
    """fourierDiff (stokesSolver.f:348-388): exact on band-limited data, Nyquist dropped."""
    n = 48
    t = 2 * np.pi * np.arange(n) / n
    f = np.exp(1j * 3 * t) + 0.5 * np.cos(7 * t) - 2j * np.sin(11 * t) + 0.25 * np.cos(24 * t)
    df = 3j * np.exp(1j * 3 * t) - 3.5 * np.sin(7 * t) - 22j * np.cos(11 * t)
    assert np.max(np.abs(O.fourier_diff(f) - df)) < 1e-12
    # odd length goes through the plain DFT branch
    n = 45
    t = 2 * np.pi * np.arange(n) / n
    assert np.max(np.abs(O.fourier_diff(np.sin(5 * t)) - 5 * np.cos(5 * t))) < 1e-12


@pytest.mark.parametrize("name", FIXTURES)
def test_rhs(golden, name):
    """bd_condition (stokesSolver.f:208-314): pipe flow 1-y^2 on the wall, zero elsewhere."""
    g = golden(name)
    geo = _geom(g)
    b = geo.rhs(0)
    no = int(g["no"])
    yo = geo.array("yo")
    assert np.array_equal(b, g["rhs"])
    assert np.allclose(b[:no], (1 + yo) * (1 - yo), rtol=0, atol=1e-15)
    assert np.all(b[no:] == 0)
    b1 = geo.rhs(1)
    assert np.all(b1[:no] == 1) and np.all(b1[no:] == 0)


# ----------------------------------------------------------------- matvec
@pytest.mark.parametrize("name", FIXTURES)
def test_matvec_trapezoid_matches_reference(golden, name):
    """matvec_DLP (stokesSolver.f:1519-1912), direct sums on both sides."""
    g = golden(name)
    assert relerr(_geom(g).matvec(g["eta"], ibary=0), g["mv_trap"]) < 1e-13


@pytest.mark.parametrize("name", FIXTURES)
def test_matvec_fmmbary_matches_reference(golden, name):
    """matvec_DLP_fmmbary (stokesSolver.f:570-1124); the reference's far field is an
    FMM with tolerance 0.5e-12."""
    g = golden(name)
    assert relerr(_geom(g).matvec(g["eta"], ibary=1), g["mv_fmmbary"]) < 1e-12


def test_near_branch_is_active(golden):
    """The tight fixture exercises wall->body, body->wall and body->body replacements."""
    g = golden("tight8_n64")
    d = g["mv_fmmbary"] - g["mv_trap"]
    no, n, m = int(g["no"]), int(g["n"]), int(g["m"])
    assert np.max(np.abs(d[:2 * no])) > 1e-6  # wall rows changed (body -> wall)
    assert np.max(np.abs(d[2 * no:2 * no + 2 * n * m])) > 1e-6  # body rows changed


def test_far_field_equals_fmm_dipole_form(golden):
    """The trapezoid double layer equals the FMM's dipole sum (bhrouts.f:41-141 with
    dip1 = mu z/4pi, dip2 = (mu conj z - conj mu z)/4pi, stokesSolver.f:636-681)."""
    g = golden("bodies10_n128")
    n, m, no = int(g["n"]), int(g["m"]), int(g["no"])
    geo = _geom(g)
    eta = g["eta"].copy()
    eta[2 * no + 2 * n * m:] = 0.0  # no stokeslets/rotlets
    v = geo.matvec(eta, ibary=0)
    # strip local terms: diagonal curvature, jump, rank one
    xb = np.concatenate([geo.array("xo"), g["xx"]])
    yb = np.concatenate([geo.array("yo"), g["yy"]])
    pxb = np.concatenate([geo.array("px0"), geo.array("px")])
    pyb = np.concatenate([geo.array("py0"), geo.array("py")])
    spb = np.concatenate([geo.array("speed0"), geo.array("speed")])
    nsrc = np.concatenate([np.full(no, no), np.full(n * m, n)])
    ex = np.concatenate([eta[:no]] + [eta[2 * no + 2 * b * n:2 * no + 2 * b * n + n] for b in range(m)])
    ey = np.concatenate([eta[no:2 * no]] + [eta[2 * no + 2 * b * n + n:2 * no + 2 * b * n + 2 * n] for b in range(m)])
    mu = (ey - 1j * ex) * spb * 2 * np.pi / nsrc
    z = pxb + 1j * pyb
    dip1 = 0.25 / np.pi * mu * z
    dip2 = 0.25 / np.pi * (mu * np.conj(z) - np.conj(mu) * z)
    zz = xb + 1j * yb
    zd = zz[:, None] - zz[None, :]
    np.fill_diagonal(zd, 1.0)
    z1 = 1.0 / zd
    terms = dip1[None, :] * z1 + dip2[None, :] * np.conj(z1) - np.conj(dip1[None, :] * z1 ** 2) * zd
    np.fill_diagonal(terms, 0.0)
    vel = terms.sum(axis=1)
    u, w = -vel.imag, vel.real
    far = np.zeros_like(v)
    far[:no], far[no:2 * no] = u[:no], w[:no]
    for b in range(m):
        far[2 * no + 2 * b * n:2 * no + 2 * b * n + n] = u[no + b * n:no + b * n + n]
        far[2 * no + 2 * b * n + n:2 * no + 2 * b * n + 2 * n] = w[no + b * n:no + b * n + n]
    # local terms restated independently from stokesSolver.f:932-987, 1102-1118
    loc = np.zeros_like(v)
    cur = np.concatenate([geo.array("cur0"), geo.array("cur")])
    td = pyb * ex - pxb * ey
    lu = -cur * td * pyb * spb / nsrc - 0.5 * ex
    lv = cur * td * pxb * spb / nsrc - 0.5 * ey
    sig = np.sum((geo.array("px0") * ex[:no] + geo.array("py0") * ey[:no]) * geo.array("speed0")) * 2 * np.pi / no
    lu[:no] += sig * geo.array("px0")
    lv[:no] += sig * geo.array("py0")
    loc[:no], loc[no:2 * no] = lu[:no], lv[:no]
    for b in range(m):
        loc[2 * no + 2 * b * n:2 * no + 2 * b * n + n] = lu[no + b * n:no + b * n + n]
        loc[2 * no + 2 * b * n + n:2 * no + 2 * b * n + 2 * n] = lv[no + b * n:no + b * n + n]
    rows = slice(0, 2 * no + 2 * n * m)
    assert relerr(v[rows], far[rows] + loc[rows]) < 1e-12


def test_constant_density_identity():
    """Stokes double layer of a constant density: with the inward normal it is
    c/2 on the curve (cancelled by the jump -c/2) and 0 in the fluid, so the
    operator maps eta = c on one body to ~0 on every velocity row."""
    n, m, no = 128, 2, 512
    th = 2 * np.pi * np.arange(n) / n
    xx = np.concatenate([-0.3 + 0.2 * np.cos(th) + 0.02 * np.cos(3 * th), 0.4 + 0.15 * np.cos(th)])
    yy = np.concatenate([0.1 + 0.2 * np.sin(th), -0.2 + 0.25 * np.sin(th)])
    geo = O.Geometry(n, m, no, xx, yy)
    eta = np.zeros(geo.nunk)
    eta[2 * no:2 * no + n] = 0.7  # eta_x on body 0
    eta[2 * no + n:2 * no + 2 * n] = -0.4  # eta_y on body 0
    v = geo.matvec(eta, ibary=0)
    assert np.max(np.abs(v[:2 * no + 2 * n * m])) < 1e-10
    # constraint rows: -(1/n) sum eta * speed = -c L / (2 pi)
    L0 = np.sum(geo.array("speed")[:n]) * 2 * np.pi / n
    c = v[2 * no + 2 * n * m:]
    assert abs(c[0] + 0.7 * L0 / (2 * np.pi)) < 1e-12
    assert abs(c[1] - 0.4 * L0 / (2 * np.pi)) < 1e-12


def test_stokeslet_rotlet_closed_form():
    """Stokeslet/rotlet completion (stokesSolver.f:992-1049) against the closed form."""
    n, m, no = 32, 1, 64
    th = 2 * np.pi * np.arange(n) / n
    geo = O.Geometry(n, m, no, 0.1 + 0.2 * np.cos(th), 0.2 * np.sin(th))
    eta = np.zeros(geo.nunk)
    s1, s2, rot = 0.3, -0.2, 0.5
    eta[-3:] = [s1, s2, rot]
    v = geo.matvec(eta, ibary=0)
    xo, yo = geo.array("xo"), geo.array("yo")
    cx, cy = geo.array("cx")[0], geo.array("cy")[0]
    rx, ry = xo - cx, yo - cy
    r2 = rx * rx + ry * ry
    u = (-0.5 * np.log(r2) * s1 + (rx * s1 + ry * s2) / r2 * rx) / (4 * np.pi) + rot * ry / r2
    w = (-0.5 * np.log(r2) * s2 + (rx * s1 + ry * s2) / r2 * ry) / (4 * np.pi) - rot * rx / r2
    assert np.allclose(v[:no], u, rtol=1e-13, atol=1e-15)
    assert np.allclose(v[no:2 * no], w, rtol=1e-13, atol=1e-15)
    assert np.allclose(v[-3:], [s1, s2, rot], rtol=0, atol=1e-15)


# ------------------------------------------------------------ barycentric
def test_exterior_interior_dlp_match_reference(golden):
    """StokesExteriorDLP / StokesInteriorDLP with nder=2 (stokesSolver.f:5233-5379, 4937-5100)."""
    g = golden("dlp_n64")
    ext = O.stokes_dlp(True, g["xs"], g["ys"], g["dx"], g["dy"], g["px"], g["py"], g["xe"], g["ye"], 2)
    inn = O.stokes_dlp(False, g["xs"], g["ys"], g["dx"], g["dy"], g["px"], g["py"], g["xi"], g["yi"], 2)
    for i in range(6):
        assert relerr(ext[i], g["ext"][i]) < 1e-11, i
        assert relerr(inn[i], g["inn"][i]) < 1e-11, i


def test_exterior_dlp_coincident_branch(golden):
    """The coincident-node branch of laplaceExteriorHolomorphic (stokesSolver.f:5492-5506)."""
    g = golden("dlp_n64")
    c = O.stokes_dlp(True, g["xs"], g["ys"], g["dx"], g["dy"], g["px"], g["py"], g["xc"], g["yc"], 2)
    ref = g["coin"]
    # velocity survives the branch; the derivative terms are cancellation-dominated
    assert relerr(c[0], ref[0]) < 1e-6 and relerr(c[1], ref[1]) < 1e-6
    assert np.all(np.isfinite(np.array(c)))


def test_barycentric_far_target_equals_trapezoid(golden):
    """Far from the curve the barycentric Stokes DLP equals the trapezoid sum."""
    g = golden("dlp_n64")
    xs, ys, dx, dy, px, py = (g[k] for k in ("xs", "ys", "dx", "dy", "px", "py"))
    geo = O.Geometry(len(xs), 1, 64, xs, ys)
    sp = geo.array("speed")
    ang = np.array([0.3, 2.1, 4.0])
    xt = np.mean(xs) + 0.6 * np.cos(ang)  # 0.6 from the centre, radius 0.2
    yt = np.mean(ys) + 0.6 * np.sin(ang)
    ux, uy = O.stokes_dlp(True, xs, ys, dx, dy, px, py, xt, yt, 1)[:2]
    rx = xt[:, None] - xs[None, :]
    ry = yt[:, None] - ys[None, :]
    r2 = rx * rx + ry * ry
    c = (rx * px + ry * py) * (rx * dx + ry * dy) / r2 ** 2 * sp * 2 / len(xs)
    assert np.allclose(ux, (c * rx).sum(1), rtol=1e-10, atol=1e-13)
    assert np.allclose(uy, (c * ry).sum(1), rtol=1e-10, atol=1e-13)


# ------------------------------------------------------------------ GMRES
@pytest.mark.parametrize("name", FIXTURES)
def test_gmres_matches_reference(golden, name):
    """DGMRES as solveBIE configures it (stokesSolver.f:481-497, gmres_routines.f)."""
    g = golden(name)
    geo = _geom(g)
    den, it, err, ierr = geo.gmres(geo.rhs(0), ibary=1, maxl=2000, tol=1e-8)
    assert ierr == 0 and it == int(g["iter_tol8"])
    assert relerr(den, g["den_tol8"]) < 1e-9
    den12, it12, err12, _ = geo.gmres(geo.rhs(0), ibary=1, maxl=2000, tol=1e-12)
    assert relerr(den12, g["den_tol12"]) < 1e-10
    # the converged density solves the system to (close to) the tolerance
    b = geo.rhs(0)
    assert np.linalg.norm(b - geo.matvec(den12, 1)) / np.linalg.norm(b) < 1e-10


def test_gmres_small_system_dense_solve():
    """GMRES to 1e-13 agrees with a dense solve of the same (small) operator."""
    n, m, no = 16, 1, 32
    th = 2 * np.pi * np.arange(n) / n
    geo = O.Geometry(n, m, no, 0.3 * np.cos(th), 0.2 * np.sin(th))
    N = geo.nunk
    A = np.stack([geo.matvec(e, ibary=0) for e in np.eye(N)], axis=1)
    b = geo.rhs(0)
    x, it, err, ierr = geo.gmres(b, ibary=0, maxl=N, tol=1e-13)
    assert relerr(x, np.linalg.solve(A, b)) < 1e-10


# ----------------------------------------------------------------- stress
@pytest.mark.parametrize("name", FIXTURES)
def test_shear_stress_matches_reference(golden, name):
    """computeShearStress / deformation_on_boundary (stokesSolver.f:3002-3096, 2618-2998)."""
    g = golden(name)
    tau = O.shear_stress(int(g["n"]), int(g["m"]), int(g["no"]), g["xx"], g["yy"], g["den_tol12"])
    assert relerr(tau, g["tau_tol12"]) < 1e-10


def test_deformation_is_traceless(golden):
    """Incompressibility: E11 + E22 = 0 at every body point."""
    g = golden("tight8_n64")
    e11, e12, e22 = _geom(g).deformation(g["den_tol12"])
    assert np.max(np.abs(e11 + e22)) < 1e-8 * np.max(np.abs(e11))


# -------------------------------------------------------------------- QoI
@pytest.mark.parametrize("name", FIXTURES)
@pytest.mark.parametrize("ibary,key", [(1, "qoi_bary"), (0, "qoi_trap")])
def test_qoi_targets_match_reference(golden, name, ibary, key):
    """computeQoiTargets + classifyPointsOld (stokesSolver.f:3099-3323, 4420-4500)."""
    g = golden(name)
    out = O.qoi_targets(int(g["n"]), int(g["m"]), int(g["no"]), ibary, g["xx"], g["yy"],
                        g["den_tol12"], g["qoi_xt"], g["qoi_yt"])
    for i in range(4):
        assert relerr(out[i], g[key][i]) < 1e-10, i
    # body centres are inside: all four quantities zeroed
    m = int(g["m"])
    centres = slice(26, 26 + m)
    assert np.all(np.array(out)[:, centres] == 0.0)


@pytest.mark.parametrize("name", ["tight8_n64", "bodies10_n128"])
def test_pressure_matches_reference(golden, name):
    """computePressure (stokesSolver.f:4644-4882) against the reference's own build."""
    g = golden(name)
    n, m, no = int(g["n"]), int(g["m"]), int(g["no"])
    p = O.pressure(n, m, no, g["xx"], g["yy"], g["den_tol12"])
    # the densities reach 1e3 while the pressure is O(10): cancellation bounds
    # the agreement of two FP64 evaluations (FFT versus direct derivative)
    assert relerr(p, g["pressure_tol12"]) < 1e-10
