"""ctypes binding of the reference's Fortran, built from source by
``make -C oracle ref`` into oracle/_ref/libstokes_ref.so.

TEST INFRASTRUCTURE ONLY.  Used in this container to write golden vectors
(tests/golden/make_golden.py) and to pin the C restatement.  The routines are
called exactly as the reference calls them internally: geometry goes into the
COMMON blocks /geometry/ and /wall/ (stokesSolver.f:58-60) through
inner_geometry / outer_geometry, the matvecs take the DGMRES MATVEC argument
list, and a tight-tolerance solve calls DGMRES with solveBIE's settings
(stokesSolver.f:481-497) and an explicit TOL instead of the hard-coded 1e-8.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_ref", "libstokes_ref.so")
NMAX = 2 ** 17  # stokesSolver.f:11
MAXBODIES = 500  # stokesSolver.f:12

_dp = ctypes.POINTER(ctypes.c_double)
_lib = None


def available() -> bool:
    return os.path.exists(LIB_PATH)


def lib():
    global _lib
    if _lib is None:
        _lib = ctypes.CDLL(LIB_PATH)
    return _lib


def _i(v):
    return ctypes.byref(ctypes.c_int32(int(v)))


def _i64(v):
    return ctypes.byref(ctypes.c_int64(int(v)))


def _p(a):
    return a.ctypes.data_as(_dp)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


class _Common:
    """Pointers into COMMON /geometry/ and /wall/ (stokesSolver.f:58-60)."""

    def __init__(self):
        L = lib()
        g = ctypes.addressof(ctypes.c_double.in_dll(L, "geometry_"))
        w = ctypes.addressof(ctypes.c_double.in_dll(L, "wall_"))
        d = 8
        self.x = g
        self.y = g + d * NMAX
        self.cx = g + 2 * d * NMAX
        self.cy = self.cx + d * MAXBODIES
        self.px = self.cy + d * MAXBODIES
        self.py = self.px + d * NMAX
        self.cur = self.py + d * NMAX
        self.speed = self.cur + d * NMAX
        self.ninner = self.speed + d * NMAX
        self.nbodies = self.ninner + 4
        self.xo = w
        self.yo = w + d * NMAX
        self.px0 = w + 2 * d * NMAX
        self.py0 = w + 3 * d * NMAX
        self.cur0 = w + 4 * d * NMAX
        self.speed0 = w + 5 * d * NMAX
        self.nouter = w + 6 * d * NMAX

    @staticmethod
    def dptr(addr):
        return ctypes.cast(addr, _dp)

    def read(self, addr, count):
        return np.ctypeslib.as_array(self.dptr(addr), shape=(count,)).copy()


def setup(n, m, no, xx, yy) -> _Common:
    """Fill the COMMON blocks as stokesSolver does (stokesSolver.f:64-72)."""
    L = lib()
    c = _Common()
    xx, yy = _f64(xx), _f64(yy)
    ctypes.c_int32.from_address(c.ninner).value = n
    ctypes.c_int32.from_address(c.nbodies).value = m
    ctypes.c_int32.from_address(c.nouter).value = no
    dp = c.dptr
    L.inner_geometry_(_i(n), _i(m), dp(c.x), dp(c.y), _p(xx), _p(yy), dp(c.px), dp(c.py),
                      dp(c.cur), dp(c.speed), dp(c.cx), dp(c.cy))
    L.outer_geometry_(_i(no), dp(c.xo), dp(c.yo), dp(c.px0), dp(c.py0), dp(c.cur0), dp(c.speed0))
    return c


def geometry_arrays(n, m, no, xx, yy) -> dict:
    c = setup(n, m, no, xx, yy)
    nb = n * m
    return {
        "px": c.read(c.px, nb), "py": c.read(c.py, nb), "cur": c.read(c.cur, nb),
        "speed": c.read(c.speed, nb), "cx": c.read(c.cx, m), "cy": c.read(c.cy, m),
        "xo": c.read(c.xo, no), "yo": c.read(c.yo, no), "px0": c.read(c.px0, no),
        "py0": c.read(c.py0, no), "cur0": c.read(c.cur0, no), "speed0": c.read(c.speed0, no),
    }


def _matvec_fn(ibary: int, ifmm: int):
    L = lib()
    if ibary == 1:
        return L.matvec_dlp_fmmbary_ if ifmm == 1 else L.matvec_dlp_bary_
    return L.matvec_dlp_fmm_ if ifmm == 1 else L.matvec_dlp_


def matvec(n, m, no, xx, yy, den, ibary=1, ifmm=1) -> np.ndarray:
    """One application of the reference operator selected as solveBIE does."""
    setup(n, m, no, xx, yy)
    den = _f64(den).copy()
    N = len(den)
    vel = np.zeros(N)
    dummy_i = ctypes.c_int32(0)
    dummy_d = ctypes.c_double(0.0)
    _matvec_fn(ibary, ifmm)(_i(N), _p(den), _p(vel), _i(0), ctypes.byref(dummy_i),
                            ctypes.byref(dummy_i), ctypes.byref(dummy_d), _i(0))
    return vel


def rhs(n, m, no, xx, yy, ibc) -> np.ndarray:
    c = setup(n, m, no, xx, yy)
    N = 2 * no + 2 * n * m + 3 * m
    out = np.zeros(N)
    lib().bd_condition_(_i(n), _i(m), _i(ibc), c.dptr(c.x), c.dptr(c.y), _i(no), c.dptr(c.xo),
                        c.dptr(c.yo), _p(out))
    return out


def solve(n, m, no, xx, yy, ibary=1, ifmm=1, ibc=0, maxl=2000, tol=1e-8, den0=None):
    """DGMRES with solveBIE's configuration and an explicit tolerance.

    Returns (den, iter, err, ierr)."""
    b = rhs(n, m, no, xx, yy, ibc)  # also fills the COMMON blocks
    N = len(b)
    x = np.zeros(N) if den0 is None else _f64(den0).copy()
    maxl = min(maxl, N)
    lrgw = 10 + N * (maxl + 6) + maxl * (maxl + 3)
    rgwk = np.zeros(lrgw)
    igwk = np.zeros(30, dtype=np.int32)
    igwk[0] = maxl
    igwk[3] = 0  # no preconditioner (stokesSolver.f:490)
    igwk[4] = -1  # no restart (stokesSolver.f:497)
    iwork = np.array([n, m, no], dtype=np.int32)
    it = ctypes.c_int32(0)
    err = ctypes.c_double(0.0)
    ierr = ctypes.c_int32(0)
    dummy_d = ctypes.c_double(0.0)
    dummy_i = ctypes.c_int32(0)
    sb = np.zeros(1)
    sx = np.zeros(1)
    rwork = np.zeros(1)
    L = lib()
    mv = ctypes.cast(_matvec_fn(ibary, ifmm), ctypes.c_void_p)
    ms = ctypes.cast(L.msolve_dlp_, ctypes.c_void_p)
    L.dgmres_(_i(N), _p(b), _p(x), _i(0), ctypes.byref(dummy_i), ctypes.byref(dummy_i),
              ctypes.byref(dummy_d), _i(0), mv, ms, _i(0), ctypes.byref(ctypes.c_double(tol)),
              _i(0), ctypes.byref(it), ctypes.byref(err), ctypes.byref(ierr), _i(0), _p(sb), _p(sx),
              _p(rgwk), _i(lrgw), igwk.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), _i(30),
              _p(rwork), iwork.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)))
    return x, int(it.value), float(err.value), int(ierr.value)


def stokessolver(n, m, no, xx, yy, ifmm=1, ibary=1, ibc=0, maxl=2000):
    """The shipped entry point (GMRES tol 1e-8), called with int64 arguments
    as DensityStress.jl:48-51 does."""
    xx, yy = _f64(xx), _f64(yy)
    N = 2 * no + 2 * n * m + 3 * m
    den = np.zeros(N)
    nits = ctypes.c_int64(0)
    lib().stokessolver_(_i64(n), _i64(m), _i64(no), _i64(ifmm), _i64(ibary), _i64(ibc), _i64(maxl),
                        _p(xx), _p(yy), _p(den), ctypes.byref(nits))
    return den, int(nits.value)


def exterior_dlp(xs, ys, dx, dy, px1, py1, xt, yt, nder=1):
    return _dlp(lib().stokesexteriordlp_, True, xs, ys, dx, dy, px1, py1, xt, yt, nder)


def interior_dlp(xs, ys, dx, dy, px1, py1, xt, yt, nder=1):
    return _dlp(lib().stokesinteriordlp_, False, xs, ys, dx, dy, px1, py1, xt, yt, nder)


def _dlp(fn, exterior, xs, ys, dx, dy, px1, py1, xt, yt, nder):
    xs, ys, dx, dy, px1, py1, xt, yt = map(_f64, (xs, ys, dx, dy, px1, py1, xt, yt))
    npts, nt = len(xs), len(xt)
    outs = [np.zeros(nt) for _ in range(6)]
    if exterior:
        # StokesExteriorDLP(npts,nptsc,nbeta,xsou,ysou,densx,densy,px,py,ntar,xtar,ytar,isou,n,nder,...)
        fn(_i(npts), _i(npts), _i(1), _p(xs), _p(ys), _p(dx), _p(dy), _p(px1), _p(py1), _i(nt),
           _p(xt), _p(yt), _i(1), _i(1), _i(nder), *[_p(o) for o in outs])
    else:
        # StokesInteriorDLP(npts,nptsc,nbeta,xsou,ysou,dens1,dens2,nx,ny,ntar,xtar,ytar,nder,...)
        fn(_i(npts), _i(npts), _i(1), _p(xs), _p(ys), _p(dx), _p(dy), _p(px1), _p(py1), _i(nt),
           _p(xt), _p(yt), _i(nder), *[_p(o) for o in outs])
    return tuple(outs)


def qoi_targets(n, m, no, ibary, xx, yy, den, xt, yt):
    """computeqoitargets_ called with int64 arguments as DensityStress.jl:81-86 does."""
    xx, yy, den = _f64(xx).copy(), _f64(yy).copy(), _f64(den)
    xt, yt = _f64(xt).copy(), _f64(yt).copy()
    nt = len(xt)
    outs = [np.zeros(nt) for _ in range(4)]
    lib().computeqoitargets_(_i64(n), _i64(m), _i64(no), _i64(ibary), _p(xx), _p(yy), _p(den),
                             _i64(nt), _p(xt), _p(yt), *[_p(o) for o in outs])
    return tuple(outs)


def shear_stress(n, m, no, xx, yy, den, ib=1) -> np.ndarray:
    """computeshearstress_ called with int64 arguments as DensityStress.jl:195-198 does."""
    xx, yy, den = _f64(xx).copy(), _f64(yy).copy(), _f64(den)
    tau = np.zeros(n * m)
    lib().computeshearstress_(_i64(n), _i64(m), _i64(no), _p(xx), _p(yy), _p(den), _i64(ib), _p(tau))
    return tau


def pressure(n, m, no, xx, yy, den) -> np.ndarray:
    """computepressure_ called with int64 arguments as PostProcess.jl:107-110 does."""
    xx, yy, den = _f64(xx).copy(), _f64(yy).copy(), _f64(den)
    out = np.zeros(n * m)
    lib().computepressure_(_i64(n), _i64(m), _i64(no), _p(xx), _p(yy), _p(den), _p(out))
    return out
