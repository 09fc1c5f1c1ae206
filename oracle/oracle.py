"""ctypes binding of the C restatement in oracle/stokes_oracle.c.

TEST INFRASTRUCTURE ONLY: the checker that tests/, __graft_entry__.smoke()
and bench.py's cpu_baseline leg compare the HIP library against.

Parity status of each restated function (how each one is pinned):
  fourier_diff, geometry  closed forms (circle, trig polynomials) and the
                          flang-built reference goldens
  rhs                     closed form (pipe flow / uniform flow on the wall)
  matvec ibary=0/1        reference goldens (matvec_DLP, matvec_DLP_fmmbary;
                          the latter to FMM accuracy), the dipole-form identity,
                          the constant-density jump identity
  exterior/interior DLP   reference goldens; holomorphic test data
  gmres / stokessolver    reference goldens (iteration count and density),
                          dense numpy solve on small systems
  deformation / stress    reference goldens; symmetric-traceless invariant
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


def build() -> None:
    subprocess.run(["make", "-s", "-C", _HERE, "liboracle.so"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.orc_geom_create.restype = ctypes.c_void_p
        L.orc_geom_create.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp, _dp]
        L.orc_geom_free.argtypes = [ctypes.c_void_p]
        L.orc_geom_array.restype = _dp
        L.orc_geom_array.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.orc_geom_set_array.restype = ctypes.c_int
        L.orc_geom_set_array.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp]
        L.orc_fourier_diff.argtypes = [ctypes.c_int, _dp, _dp]
        L.orc_rhs.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp]
        L.orc_matvec.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, _dp]
        dlp_args = [ctypes.c_int, _dp, _dp, _dp, _dp, _dp, _dp, ctypes.c_int, _dp, _dp, ctypes.c_int,
                    _dp, _dp, _dp, _dp, _dp, _dp]
        L.orc_stokes_exterior_dlp.argtypes = dlp_args
        L.orc_stokes_interior_dlp.argtypes = dlp_args
        L.orc_gmres.restype = ctypes.c_int
        L.orc_gmres.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, _dp, ctypes.c_int, ctypes.c_double,
                                _dp, _ip]
        L.orc_deformation.argtypes = [ctypes.c_void_p, _dp, _dp, _dp, _dp]
        L.orc_shear_stress.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp, _dp, _dp, _dp]
        L.orc_qoi_targets.argtypes = [ctypes.c_int] * 4 + [_dp, _dp, _dp, ctypes.c_int] + [_dp] * 6
        L.orc_set_threads.argtypes = [ctypes.c_int]
        L.orc_get_threads.restype = ctypes.c_int
        _lib = L
    return _lib


def _p(a: np.ndarray):
    return a.ctypes.data_as(_dp)


def _f64(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64)


def set_threads(n: int) -> None:
    lib().orc_set_threads(int(n))


def get_threads() -> int:
    return int(lib().orc_get_threads())


def nunk(n: int, m: int, no: int) -> int:
    return 2 * no + 2 * n * m + 3 * m


class Geometry:
    """Oracle geometry of m bodies x n points plus the fixed wall of no points."""

    NAMES = ("px", "py", "cur", "speed", "cx", "cy", "xo", "yo", "px0", "py0", "cur0", "speed0")

    def __init__(self, n: int, m: int, no: int, xx, yy):
        self.n, self.m, self.no = int(n), int(m), int(no)
        self.xx, self.yy = _f64(xx), _f64(yy)
        self._h = lib().orc_geom_create(self.n, self.m, self.no, _p(self.xx), _p(self.yy))

    def __del__(self):
        if getattr(self, "_h", None):
            lib().orc_geom_free(self._h)
            self._h = None

    @property
    def nunk(self) -> int:
        return nunk(self.n, self.m, self.no)

    def array(self, name: str) -> np.ndarray:
        which = self.NAMES.index(name)
        size = {"cx": self.m, "cy": self.m}.get(name, self.n * self.m if which < 4 else self.no)
        ptr = lib().orc_geom_array(self._h, which)
        return np.ctypeslib.as_array(ptr, shape=(size,)).copy() if size else np.zeros(0)

    def set_array(self, name: str, values) -> None:
        """Overwrite one geometry array (same name and length as ``array``)."""
        which = self.NAMES.index(name)
        v = _f64(values)
        size = {"cx": self.m, "cy": self.m}.get(name, self.n * self.m if which < 4 else self.no)
        if len(v) != size:
            raise ValueError(f"{name} holds {size} values, got {len(v)}")
        if lib().orc_geom_set_array(self._h, which, _p(v)) != 0:
            raise RuntimeError("orc_geom_set_array failed")

    def rhs(self, ibc: int) -> np.ndarray:
        out = np.zeros(self.nunk)
        lib().orc_rhs(self._h, int(ibc), _p(out))
        return out

    def matvec(self, den, ibary: int = 1) -> np.ndarray:
        den = _f64(den)
        out = np.zeros(self.nunk)
        lib().orc_matvec(self._h, int(ibary), _p(den), _p(out))
        return out

    def gmres(self, rhs, x0=None, ibary: int = 1, maxl: int = 8000, tol: float = 1e-8):
        rhs = _f64(rhs)
        x = np.zeros(self.nunk) if x0 is None else _f64(x0).copy()
        err = ctypes.c_double(0.0)
        ierr = ctypes.c_int(0)
        it = lib().orc_gmres(self._h, int(ibary), _p(rhs), _p(x), int(maxl), float(tol),
                             ctypes.byref(err), ctypes.byref(ierr))
        return x, int(it), float(err.value), int(ierr.value)

    def shear_stress(self, den) -> np.ndarray:
        """computeShearStress (ib=1) on this geometry: tau = t . (p_y, -p_x),
        t = -2 E n (stokesSolver.f:3075-3091)."""
        e11, e12, e22 = self.deformation(den)
        px, py = self.array("px"), self.array("py")
        tx = -2.0 * e11 * px - 2.0 * e12 * py
        ty = -2.0 * e12 * px - 2.0 * e22 * py
        return tx * py - ty * px

    def deformation(self, den):
        den = _f64(den)
        nb = self.n * self.m
        e11, e12, e22 = np.zeros(nb), np.zeros(nb), np.zeros(nb)
        lib().orc_deformation(self._h, _p(den), _p(e11), _p(e12), _p(e22))
        return e11, e12, e22


def fourier_diff(f: np.ndarray) -> np.ndarray:
    f = np.asarray(f, dtype=np.complex128)
    re, im = _f64(f.real).copy(), _f64(f.imag).copy()
    lib().orc_fourier_diff(len(f), _p(re), _p(im))
    return re + 1j * im


def stokes_dlp(exterior: bool, xs, ys, dx, dy, px1, py1, xt, yt, nder: int = 1):
    xs, ys, dx, dy, px1, py1, xt, yt = map(_f64, (xs, ys, dx, dy, px1, py1, xt, yt))
    nt = len(xt)
    outs = [np.zeros(nt) for _ in range(6)]
    fn = lib().orc_stokes_exterior_dlp if exterior else lib().orc_stokes_interior_dlp
    fn(len(xs), _p(xs), _p(ys), _p(dx), _p(dy), _p(px1), _p(py1), nt, _p(xt), _p(yt), int(nder),
       *[_p(o) for o in outs])
    return tuple(outs)


def stokessolver(n, m, no, xx, yy, ibary=1, ibc=0, maxl=8000, tol=1e-8, den0=None):
    """Density solve; returns (den, iterations)."""
    g = Geometry(n, m, no, xx, yy)
    x, it, _err, _ierr = g.gmres(g.rhs(ibc), den0, ibary=ibary, maxl=maxl, tol=tol)
    return x, it


def qoi_targets(n, m, no, ibary, xx, yy, den, xt, yt):
    """(u, v, p, vorticity) at targets, computeQoiTargets."""
    xx, yy, den, xt, yt = map(_f64, (xx, yy, den, xt, yt))
    nt = len(xt)
    outs = [np.zeros(nt) for _ in range(4)]
    lib().orc_qoi_targets(int(n), int(m), int(no), int(ibary), _p(xx), _p(yy), _p(den), nt,
                          _p(xt), _p(yt), *[_p(o) for o in outs])
    return tuple(outs)


def shear_stress(n, m, no, xx, yy, den) -> np.ndarray:
    xx, yy, den = _f64(xx), _f64(yy), _f64(den)
    tau = np.zeros(n * m)
    lib().orc_shear_stress(int(n), int(m), int(no), _p(xx), _p(yy), _p(den), _p(tau))
    return tau


def pressure(n, m, no, xx, yy, den) -> np.ndarray:
    """computePressure (stokesSolver.f:4644-4882) restated in C."""
    xx, yy, den = _f64(xx), _f64(yy), _f64(den)
    out = np.zeros(n * m)
    lib().orc_pressure(int(n), int(m), int(no), _p(xx), _p(yy), _p(den), _p(out))
    return out
