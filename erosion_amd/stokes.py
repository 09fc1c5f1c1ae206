"""ctypes binding of libstokes_hip.so (C-ABI in include/erosion_stokes.h).

The library is the only compute path: loading fails loudly when it has not
been built (``python -m erosion_amd.build``) and every call checks the
library's status channel, raising on error.
"""
from __future__ import annotations

import ctypes
import os
from typing import Optional, Tuple

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libstokes_hip.so")

_i64p = ctypes.POINTER(ctypes.c_int64)
_dp = ctypes.POINTER(ctypes.c_double)
_lib = None


class StokesError(RuntimeError):
    pass


def lib():
    """Load libstokes_hip.so; raise if it is missing."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise StokesError(f"{LIB_PATH} not built: run `python -m erosion_amd.build`")
        L = ctypes.CDLL(LIB_PATH)
        L.stokessolver_.argtypes = [_i64p] * 7 + [_dp, _dp, _dp, _i64p]
        L.stokessolver_.restype = None
        L.computeshearstress_.argtypes = [_i64p, _i64p, _i64p, _dp, _dp, _dp, _i64p, _dp]
        L.computeshearstress_.restype = None
        L.computeqoitargets_.argtypes = [_i64p] * 4 + [_dp, _dp, _dp, _i64p] + [_dp] * 6
        L.computeqoitargets_.restype = None
        L.computepressure_.argtypes = [_i64p, _i64p, _i64p, _dp, _dp, _dp, _dp]
        L.computepressure_.restype = None
        L.erosion_set_gmres_tol.argtypes = [ctypes.c_double]
        L.erosion_set_gmres_tol.restype = ctypes.c_int
        L.erosion_get_gmres_tol.restype = ctypes.c_double
        L.erosion_last_status.restype = ctypes.c_int
        L.erosion_last_error.restype = ctypes.c_char_p
        L.erosion_last_solve_info.argtypes = [_dp]
        L.erosion_last_solve_info.restype = ctypes.c_int
        L.erosion_matvec.argtypes = [ctypes.c_int64] * 4 + [_dp, _dp, _dp, _dp]
        L.erosion_matvec.restype = ctypes.c_int
        L.erosion_time_matvec.argtypes = [ctypes.c_int64] * 4 + [_dp, _dp, _dp, ctypes.c_int64, _dp]
        L.erosion_time_matvec.restype = ctypes.c_int
        L.erosion_kernel_timing.argtypes = [ctypes.c_int]
        L.erosion_kernel_timing.restype = ctypes.c_int
        L.erosion_kernel_timers.argtypes = [_dp]
        L.erosion_kernel_timers.restype = ctypes.c_int
        L.erosion_set_precision.argtypes = [ctypes.c_int]
        L.erosion_set_precision.restype = ctypes.c_int
        L.erosion_get_precision.restype = ctypes.c_int
        L.erosion_set_mixed_switch.argtypes = [ctypes.c_double]
        L.erosion_set_mixed_switch.restype = ctypes.c_int
        L.erosion_set_preconditioner.argtypes = [ctypes.c_int] * 3
        L.erosion_set_preconditioner.restype = ctypes.c_int
        L.erosion_get_preconditioner.argtypes = [ctypes.POINTER(ctypes.c_int)]
        L.erosion_get_preconditioner.restype = ctypes.c_int
        L.erosion_set_farfield_shape.argtypes = [ctypes.c_int]
        L.erosion_set_farfield_shape.restype = ctypes.c_int
        L.erosion_farfield_shape_count.argtypes = []
        L.erosion_farfield_shape_count.restype = ctypes.c_int
        L.erosion_set_graphs.argtypes = [ctypes.c_int]
        L.erosion_set_graphs.restype = ctypes.c_int
        L.erosion_graph_stats.argtypes = [ctypes.POINTER(ctypes.c_longlong)]
        L.erosion_graph_stats.restype = ctypes.c_int
        L.erosion_precond_apply.argtypes = [ctypes.c_int64] * 3 + [_dp] * 4
        L.erosion_precond_apply.restype = ctypes.c_int
        L.erosion_rungekutta2.argtypes = [ctypes.c_int64] * 3 + [_dp] * 6 + [_i64p, _dp]
        L.erosion_stepper_history.argtypes = [_i64p]
        L.erosion_stepper_history.restype = ctypes.c_int
        L.erosion_rungekutta2.restype = ctypes.c_int
        L.erosion_stepper_reset.argtypes = []
        L.erosion_stepper_reset.restype = ctypes.c_int
        L.erosion_spectral.argtypes = [ctypes.c_int, ctypes.c_int64, ctypes.c_int64, _dp, _dp, _dp]
        L.erosion_spectral.restype = ctypes.c_int
        L.erosion_set_guess_space.argtypes = [ctypes.c_int]
        L.erosion_set_guess_space.restype = ctypes.c_int
        L.erosion_last_preconditioned.argtypes = [_dp, ctypes.c_int64]
        L.erosion_last_preconditioned.restype = ctypes.c_int
        L.erosion_get_guess_space.argtypes = []
        L.erosion_get_guess_space.restype = ctypes.c_int
        _lib = L
    return _lib


def _p(a: np.ndarray):
    return a.ctypes.data_as(_dp)


def _ref(v: int):
    return ctypes.byref(ctypes.c_int64(int(v)))


def _f64(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64)


def nunk(npts: int, nbods: int, nouter: int) -> int:
    return 2 * nouter + 2 * npts * nbods + 3 * nbods


def _lengths(npts: int, nbods: int, nouter: int, xx, yy, den=None) -> None:
    """The library reads exactly these lengths from the caller's buffers."""
    if len(xx) < npts * nbods or len(yy) < npts * nbods:
        raise StokesError(f"xx, yy hold {len(xx)}, {len(yy)} values; need {npts * nbods}")
    if den is not None and len(den) < nunk(npts, nbods, nouter):
        raise StokesError(f"density holds {len(den)} values; need {nunk(npts, nbods, nouter)}")


def last_status() -> int:
    return int(lib().erosion_last_status())


def last_error() -> str:
    return lib().erosion_last_error().decode()


def _check(where: str) -> None:
    st = last_status()
    if st < 0:
        raise StokesError(f"{where}: {last_error()}")


def set_gmres_tol(tol: float) -> None:
    if lib().erosion_set_gmres_tol(float(tol)) != 0:
        raise StokesError(f"invalid GMRES tolerance {tol}")


def get_gmres_tol() -> float:
    return float(lib().erosion_get_gmres_tol())


def last_solve_info() -> dict:
    info = np.zeros(40)
    lib().erosion_last_solve_info(_p(info))

    def record(o):
        return {"stop": STOP_REASONS.get(int(info[o]), str(int(info[o]))),
                "true_relres": [float(v) for v in info[o + 1:o + 5]],
                "iters": [int(v) for v in info[o + 5:o + 9]],
                "fp64_floor_relres": float(info[o + 9])}

    return {"iters": int(info[0]), "implicit_relres": float(info[1]),
            "true_relres": float(info[2]), "seconds": float(info[3]), "reorth": int(info[4]),
            "ff_targets_per_lane": int(info[5]), "ff_splits": int(info[6]),
            "preconditioner": int(info[7]), "fp32_iters": int(info[8]),
            "tol_miss": int(info[9]), "tol_misses_total": int(info[10]), "cycles": int(info[11]),
            "r0_relres": float(info[12]), "cycle_record": record(13),
            "latest_miss_record": record(23), "tol_misses_at_floor_total": int(info[33])}


# why a solve stopped (erosion_last_solve_info, info[13] and info[23])
STOP_REASONS = {0: "met tolerance", 1: "restart did not halve the residual", 2: "cycle limit",
                3: "maxl spent", 4: "GMRES status nonzero"}


GUESS_DENSITY, GUESS_PRECONDITIONED = 0, 1


def set_guess_space(space: int) -> None:
    """Space of the initial guess stokessolver_ takes: GUESS_DENSITY (x0) or
    GUESS_PRECONDITIONED (u0 = P x0 of the right preconditioner)."""
    if lib().erosion_set_guess_space(int(space)) != 0:
        raise StokesError(f"invalid guess space {space}")


FARFIELD_DIRECT, FARFIELD_AUTO = 0, 1


def set_farfield_method(method: int) -> None:
    """FARFIELD_AUTO (default): the direct sums or the cluster
    expansions, whichever the autotuner measures faster; FARFIELD_DIRECT:
    the direct all-pairs sums only."""
    L = lib()
    L.erosion_set_farfield_method.argtypes = [ctypes.c_int]
    L.erosion_set_farfield_method.restype = ctypes.c_int
    if L.erosion_set_farfield_method(int(method)) != 0:
        raise StokesError(f"invalid far-field method {method}")


def get_guess_space() -> int:
    return int(lib().erosion_get_guess_space())


def last_preconditioned(npts: int, nbods: int, nouter: int) -> np.ndarray:
    """u = P x of the last solve (which must have run with GUESS_PRECONDITIONED)."""
    u = np.zeros(nunk(npts, nbods, nouter))
    if lib().erosion_last_preconditioned(_p(u), len(u)) != 0:
        raise StokesError("the last solve did not track the preconditioned variable")
    return u


def stokessolver(npts: int, nbods: int, nouter: int, ifmm: int, ibary: int, ibc: int, maxl: int,
                 xx, yy, den: Optional[np.ndarray] = None) -> Tuple[np.ndarray, int]:
    """Call stokessolver_ exactly as DensityStress.jl:48-51 does.

    ``den`` is the initial guess (zeros when omitted); returns (den, nits)."""
    xx, yy = _f64(xx), _f64(yy)
    _lengths(npts, nbods, nouter, xx, yy, den)
    out = np.zeros(nunk(npts, nbods, nouter)) if den is None else _f64(den).copy()
    nits = ctypes.c_int64(0)
    lib().stokessolver_(_ref(npts), _ref(nbods), _ref(nouter), _ref(ifmm), _ref(ibary), _ref(ibc),
                        _ref(maxl), _p(xx), _p(yy), _p(out), ctypes.byref(nits))
    _check("stokessolver_")
    return out, int(nits.value)


def computeshearstress(npts: int, nbods: int, nouter: int, xx, yy, den, ibary: int = 1) -> np.ndarray:
    """Call computeshearstress_ exactly as DensityStress.jl:195-198 does."""
    xx, yy, den = _f64(xx), _f64(yy), _f64(den)
    _lengths(npts, nbods, nouter, xx, yy, den)
    tau = np.zeros(npts * nbods)
    lib().computeshearstress_(_ref(npts), _ref(nbods), _ref(nouter), _p(xx), _p(yy), _p(den),
                              _ref(ibary), _p(tau))
    _check("computeshearstress_")
    return tau


def computepressure(npts: int, nbods: int, nouter: int, xx, yy, den) -> np.ndarray:
    """Call computepressure_ exactly as PostProcess.jl:107-110 does."""
    xx, yy, den = _f64(xx), _f64(yy), _f64(den)
    _lengths(npts, nbods, nouter, xx, yy, den)
    out = np.zeros(npts * nbods)
    lib().computepressure_(_ref(npts), _ref(nbods), _ref(nouter), _p(xx), _p(yy), _p(den),
                           _p(out))
    _check("computepressure_")
    return out


def geometry(npts: int, nbods: int, nouter: int, xx, yy) -> dict:
    """Normals, curvature, speed (wall first, then bodies) and body centres
    as the library builds them (test hook)."""
    xx, yy = _f64(xx), _f64(yy)
    _lengths(npts, nbods, nouter, xx, yy)
    np_ = nouter + npts * nbods
    out = np.zeros(4 * np_ + 2 * nbods)
    L = lib()
    L.erosion_geometry.argtypes = [ctypes.c_int64] * 3 + [_dp, _dp, _dp]
    L.erosion_geometry.restype = ctypes.c_int
    if L.erosion_geometry(npts, nbods, nouter, _p(xx), _p(yy), _p(out)) != 0:
        raise StokesError(f"erosion_geometry: {last_error()}")
    px, py, cur, sp = (out[a * np_:(a + 1) * np_] for a in range(4))
    return {"px": px, "py": py, "cur": cur, "speed": sp,
            "cx": out[4 * np_:4 * np_ + nbods], "cy": out[4 * np_ + nbods:]}


def time_matvec(npts: int, nbods: int, nouter: int, xx, yy, den, ibary: int = 1,
                reps: int = 10) -> dict:
    """Operator and far-field kernel durations measured with HIP events."""
    xx, yy, den = _f64(xx), _f64(yy), _f64(den)
    _lengths(npts, nbods, nouter, xx, yy, den)
    out = np.zeros(8)
    if lib().erosion_time_matvec(npts, nbods, nouter, ibary, _p(xx), _p(yy), _p(den), int(reps),
                                 _p(out)) != 0:
        raise StokesError(f"erosion_time_matvec: {last_error()}")
    return {"matvec_ms": out[0], "farfield_ms": out[1], "pairs": out[2], "near_entries": int(out[3]),
            "near_ms": out[4], "splits": int(out[5]), "gather_bytes": out[6]}


TIMER_FAMILIES = ("farfield", "near", "stress_eval", "gram_schmidt", "precond", "farfield32",
                  "fmm", "fmm_direct", "fmm_expand")


def kernel_timing(enable: bool, period: int = 1) -> None:
    """Start (clearing totals) or stop HIP-event timing of the hot kernels.
    period > 1: a GMRES iteration's launches are timed in one iteration of
    every `period`, each standing for `period` (totals stay estimates of all
    launches, per-launch means are measured)."""
    if period < 1:
        raise StokesError("period must be at least 1")
    if lib().erosion_kernel_timing(int(period) if enable else 0) != 0:
        raise StokesError(f"erosion_kernel_timing: {last_error()}")


def kernel_timers() -> dict:
    """{family: {"ms": total, "launches": n, "work": total work}} since enabling."""
    out = np.zeros(3 * len(TIMER_FAMILIES))
    if lib().erosion_kernel_timers(_p(out)) != 0:
        raise StokesError(f"erosion_kernel_timers: {last_error()}")
    return {f: {"ms": out[3 * i], "launches": int(out[3 * i + 1]), "work": out[3 * i + 2]}
            for i, f in enumerate(TIMER_FAMILIES)}


def computeqoitargets(npts: int, nbods: int, nouter: int, ibary: int, xx, yy, den, xtar, ytar):
    """Call computeqoitargets_ exactly as DensityStress.jl:81-86 does; returns (u, v, p, vort)."""
    xx, yy, den, xtar, ytar = map(_f64, (xx, yy, den, xtar, ytar))
    _lengths(npts, nbods, nouter, xx, yy, den)
    if len(ytar) != len(xtar):
        raise StokesError("xtar and ytar differ in length")
    nt = len(xtar)
    outs = [np.zeros(nt) for _ in range(4)]
    lib().computeqoitargets_(_ref(npts), _ref(nbods), _ref(nouter), _ref(ibary), _p(xx), _p(yy),
                             _p(den), _ref(nt), _p(xtar), _p(ytar), *[_p(o) for o in outs])
    _check("computeqoitargets_")
    return tuple(outs)


def matvec(npts: int, nbods: int, nouter: int, xx, yy, den, ibary: int = 1) -> np.ndarray:
    """One application of the completed double-layer operator."""
    xx, yy, den = _f64(xx), _f64(yy), _f64(den)
    _lengths(npts, nbods, nouter, xx, yy, den)
    vel = np.zeros(nunk(npts, nbods, nouter))
    if lib().erosion_matvec(npts, nbods, nouter, ibary, _p(xx), _p(yy), _p(den), _p(vel)) != 0:
        raise StokesError(f"erosion_matvec: {last_error()}")
    return vel


PRECOND_NONE = 0
PRECOND_CURVE_FOURIER = 1


def set_preconditioner(kind: int, body_modes: int = 32, wall_modes: int = 16) -> None:
    """Select the right preconditioner of stokessolver (0 none, as the reference;
    1 per-curve two-level Fourier).  Only the iteration count changes."""
    if lib().erosion_set_preconditioner(int(kind), int(body_modes), int(wall_modes)) != 0:
        raise StokesError(f"invalid preconditioner kind={kind} body_modes={body_modes} "
                          f"wall_modes={wall_modes}")


def set_farfield_shape(index: int) -> None:
    """Fix the far-field launch shape (-1: the autotuner chooses)."""
    if lib().erosion_set_farfield_shape(int(index)) != 0:
        raise StokesError(f"far-field shape {index} out of range")


def farfield_shape_count() -> int:
    return int(lib().erosion_farfield_shape_count())


def set_fmm_one_stream(on: bool) -> None:
    """Run the cluster far field's expansions on the library stream ahead of
    its direct sums (a measurement hook for hardware-counter passes; results
    are unchanged)."""
    L = lib()
    L.erosion_set_fmm_one_stream.argtypes = [ctypes.c_int]
    L.erosion_set_fmm_one_stream.restype = ctypes.c_int
    if L.erosion_set_fmm_one_stream(1 if on else 0) != 0:
        raise StokesError(f"erosion_set_fmm_one_stream: {last_error()}")


def set_graphs(on: bool) -> None:
    """Capture GMRES iteration chunks as HIP graphs (the second time a chunk
    comes under the same sizes and buffers) and replay them on later solves
    (results identical to direct launches)."""
    if lib().erosion_set_graphs(1 if on else 0) != 0:
        raise StokesError(f"erosion_set_graphs: {last_error()}")


def graph_stats() -> dict:
    out = (ctypes.c_longlong * 4)()
    if lib().erosion_graph_stats(out) != 0:
        raise StokesError(f"erosion_graph_stats: {last_error()}")
    return {"on": bool(out[0]), "captures": int(out[1]), "replays": int(out[2]),
            "kept": int(out[3])}


def get_preconditioner() -> Tuple[int, int, int]:
    out = (ctypes.c_int * 3)()
    lib().erosion_get_preconditioner(out)
    return int(out[0]), int(out[1]), int(out[2])


def precond_apply(npts: int, nbods: int, nouter: int, xx, yy, r) -> np.ndarray:
    """z = P^{-1} r for the selected preconditioner on this geometry (test hook)."""
    xx, yy, r = _f64(xx), _f64(yy), _f64(r)
    _lengths(npts, nbods, nouter, xx, yy, r)
    z = np.zeros(nunk(npts, nbods, nouter))
    if lib().erosion_precond_apply(npts, nbods, nouter, _p(xx), _p(yy), _p(r), _p(z)) != 0:
        raise StokesError(f"erosion_precond_apply: {last_error()}")
    return z


PRECISION_FP64 = 0
PRECISION_MIXED = 1
PRECISION_FP32_FARFIELD = 2


def set_precision(mode: int) -> None:
    """Far-field precision of stokessolver: 0 FP64 (as the reference), 1 mixed
    (FP32 far field once the residual is small, FP64 true residual and
    restart), 2 FP32 far field everywhere (tests and measurement)."""
    if lib().erosion_set_precision(int(mode)) != 0:
        raise StokesError(f"invalid precision mode {mode}")


def get_precision() -> int:
    return int(lib().erosion_get_precision())


def set_mixed_switch(factor: float) -> None:
    """Mixed mode: FP32 far field once the implicit residual <= factor * tol."""
    if lib().erosion_set_mixed_switch(float(factor)) != 0:
        raise StokesError(f"invalid mixed switch factor {factor}")


def solve_batch(geoms, ibary: int = 1, ibc: int = 0, maxl: int = 8000, workers: int = 4,
                stress: bool = False, guesses=None):
    """Solve many independent geometries in one call (erosion_solve_batch).

    ``geoms`` is a sequence of (npts, nbods, nouter, xx, yy); returns a list of
    dicts with den, nits, status and (with ``stress``) tau per geometry."""
    k = len(geoms)
    n = np.array([g[0] for g in geoms], dtype=np.int64)
    m = np.array([g[1] for g in geoms], dtype=np.int64)
    no = np.array([g[2] for g in geoms], dtype=np.int64)
    xs = [_f64(g[3]) for g in geoms]
    ys = [_f64(g[4]) for g in geoms]
    for i in range(k):
        _lengths(int(n[i]), int(m[i]), int(no[i]), xs[i], ys[i])
    dens = [np.zeros(nunk(int(n[i]), int(m[i]), int(no[i]))) if guesses is None or guesses[i] is None
            else _f64(guesses[i]).copy() for i in range(k)]
    taus = [np.zeros(int(n[i] * m[i])) for i in range(k)] if stress else None
    P = ctypes.POINTER(ctypes.c_double) * k
    px = P(*[_p(a) for a in xs])
    py = P(*[_p(a) for a in ys])
    pd = P(*[_p(a) for a in dens])
    pt = P(*[_p(a) for a in taus]) if stress else None
    nits = np.zeros(k, dtype=np.int64)
    status = np.zeros(k, dtype=np.int32)
    relres = np.zeros(k)
    L = lib()
    L.erosion_solve_batch.restype = ctypes.c_int
    rc = L.erosion_solve_batch(ctypes.c_int64(k), n.ctypes.data_as(_i64p), m.ctypes.data_as(_i64p),
                               no.ctypes.data_as(_i64p), px, py, pd, pt, ctypes.c_int64(ibary),
                               ctypes.c_int64(ibc), ctypes.c_int64(maxl), ctypes.c_int64(workers),
                               nits.ctypes.data_as(_i64p),
                               status.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), _p(relres))
    out = [dict(den=dens[i], nits=int(nits[i]), status=int(status[i]),
                tau=taus[i] if stress else None, true_relres=float(relres[i]))
           for i in range(k)]
    if rc != 0 and all(o["status"] != 0 for o in out):
        raise StokesError(f"erosion_solve_batch: {last_error()}")
    return out


RK2_INFO = ("iters_stage1", "iters_stage2", "deleted", "umax_stage1", "umax_stage2",
            "true_relres_max", "positive_mterm", "flux_warnings", "warm_starts",
            "given_guesses")


def rungekutta2(npts: int, nouter: int, params, theta, length, xsm, ysm, guess=None):
    """One erosion step on the device (erosion_rungekutta2).

    ``params`` = (dt, epsilon, sigma, fixarea, fixpdrop, ibc, ibary, maxl, warm);
    theta is (nbods, npts); ``guess`` an optional first-stage GMRES guess used
    when no kept density fits.  Returns (theta, len, xsm, ysm, info) of the
    bodies that remain; the inputs are not modified."""
    theta = np.array(theta, dtype=np.float64, order="C", copy=True).reshape(-1)
    m = len(theta) // npts if npts else 0
    if len(theta) != m * npts:
        raise StokesError("theta must hold npts values per body")
    ln, xs, ys = (np.array(a, dtype=np.float64, copy=True) for a in (length, xsm, ysm))
    if not (len(ln) == len(xs) == len(ys) == m):
        raise StokesError("len, xsm, ysm need one value per body")
    prm = _f64(params)
    if len(prm) != 9:
        raise StokesError("params holds 9 values")
    g = None
    if guess is not None:
        g = _f64(guess)
        if len(g) != nunk(npts, m, nouter):
            raise StokesError(f"guess holds {len(g)} values; need {nunk(npts, m, nouter)}")
    info = np.zeros(len(RK2_INFO))
    mo = ctypes.c_int64(0)
    if lib().erosion_rungekutta2(npts, m, nouter, _p(prm), _p(theta), _p(ln), _p(xs), _p(ys),
                                 _p(g) if g is not None else None, ctypes.byref(mo),
                                 _p(info)) != 0:
        raise StokesError(f"erosion_rungekutta2: {last_error()}")
    k = int(mo.value)
    return (theta[:k * npts].reshape(k, npts), ln[:k], xs[:k], ys[:k],
            dict(zip(RK2_INFO, info.tolist())))


def stepper_history() -> dict:
    """Densities the device stepper keeps for guesses: count, body count, length."""
    out = np.zeros(3, dtype=np.int64)
    if lib().erosion_stepper_history(out.ctypes.data_as(_i64p)) != 0:
        raise StokesError(f"erosion_stepper_history: {last_error()}")
    return {"count": int(out[0]), "bodies": int(out[1]), "length": int(out[2])}


def stepper_reset() -> None:
    """Forget the densities the device stepper keeps for its GMRES guesses."""
    if lib().erosion_stepper_reset() != 0:
        raise StokesError(f"erosion_stepper_reset: {last_error()}")


SPECDIFF, SPECINT, EXPSMOOTH, GETXY = 0, 1, 2, 3


def spectral(op: int, rows, fac=None) -> np.ndarray:
    """The device stepper's spectral operations on each row (test hook)."""
    a = np.ascontiguousarray(rows, dtype=np.float64)
    r, n = a.shape
    f = _f64(fac) if fac is not None else np.zeros(1)
    out = np.zeros(2 * r * n if op == GETXY else r * n)
    if lib().erosion_spectral(int(op), n, r, _p(a), _p(f), _p(out)) != 0:
        raise StokesError(f"erosion_spectral: {last_error()}")
    return out.reshape(2, r, n) if op == GETXY else out.reshape(r, n)
