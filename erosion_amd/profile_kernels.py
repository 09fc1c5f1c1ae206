"""Kernel workloads for timing and hardware-counter runs
(python -m erosion_amd.profile_kernels).

Runs, on one workload (default: 100 bodies x 256 points):

* with --fp64-rate, the FP64 FMA rate the chip sustains
  (erosion_measure_fp64);
* device-to-device copies of a known size (erosion_measure_copy), so that a
  profiler's per-dispatch byte counters for the other kernels can be read
  against those of a copy whose traffic is known: 2 x bytes per copy;
* a fixed number of FP64 and FP32 far-field operator applications
  (erosion_time_matvec), reporting the far field's FLOP rate (20 FLOP per
  ordered source-target pair) and, where the geometry has near pairs, the
  barycentric near-pair gather pass against HBM (per near entry the source
  curve's nodes x 14 doubles, plus 16 bytes written);
* one shear-stress evaluation (computeshearstress_) under the library's kernel
  timers, reporting the barycentric stress pass's FLOP rate (130 FLOP per
  target-node pair);
* with --cluster, a fixed number of applications of the cluster far field
  (multipoles, translations and direct sums of well-separated bodies and wall
  segments), with --one-stream its expansions on the library stream ahead of
  the direct sums (for hardware-counter passes, which serialize dispatches).
"""
from __future__ import annotations

import argparse
import ctypes
import sys

import numpy as np

from . import stokes
from .geometry import WORKLOADS, workload_xy

PEAK_FP64_TFLOPS = 78.6
PEAK_FP32_TFLOPS = 157.3
PEAK_HBM_GBS = 8000.0
STRESS_FLOP_PER_PAIR = 130.0


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--workload", default="bodies100_n256", choices=sorted(WORKLOADS))
    ap.add_argument("--reps", type=int, default=5)
    ap.add_argument("--copy-mib", type=int, default=256)
    ap.add_argument("--no-stress", action="store_true")
    ap.add_argument("--shapes", action="store_true",
                    help="time every far-field launch shape (FP64 and FP32)")
    ap.add_argument("--fp64-rate", action="store_true",
                    help="the FP64 FMA rate the chip sustains at 4, 8 and 16 waves per SIMD")
    ap.add_argument("--no-direct", action="store_true",
                    help="skip the all-pairs FP64 and FP32 far field")
    ap.add_argument("--cluster", action="store_true",
                    help="time the cluster far field (FP64)")
    ap.add_argument("--one-stream", action="store_true",
                    help="with --cluster: expansions on the library stream")
    a = ap.parse_args(argv)
    L = stokes.lib()
    if a.fp64_rate:
        L.erosion_measure_fp64.argtypes = [ctypes.c_int64, ctypes.c_int64,
                                           ctypes.POINTER(ctypes.c_double)]
        L.erosion_measure_fp64.restype = ctypes.c_int
        iters = 20000
        for blocks in (1024, 2048, 4096):  # 4, 8, 16 waves of 256 CUs x 4 SIMDs
            ms = ctypes.c_double(0.0)
            if L.erosion_measure_fp64(blocks, iters, ctypes.byref(ms)) != 0:
                raise stokes.StokesError(stokes.last_error())
            fma = 8.0 * iters * blocks * 256
            rate = fma / (ms.value * 1e-3) / 1e12
            print(f"FP64 FMA chains, {blocks * 4 // 1024} waves per SIMD: {ms.value:.3f} ms, "
                  f"{rate:.2f} T FMA/s = {2 * rate:.1f} TFLOP/s ({2 * rate / PEAK_FP64_TFLOPS:.3f} "
                  f"of {PEAK_FP64_TFLOPS})", flush=True)
    if a.copy_mib > 0:
        L.erosion_measure_copy.argtypes = [ctypes.c_int64, ctypes.c_int64,
                                           ctypes.POINTER(ctypes.c_double)]
        L.erosion_measure_copy.restype = ctypes.c_int
        ms = ctypes.c_double(0.0)
        nbytes = a.copy_mib << 20
        if L.erosion_measure_copy(nbytes, a.reps, ctypes.byref(ms)) != 0:
            raise stokes.StokesError(stokes.last_error())
        print(f"copy {a.copy_mib} MiB: {ms.value:.3f} ms, {2 * nbytes / ms.value / 1e6:.0f} GB/s",
              flush=True)
    n, m, no, xx, yy = workload_xy(WORKLOADS[a.workload])
    print(f"workload {a.workload}: {m} bodies x {n} points, wall {no}", flush=True)
    eta = np.random.default_rng(7).standard_normal(2 * no + 2 * n * m + 3 * m)
    direct_modes = () if a.no_direct else (
        (stokes.PRECISION_FP64, "FP64", PEAK_FP64_TFLOPS),
        (stokes.PRECISION_FP32_FARFIELD, "FP32", PEAK_FP32_TFLOPS))
    for mode, name, peak in direct_modes:
        stokes.set_precision(mode)
        r = stokes.time_matvec(n, m, no, xx, yy, eta, reps=a.reps)
        tf = 20.0 * r["pairs"] / (r["farfield_ms"] * 1e-3) / 1e12
        print(f"{name} far field: {r['farfield_ms']:.4f} ms per launch, {tf:.2f} TFLOP/s "
              f"({tf / peak:.3f} of {peak}), operator {r['matvec_ms']:.4f} ms", flush=True)
        if r["near_entries"] > 0 and mode == stokes.PRECISION_FP64:
            gbs = r["gather_bytes"] / (r["near_ms"] * 1e-3) / 1e9
            print(f"near-pair gather pass: {r['near_entries']} entries, "
                  f"{r['gather_bytes'] / 1e6:.2f} MB, {r['near_ms']:.4f} ms, {gbs:.0f} GB/s "
                  f"({gbs / PEAK_HBM_GBS:.3f} of HBM)", flush=True)
    if a.shapes:
        for mode, name in ((stokes.PRECISION_FP64, "FP64"), (stokes.PRECISION_FP32_FARFIELD, "FP32")):
            stokes.set_precision(mode)
            for k in range(stokes.farfield_shape_count()):
                stokes.set_farfield_shape(k)
                r = stokes.time_matvec(n, m, no, xx, yy, eta, reps=a.reps)
                tf = 20.0 * r["pairs"] / (r["farfield_ms"] * 1e-3) / 1e12
                print(f"{name} shape {k}: {r['farfield_ms']:.4f} ms per launch, {tf:.2f} TFLOP/s, "
                      f"{r['splits']} splits", flush=True)
        stokes.set_farfield_shape(-1)
    stokes.set_precision(stokes.PRECISION_FP64)
    if a.cluster:
        stokes.set_fmm_one_stream(a.one_stream)
        stokes.set_farfield_shape(stokes.farfield_shape_count() - 1)
        try:
            r = stokes.time_matvec(n, m, no, xx, yy, eta, reps=a.reps)
        finally:
            stokes.set_farfield_shape(-1)
            stokes.set_fmm_one_stream(False)
        print(f"cluster far field ({'one stream' if a.one_stream else 'two streams'}): "
              f"{r['farfield_ms']:.4f} ms per application, operator {r['matvec_ms']:.4f} ms",
              flush=True)
    if not a.no_stress:
        stokes.set_gmres_tol(1e-8)
        stokes.set_preconditioner(stokes.PRECOND_CURVE_FOURIER)
        try:
            den, _ = stokes.stokessolver(n, m, no, 1, 1, 0, 4000, xx, yy)
        finally:
            stokes.set_preconditioner(stokes.PRECOND_NONE)
        stokes.computeshearstress(n, m, no, xx, yy, den, 1)  # set-up outside the timed call
        stokes.kernel_timing(True)
        stokes.computeshearstress(n, m, no, xx, yy, den, 1)
        kt = stokes.kernel_timers()
        stokes.kernel_timing(False)
        se = kt["stress_eval"]
        if se["launches"]:
            ms = se["ms"] / se["launches"]
            tf = STRESS_FLOP_PER_PAIR * se["work"] / se["launches"] / (ms * 1e-3) / 1e12
            print(f"stress barycentric pass: {ms:.3f} ms per launch, {se['work'] / se['launches']:.3e} "
                  f"target-node pairs, {tf:.2f} TFLOP/s ({tf / PEAK_FP64_TFLOPS:.3f} of "
                  f"{PEAK_FP64_TFLOPS})", flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
