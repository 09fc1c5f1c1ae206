"""Kernel workloads for hardware-counter runs (python -m erosion_amd.profile_kernels).

Runs, on the 100-body x 256-point workload, a fixed number of FP64 and FP32
far-field operator applications (erosion_time_matvec) and device-to-device
copies of a known size (erosion_measure_copy), so that a profiler's per-
dispatch byte counters for the far-field kernels can be read against those of
a copy whose traffic is known: 2 x bytes per copy (read and write).
"""
from __future__ import annotations

import argparse
import ctypes
import sys

import numpy as np

from . import stokes
from .geometry import WORKLOADS, workload_xy


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--reps", type=int, default=5)
    ap.add_argument("--copy-mib", type=int, default=256)
    a = ap.parse_args(argv)
    L = stokes.lib()
    L.erosion_measure_copy.argtypes = [ctypes.c_int64, ctypes.c_int64,
                                       ctypes.POINTER(ctypes.c_double)]
    L.erosion_measure_copy.restype = ctypes.c_int
    ms = ctypes.c_double(0.0)
    nbytes = a.copy_mib << 20
    if L.erosion_measure_copy(nbytes, a.reps, ctypes.byref(ms)) != 0:
        raise stokes.StokesError(stokes.last_error())
    print(f"copy {a.copy_mib} MiB: {ms.value:.3f} ms, {2 * nbytes / ms.value / 1e6:.0f} GB/s",
          flush=True)
    n, m, no, xx, yy = workload_xy(WORKLOADS["bodies100_n256"])
    eta = np.random.default_rng(7).standard_normal(2 * no + 2 * n * m + 3 * m)
    for mode, name in ((stokes.PRECISION_FP64, "FP64"), (stokes.PRECISION_FP32_FARFIELD, "FP32")):
        stokes.set_precision(mode)
        r = stokes.time_matvec(n, m, no, xx, yy, eta, reps=a.reps)
        print(f"{name} far field: {r['farfield_ms']:.4f} ms per launch, operator "
              f"{r['matvec_ms']:.4f} ms", flush=True)
    stokes.set_precision(stokes.PRECISION_FP64)
    return 0


if __name__ == "__main__":
    sys.exit(main())
