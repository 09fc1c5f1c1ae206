"""Mixed precision: FP32 far field inside FP64-refined GMRES (GPU).

The FP32 far field is checked against the FP64 operator, and mixed solves
against FP64 solves of the same geometry: the FP64 true residual meets the
tolerance and the density and shear stress agree to what the tolerance allows
(shear stress within 1e-8 relative).
"""
import numpy as np
import pytest

from conftest import load_golden, relerr

pytestmark = pytest.mark.gpu


@pytest.fixture
def st():
    from erosion_amd import stokes
    stokes.set_gmres_tol(1e-8)
    yield stokes
    stokes.set_precision(0)
    stokes.set_preconditioner(0)
    stokes.set_gmres_tol(1e-8)


def _full():
    from erosion_amd.geometry import WORKLOADS, workload_xy
    return workload_xy(WORKLOADS["bodies100_n256"])


@pytest.mark.parametrize("case", ["tight8_n64", "bodies10_n128", "bodies100_n256"])
def test_fp32_farfield_operator_accuracy(st, case):
    if case == "bodies100_n256":
        n, m, no, xx, yy = _full()
    else:
        g = load_golden(case)
        n, m, no, xx, yy = int(g["n"]), int(g["m"]), int(g["no"]), g["xx"], g["yy"]
    N = 2 * no + 2 * n * m + 3 * m
    eta = np.random.default_rng(21).standard_normal(N)
    st.set_precision(0)
    v64 = st.matvec(n, m, no, xx, yy, eta)
    st.set_precision(2)
    v32 = st.matvec(n, m, no, xx, yy, eta)
    err = relerr(v32, v64)
    rms = np.sqrt(np.mean((v32 - v64) ** 2) / np.mean(v64 ** 2))
    print(f"{case}: FP32 far field max rel err {err:.2e}, rms {rms:.2e}")
    assert err < 1e-5 and rms < 1e-6


def test_fp32_farfield_every_shape(st):
    """Each far-field launch shape in FP32 (shapes asking for 8 targets per
    lane run the FP32 kernel's 4 and must still cover every target), on 20
    bodies x 256 points."""
    from erosion_amd.geometry import WORKLOADS, workload_xy
    n, m, no, xx, yy = workload_xy(WORKLOADS["bodies100_n256"], nbodies=20)
    N = 2 * no + 2 * n * m + 3 * m
    eta = np.random.default_rng(12).standard_normal(N)
    st.set_precision(0)
    v64 = st.matvec(n, m, no, xx, yy, eta)
    st.set_precision(2)
    try:
        for k in range(st.farfield_shape_count()):
            st.set_farfield_shape(k)
            v32 = st.matvec(n, m, no, xx, yy, eta)
            err = relerr(v32, v64)
            rms = np.sqrt(np.mean((v32 - v64) ** 2) / np.mean(v64 ** 2))
            assert err < 1e-6 and rms < 1e-7, (k, err, rms)
    finally:
        st.set_farfield_shape(-1)


@pytest.mark.parametrize("pc", [0, 1])
def test_mixed_solve_bodies10(st, pc):
    """With a direct far-field shape fixed, so that mixed precision uses the FP32
    far field as its residual switch asks (left to the autotuner, the choice at
    this small size rests on launch-overhead-bound timings)."""
    g = load_golden("bodies10_n128")
    n, m, no, xx, yy = int(g["n"]), int(g["m"]), int(g["no"]), g["xx"], g["yy"]
    st.set_gmres_tol(1e-10)
    st.set_preconditioner(pc)
    st.set_precision(0)
    st.set_farfield_shape(0)
    try:
        d64, it64 = st.stokessolver(n, m, no, 1, 1, 0, 8000, xx, yy)
        st.set_precision(1)
        d32, it32 = st.stokessolver(n, m, no, 1, 1, 0, 8000, xx, yy)
    finally:
        st.set_farfield_shape(-1)
    info = st.last_solve_info()
    print(f"pc={pc}: fp64 {it64} iterations, mixed {it32} ({info['fp32_iters']} FP32), "
          f"true relres {info['true_relres']:.2e}")
    assert st.last_status() == 0
    assert info["true_relres"] <= 1e-10
    # unpreconditioned GMRES here falls from a plateau to the tolerance within
    # one 16-iteration chunk, before the residual reaches the FP32 threshold
    assert info["fp32_iters"] > 0 or pc == 0
    assert relerr(d32, g["den_tol12"]) < 1e-8
    st.set_precision(0)
    t64 = st.computeshearstress(n, m, no, xx, yy, d64)
    t32 = st.computeshearstress(n, m, no, xx, yy, d32)
    assert relerr(t32, t64) < 1e-8


def test_mixed_solve_full_size(st):
    """100 bodies x 256 points on the direct far field: there the FP32 far
    field is faster than the FP64 one and mixed precision uses it (the FP64
    cluster far field is faster still, and a mixed solve then keeps FP64)."""
    n, m, no, xx, yy = _full()
    st.set_gmres_tol(1e-10)
    st.set_preconditioner(1)
    st.set_precision(0)
    st.set_farfield_shape(0)
    try:
        d64, it64 = st.stokessolver(n, m, no, 1, 1, 0, 8000, xx, yy)
        st.set_precision(1)
        d32, it32 = st.stokessolver(n, m, no, 1, 1, 0, 8000, xx, yy)
    finally:
        st.set_farfield_shape(-1)
    info = st.last_solve_info()
    print(f"full size: fp64 {it64} iterations, mixed {it32} ({info['fp32_iters']} FP32), "
          f"true relres {info['true_relres']:.2e}")
    assert info["true_relres"] <= 1e-10 and info["fp32_iters"] > 0
    st.set_precision(0)
    t64 = st.computeshearstress(n, m, no, xx, yy, d64)
    t32 = st.computeshearstress(n, m, no, xx, yy, d32)
    print(f"full size: stress rel diff {relerr(t32, t64):.2e}, density {relerr(d32, d64):.2e}")
    assert relerr(t32, t64) < 1e-8


def test_precision_mode_validation(st):
    with pytest.raises(st.StokesError):
        st.set_precision(3)
    st.set_precision(1)
    assert st.get_precision() == 1


_FRESH_PROCESS_FP32 = r"""
import numpy as np
from erosion_amd import stokes
from erosion_amd.geometry import WORKLOADS, workload_xy

def run(nbodies, seed):
    n, m, no, xx, yy = workload_xy(WORKLOADS["bodies100_n256"], nbodies=nbodies)
    eta = np.random.default_rng(seed).standard_normal(2 * no + 2 * n * m + 3 * m)
    return n, m, no, xx, yy, eta

# 1. the process's first operator application is an FP32 one
n, m, no, xx, yy, eta = run(20, 31)
stokes.set_precision(2)
v32 = stokes.matvec(n, m, no, xx, yy, eta)
stokes.set_precision(0)
v64 = stokes.matvec(n, m, no, xx, yy, eta)
# 2. the kernel-timing sequence: FP64, then FP32, at a larger geometry than
#    any FP32 source tiles formed so far
n, m, no, xx, yy, eta = run(40, 32)
stokes.time_matvec(n, m, no, xx, yy, eta, reps=1)
stokes.set_precision(2)
stokes.time_matvec(n, m, no, xx, yy, eta, reps=1)
w32 = stokes.matvec(n, m, no, xx, yy, eta)
stokes.set_precision(0)
w64 = stokes.matvec(n, m, no, xx, yy, eta)
for a, b in ((v32, v64), (w32, w64)):
    err = np.max(np.abs(a - b)) / np.max(np.abs(b))
    rms = np.sqrt(np.mean((a - b) ** 2) / np.mean(b ** 2))
    print("RESULT", err, rms, flush=True)
"""


def test_fp32_farfield_first_in_fresh_process():
    """The FP32 far field's autotuning launches read the FP32 source tiles of
    the current geometry: an FP32 application that comes first in a process,
    and one at a geometry larger than any before it, must form them before
    timing any shape.  Runs in a child process, where no earlier FP32 solve has
    left source tiles behind."""
    import os
    import subprocess
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    env = dict(os.environ, PYTHONPATH=root + os.pathsep + os.environ.get("PYTHONPATH", ""))
    r = subprocess.run([sys.executable, "-c", _FRESH_PROCESS_FP32], cwd=root, env=env,
                       capture_output=True, text=True, timeout=180)
    print(r.stdout[-2000:], r.stderr[-2000:])
    assert r.returncode == 0
    res = [tuple(map(float, ln.split()[1:])) for ln in r.stdout.splitlines()
           if ln.startswith("RESULT")]
    assert len(res) == 2
    for err, rms in res:
        assert err < 1e-5 and rms < 1e-6, (err, rms)
