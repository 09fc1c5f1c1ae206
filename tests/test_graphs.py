"""HIP-graph replay of GMRES iteration chunks (erosion_set_graphs).

A chunk runs directly the first time it comes under a signature, is captured
the second time and replayed after that.  A replayed chunk must give the
densities of direct launches bit for bit: on the solve that captures the
graphs, on later solves that only replay them, and after other work (another
geometry, with its own graphs) has run in between.
"""
import os
import subprocess
import sys
import textwrap

import numpy as np
import pytest

from erosion_amd import stokes
from erosion_amd.geometry import WORKLOADS, make_bodies, workload_xy
from erosion_amd.theta_len import getnxy, new_thlenden

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def _reset():
    yield
    stokes.set_graphs(False)
    stokes.set_preconditioner(stokes.PRECOND_NONE)
    stokes.set_precision(stokes.PRECISION_FP64)
    stokes.set_gmres_tol(1e-8)


def _configure(accel: bool) -> None:
    stokes.set_gmres_tol(1e-10)
    stokes.set_preconditioner(stokes.PRECOND_CURVE_FOURIER if accel else stokes.PRECOND_NONE)
    stokes.set_precision(stokes.PRECISION_MIXED if accel else stokes.PRECISION_FP64)


def _tight():
    thl = make_bodies(30, 128, 0.05, 0.12, 0.002, 31)  # near pairs against bodies and wall
    m, xx, yy = getnxy(new_thlenden(thl))
    return 128, m, 1024, xx, yy


@pytest.mark.parametrize("accel", [False, True])
def test_graph_replay_matches_direct(accel):
    _configure(accel)
    n, m, no, xx, yy = workload_xy(WORKLOADS["bodies100_n256"], nbodies=20)
    tn, tm, tno, txx, tyy = _tight()
    d0, it0 = stokes.stokessolver(n, m, no, 1, 1, 0, 3000, xx, yy)
    e0, jt0 = stokes.stokessolver(tn, tm, tno, 1, 1, 0, 3000, txx, tyy)

    stokes.set_graphs(True)
    s0 = stokes.graph_stats()
    d1, it1 = stokes.stokessolver(n, m, no, 1, 1, 0, 3000, xx, yy)  # first sight: direct
    s1 = stokes.graph_stats()
    assert s1["captures"] == s0["captures"] and s1["replays"] == s0["replays"]
    d2, it2 = stokes.stokessolver(n, m, no, 1, 1, 0, 3000, xx, yy)  # captures
    s2 = stokes.graph_stats()
    assert s2["captures"] > s1["captures"] and s2["replays"] > s1["replays"]
    d3, it3 = stokes.stokessolver(n, m, no, 1, 1, 0, 3000, xx, yy)  # replays only
    s3 = stokes.graph_stats()
    assert s3["captures"] == s2["captures"] and s3["replays"] > s2["replays"]
    # other work in between: another geometry (its own graphs), a stress call
    e1, jt1 = stokes.stokessolver(tn, tm, tno, 1, 1, 0, 3000, txx, tyy)
    e2, jt2 = stokes.stokessolver(tn, tm, tno, 1, 1, 0, 3000, txx, tyy)
    stokes.computeshearstress(tn, tm, tno, txx, tyy, e2, 1)
    d4, it4 = stokes.stokessolver(n, m, no, 1, 1, 0, 3000, xx, yy)
    d5, it5 = stokes.stokessolver(n, m, no, 1, 1, 0, 3000, xx, yy)
    for d, it in ((d1, it1), (d2, it2), (d3, it3), (d4, it4), (d5, it5)):
        assert it == it0
        assert np.array_equal(d, d0)
    assert jt1 == jt0 and np.array_equal(e1, e0)
    assert jt2 == jt0 and np.array_equal(e2, e0)
    assert stokes.last_solve_info()["true_relres"] < (1e-10 if accel else 1e-9)


def test_graphs_off_frees_and_direct_path_unchanged():
    _configure(True)
    n, m, no, xx, yy = workload_xy(WORKLOADS["bodies100_n256"], nbodies=10)
    stokes.set_graphs(True)
    stokes.stokessolver(n, m, no, 1, 1, 0, 3000, xx, yy)  # first sight: direct
    d1, _ = stokes.stokessolver(n, m, no, 1, 1, 0, 3000, xx, yy)  # captures
    assert stokes.graph_stats()["kept"] > 0
    stokes.set_graphs(False)
    assert stokes.graph_stats()["kept"] == 0
    d2, _ = stokes.stokessolver(n, m, no, 1, 1, 0, 3000, xx, yy)
    assert np.array_equal(d1, d2)


def test_graphs_in_batched_solves():
    """Worker contexts capture and replay their own chunks.  Each worker tunes
    its far-field shape separately, so a geometry that lands on another worker
    may sum in another order: the bar is the solve tolerance, not bit equality."""
    _configure(True)
    geos = [workload_xy(WORKLOADS["bodies100_n256"], nbodies=k) for k in (8, 12, 8, 12)]
    ref = stokes.solve_batch(geos, workers=2)
    stokes.set_graphs(True)
    for _ in range(2):
        out = stokes.solve_batch(geos, workers=2)
        for a, b in zip(out, ref):
            assert a["status"] == 0 and abs(a["nits"] - b["nits"]) <= 2
            assert np.max(np.abs(a["den"] - b["den"])) < 1e-9 * np.max(np.abs(b["den"]))


def test_graphs_on_from_the_first_call_of_a_process(tmp_path):
    """Graphs switched on before any other library call, in a fresh process:
    the set-up a graph solve runs first (far-field tuning, which reads the
    gathered density) must not depend on buffers an earlier call allocated.
    Covers the main context and fresh batch workers."""
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    out = tmp_path / "den.npz"
    code = textwrap.dedent(f"""
        import sys
        sys.path.insert(0, {root!r})
        import numpy as np
        from erosion_amd import stokes
        from erosion_amd.geometry import WORKLOADS, workload_xy
        stokes.set_graphs(True)
        stokes.set_gmres_tol(1e-10)
        stokes.set_preconditioner(stokes.PRECOND_CURVE_FOURIER)
        stokes.set_precision(stokes.PRECISION_MIXED)
        n, m, no, xx, yy = workload_xy(WORKLOADS["bodies100_n256"], nbodies=10)
        print("first solves with graphs (direct, then captured)", flush=True)
        for _ in range(2):
            d, it = stokes.stokessolver(n, m, no, 1, 1, 0, 3000, xx, yy)
        print("batches with graphs on fresh workers", flush=True)
        for _ in range(2):
            b = stokes.solve_batch([(n, m, no, xx, yy)] * 2, workers=2)
        np.savez({str(out)!r}, d=d, it=it, b0=b[0]["den"], b1=b[1]["den"],
                 captures=stokes.graph_stats()["captures"])
    """)
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    res = np.load(out)
    assert int(res["captures"]) > 0
    _configure(True)
    n, m, no, xx, yy = workload_xy(WORKLOADS["bodies100_n256"], nbodies=10)
    d0, it0 = stokes.stokessolver(n, m, no, 1, 1, 0, 3000, xx, yy)
    for d in (res["d"], res["b0"], res["b1"]):
        assert np.max(np.abs(d - d0)) < 1e-9 * np.max(np.abs(d0))
    assert abs(int(res["it"]) - it0) <= 2
