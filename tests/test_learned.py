"""Learned warm start: the CPU network oracle, features and targets (CPU),
and the GPU network against the oracle plus the trained guess's gain (GPU)."""
import os

import numpy as np
import pytest

from oracle import mlp as OM

SHAPE = (128, 256, 3, 69)
SMALL = (20, 24, 2, 7)


def _rand_net(shape, seed, scale=1.0):
    rng = np.random.default_rng(seed)
    p = rng.standard_normal(OM.param_count(*shape)) * scale / np.sqrt(shape[1])
    return p.astype(np.float32)


# ---------------------------------------------------------------- CPU
def test_bf16_round_known_values():
    # BF16 keeps 8 significant bits: the spacing just above 1 is 2^-7
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, 1.0 + 2 ** -7 + 2 ** -9, 3.0000001,
                  -2.5, 0.0], np.float32)
    r = OM.bf16_round(x)
    # ties (1 + 2^-8, 1 + 3 2^-8) go to the even neighbour; others to the nearest
    assert list(r) == [1.0, 1.0, 1.0 + 2 ** -6, 1.0 + 2 ** -7, 3.0, -2.5, 0.0]


def test_oracle_backward_matches_finite_differences():
    rng = np.random.default_rng(1)
    p = _rand_net(SMALL, 2).astype(np.float64)
    x = rng.standard_normal((9, SMALL[0]))
    t = rng.standard_normal((9, SMALL[3]))
    _, g = OM.loss_and_gradients(SMALL, p, x, t, bf16=False)
    h = 1e-6
    idx = rng.choice(len(p), 40, replace=False)
    for i in idx:
        e = np.zeros_like(p)
        e[i] = h
        lp, _ = OM.loss_and_gradients(SMALL, p + e, x, t, bf16=False)
        lm, _ = OM.loss_and_gradients(SMALL, p - e, x, t, bf16=False)
        assert abs((lp - lm) / (2 * h) - g[i]) < 1e-7 * max(1.0, abs(g[i]))


def test_planted_oracle_mistake_exceeds_gpu_tolerance():
    """Dropping the tanh derivative changes the gradient far beyond the 2e-3
    the GPU check allows, so that check can see such a mistake."""
    rng = np.random.default_rng(4)
    p = _rand_net(SHAPE, 5)
    x = rng.standard_normal((64, SHAPE[0])).astype(np.float32)
    t = rng.standard_normal((64, SHAPE[3])).astype(np.float32)
    _, g = OM.loss_and_gradients(SHAPE, p, x, t)
    # the oracle's backward pass with the factor 1 - h^2 left out
    hs = OM.forward(SHAPE, p, x)
    gcur = 2.0 * (hs[-1] - t) / hs[-1].size
    gbad = np.zeros_like(g)
    L = OM.layout(*SHAPE)
    for l in range(SHAPE[2], -1, -1):
        ni, no, ow, ob = L[l]
        gbad[ow:ow + no * ni] = (OM.bf16_round(gcur.T) @ OM.bf16_round(hs[l])).ravel()
        gbad[ob:ob + no] = gcur.sum(axis=0)
        if l:
            W = p[ow:ow + no * ni].astype(np.float64).reshape(no, ni)
            gcur = OM.bf16_round(gcur) @ OM.bf16_round(W)
    # the planted mistake only touches the hidden layers; the output layer agrees
    ni, no, ow, ob = L[-1]
    assert np.allclose(gbad[ow:], g[ow:])
    assert np.linalg.norm(gbad - g) / np.linalg.norm(g) > 0.05


def test_targets_roundtrip_band_limited_solution():
    from erosion_amd import learned as LR
    n, m, no = 64, 3, 32
    rng = np.random.default_rng(0)
    T = rng.standard_normal((m, LR.NOUT))
    wall = rng.standard_normal(2 * no)
    u = LR.assemble_solution(T, n, m, no, wall)
    assert np.abs(LR.solution_targets(u, n, m, no) - T).max() < 1e-10
    assert np.array_equal(u[:2 * no], wall)
    # body b's x part is its Fourier series: mode-0 coefficient / sqrt-free mean
    o = 2 * no
    assert abs(np.mean(u[o:o + n]) - T[0, 0]) < 1e-12


def test_features_shape_and_neighbours():
    from erosion_amd import learned as LR
    from erosion_amd.geometry import make_bodies
    from erosion_amd.theta_len import new_thlenden
    th = new_thlenden(make_bodies(12, 64, 0.05, 0.08, 0.05, 3))
    F = LR.body_features(th)
    assert F.shape == (12, LR.NIN)
    assert np.all(F[:, 125:] == 0.0)
    # first neighbour's relative centre is the nearest other centre
    ctr = np.array([[t.xsm, t.ysm] for t in th.thlenvec])
    d = np.linalg.norm(ctr[:, None] - ctr[None], axis=2)
    np.fill_diagonal(d, np.inf)
    base = 2 * LR.NMODES + 5
    for b in range(12):
        assert np.allclose(F[b, base:base + 2], ctr[np.argmin(d[b])] - ctr[b])


# ---------------------------------------------------------------- GPU
gpu = pytest.mark.gpu


@gpu
def test_gpu_forward_matches_oracle():
    from erosion_amd import learned as LR
    rng = np.random.default_rng(8)
    p = _rand_net(SHAPE, 9)
    x = rng.standard_normal((300, SHAPE[0])).astype(np.float32)
    assert LR.param_count(SHAPE) == OM.param_count(*SHAPE)
    y = LR.mlp_forward(p, x, SHAPE)
    yo = OM.forward(SHAPE, p, x)[-1]
    err = np.linalg.norm(y - yo) / np.linalg.norm(yo)
    print(f"forward relative error vs oracle {err:.2e}")
    assert err < 2e-3


@gpu
def test_gpu_gradients_match_oracle():
    from erosion_amd import learned as LR
    rng = np.random.default_rng(10)
    p = _rand_net(SHAPE, 11)
    x = rng.standard_normal((200, SHAPE[0])).astype(np.float32)
    t = rng.standard_normal((200, SHAPE[3])).astype(np.float32)
    loss, g = LR.mlp_gradients(p, x, t, SHAPE)
    lo, go = OM.loss_and_gradients(SHAPE, p, x, t)
    err = np.linalg.norm(g - go) / np.linalg.norm(go)
    print(f"loss {loss:.6f} vs {lo:.6f}; gradient relative error {err:.2e}")
    assert abs(loss - lo) < 1e-3 * lo
    assert err < 2e-3
    # per layer, so that a small layer cannot hide behind a large one
    for ni, no, ow, ob in OM.layout(*SHAPE):
        gw, gwo = g[ow:ow + ni * no], go[ow:ow + ni * no]
        assert np.linalg.norm(gw - gwo) / np.linalg.norm(gwo) < 5e-3
        assert np.linalg.norm(g[ob:ob + no] - go[ob:ob + no]) / np.linalg.norm(go[ob:ob + no]) < 5e-3


@gpu
def test_gpu_training_fits_a_smooth_map():
    from erosion_amd import learned as LR
    rng = np.random.default_rng(12)
    shape = (16, 64, 3, 4)
    x = rng.standard_normal((4096, 16)).astype(np.float32)
    A = rng.standard_normal((16, 4)) / 4
    t = np.tanh(x @ A).astype(np.float32)
    p0 = LR.init_params(shape, seed=1)
    p, hist = LR.mlp_train(p0, x, t, epochs=30, batch=256, lr=3e-3, seed=2, shape=shape)
    print(f"loss {hist[0]:.4f} -> {hist[-1]:.5f}")
    assert hist[-1] < 0.05 * hist[0]


@gpu
@pytest.mark.skipif(not os.path.exists(
    os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                 "erosion_amd", "data", "warmstart_mlp.npz")), reason="weights not trained")
def test_learned_guess_never_above_zeros_held_out():
    """Held-out 100-body geometry (seed outside the training runs),
    preconditioned GMRES at 1e-10 in the preconditioned variable: the
    network's guess, scaled by the factor that minimises its residual, starts
    no higher than zeros and costs at most two iterations more.  (The
    unscaled guess starts far above ||b||: each body's rigid-motion part of u
    would have to be right to about 1e-3, which the network is not.)  The
    test records what the guess does; it does not claim a gain."""
    from conftest import relerr
    from erosion_amd import learned as LR
    from erosion_amd import stokes
    from erosion_amd.geometry import make_bodies
    from erosion_amd.theta_len import getnxy, new_thlenden
    n, no = 256, 2048
    g = LR.LearnedGuess()
    th = new_thlenden(make_bodies(100, n, 0.02, 0.05, 0.02, 2300))
    m, xx, yy = getnxy(th)
    stokes.set_gmres_tol(1e-10)
    stokes.set_preconditioner(stokes.PRECOND_CURVE_FOURIER)
    stokes.set_guess_space(stokes.GUESS_PRECONDITIONED)
    try:
        d0, it0 = stokes.stokessolver(n, m, no, 1, 1, 0, 8000, xx, yy)
        raw = g.predict(th, no, scale=False)
        _, itr = stokes.stokessolver(n, m, no, 1, 1, 0, 8000, xx, yy, den=raw)
        r_raw = stokes.last_solve_info()["r0_relres"]
        d1, it1 = stokes.stokessolver(n, m, no, 1, 1, 0, 8000, xx, yy, den=g.predict(th, no))
        info = stokes.last_solve_info()
        print(f"learned guess: raw starting residual {r_raw:.3e} ({itr} iterations), scaled "
              f"{info['r0_relres']:.3e} ({it1}); zeros {it0}")
        assert stokes.last_status() == 0 and info["true_relres"] <= 1e-10
        assert relerr(d1, d0) < 1e-7
        assert info["r0_relres"] <= 1.0 and it1 <= it0 + 2
    finally:
        stokes.set_guess_space(stokes.GUESS_DENSITY)
        stokes.set_preconditioner(stokes.PRECOND_NONE)
        stokes.set_gmres_tol(1e-8)
