"""CPU pins of the oracle's warm-start surrogate definitions (features,
targets, MLP loss and gradient).  The MLP has no counterpart in the
reference; these checks pin it to what the mathematics fixes: finite
differences of the loss, closed forms on uniform states, scale invariance
of the Euler equations, and the Picard fixed point."""
import numpy as np
import pytest

import pyoracle as O
from layouts import single_level


def _rand_params(dims, rng, scale=0.3):
    nin_p, nh, nout_p, nout = dims
    n = nh * nin_p + nh + nh * nh + nh + nout_p * nh + nout_p
    return (rng.standard_normal(n) * scale).astype(np.float32)


def test_loss_gradient_matches_finite_differences():
    rng = np.random.default_rng(3)
    dims = (16, 128, 16, 13)
    par = _rand_params(dims, rng, 0.1)
    x = rng.standard_normal((4, 16)).astype(np.float32)
    y = rng.standard_normal((4, 16)).astype(np.float32)
    loss, g, _ = O.mlp_loss_grad(dims, par, x, y, bf16=False)
    # float32 parameters: central differences in float64 through a float64
    # copy is not possible via the float ABI, so perturb representable steps
    idx = rng.choice(par.size, 40, replace=False)
    for i in idx:
        h = np.float32(2.0 ** -10) * max(np.float32(1.0), abs(par[i]))
        pp, pm = par.copy(), par.copy()
        pp[i] += h
        pm[i] -= h
        dp = float(pp[i]) - float(pm[i])
        lp = O.mlp_loss_grad(dims, pp, x, y, bf16=False, want_grad=False)[0]
        lm = O.mlp_loss_grad(dims, pm, x, y, bf16=False, want_grad=False)[0]
        fd = (lp - lm) / dp
        assert abs(fd - g[i]) <= 1e-5 * max(1.0, abs(g[i])) + 1e-7, (i, fd, g[i])


def _bf16(v):
    u = np.array([v], dtype=np.float32).view(np.uint32)[0]
    u = (int(u) + 0x7FFF + ((int(u) >> 16) & 1)) & 0xFFFF0000
    return float(np.array([u], dtype=np.uint32).view(np.float32)[0])


def test_bf16_rounding_is_round_to_nearest_even():
    """With W1 = W2 = W3 = 0, b2 = beta, y = -t the only nonzero weight
    gradient is dW3[0][k] = bf16(2t) * bf16(gelu(beta)); t = (1 + 2^-8)/2
    puts bf16(2t) on a tie that round-to-nearest-even sends to 1.0."""
    from math import erf, sqrt
    dims = (16, 128, 16, 1)
    nin_p, nh, nout_p, nout = dims
    par = np.zeros(nh * nin_p + nh + nh * nh + nh + nout_p * nh + nout_p, dtype=np.float32)
    beta = 0.8125
    b2_off = nh * nin_p + nh + nh * nh
    par[b2_off:b2_off + nh] = beta
    x = np.zeros((1, 16), dtype=np.float32)
    gl = 0.5 * beta * (1.0 + erf(beta / sqrt(2.0)))
    for t, g_bf in (((1 + 2 ** -8) / 2, 1.0), ((1 + 3 * 2 ** -8) / 2, 1 + 2 ** -6), ((1 + 2 ** -9) / 2, 1.0)):
        y = np.zeros((1, 16), dtype=np.float32)
        y[0, 0] = -t
        _, g, _ = O.mlp_loss_grad(dims, par, x, y, bf16=True)
        assert _bf16(2 * t) == g_bf
        w3_off = b2_off + nh
        assert g[w3_off] == g_bf * _bf16(gl)
        _, g0, _ = O.mlp_loss_grad(dims, par, x, y, bf16=False)
        assert np.isclose(g0[w3_off], np.float32(2 * t) * gl, rtol=1e-7)


def _elem(Uw, s):
    return Uw[:, :, s]


@pytest.fixture(scope="module")
def vortex():
    T = O.Tables(3, 2)
    levels = single_level(8)
    H = O.Hier(2, levels, [8, 8], [0, 0], [10, 10])
    Uw = O.project_ic(T, H, 0, [5.0, 5.0, 5.0, 1.0, 0.0])
    return T, H, Uw


def test_uniform_state_closed_forms():
    T = O.Tables(3, 2)
    H = O.Hier(2, single_level(4), [4, 4], [0, 0], [1, 1])
    rho, u, v, p, g = 1.3, 0.7, -0.2, 2.0, 1.4
    Uw = O.project_ic(T, H, 2, [rho, u, v, p], g)
    U = _elem(Uw[0], 7)
    tau = np.array([0.05, 0.03])
    x = O.mlp_features(T, U, tau, g)
    c = np.sqrt(g * p / rho)
    assert np.isclose(x[T.Q * T.Np_s], tau[0] * (abs(u) + c) + tau[1] * (abs(v) + c), rtol=1e-13)
    assert np.allclose(x[T.Q * T.Np_s + 1:], tau)
    assert np.isclose(x[0], 1.0) and np.isclose(x[T.Np_s], u, rtol=1e-13)
    # a uniform state is a steady solution: the converged predictor equals
    # the time-constant extension A0 U, so the target vanishes
    Hs = O.predictor_iters(T, H, 0, Uw[0], 0.01, 12, g)
    y = O.mlp_target(T, U, Hs[:, :, 7], tau)
    assert np.abs(y).max() < 1e-10


def test_features_and_targets_are_scale_invariant(vortex):
    T, H, Uw = vortex
    Dt = 0.05
    tau = np.array([Dt / 1.25, Dt / 1.25])
    Hs = O.predictor_iters(T, H, 0, Uw[0], Dt, 12)
    Hs2 = O.predictor_iters(T, H, 0, 3.0 * Uw[0], Dt, 12)
    for s in (44, 55, 54):   # cells (3,3), (4,4), (3,4) near the vortex core
        U = _elem(Uw[0], s)
        x1, x2 = O.mlp_features(T, U, tau), O.mlp_features(T, 3.0 * U, tau)
        assert np.allclose(x1, x2, rtol=1e-12, atol=1e-14)
        y1 = O.mlp_target(T, U, Hs[:, :, s], tau)
        y2 = O.mlp_target(T, 3.0 * U, Hs2[:, :, s], tau)
        assert np.allclose(y1, y2, rtol=1e-9, atol=1e-11)
        assert 1e-3 < np.abs(y1).max() < 1e3


def test_converged_predictor_is_a_fixed_point(vortex):
    T, H, Uw = vortex
    Dt = 0.05
    a = O.predictor_iters(T, H, 0, Uw[0], Dt, 30)
    b = O.predictor_iters(T, H, 0, Uw[0], Dt, 31)
    assert np.abs(a - b).max() < 1e-13 * np.abs(a).max()
    # and P iterations is the reference predictor
    assert np.array_equal(O.predictor_iters(T, H, 0, Uw[0], Dt, T.P), O.predictor(T, H, 0, Uw[0], Dt))
