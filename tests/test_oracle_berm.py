"""Pins of the Bermudan / continuation-network oracle (CPU only): the network gradient against FP64
central differences, BF16 rounding against hand-worked ties, one Adam step against its closed form,
the bond-option closed form against Monte Carlo, and the LSM chain on a single-exercise contract
against that closed form."""
import math

import numpy as np
import pytest

import oracle
from oracle import mlp, swaption

MODEL = (0.10, 0.03, 0.01, 0.02)


def _theta(seed):
    rng = np.random.default_rng(seed)
    th = np.zeros(mlp.N_PARAMS)
    th[mlp.OFF_W1:mlp.OFF_W2] = rng.normal(0, 0.6, 256)
    th[mlp.OFF_W2:mlp.OFF_B2] = rng.normal(0, 1 / 8, 4096)
    th[mlp.OFF_B2:mlp.OFF_W3] = rng.normal(0, 0.1, 64)
    th[mlp.OFF_W3:mlp.OFF_B3] = rng.normal(0, 1 / 8, 4096)
    th[mlp.OFF_B3:mlp.OFF_W4] = rng.normal(0, 0.1, 64)
    th[mlp.OFF_W4:mlp.OFF_B4] = rng.normal(0, 1 / 8, 64)
    th[mlp.OFF_B4] = 0.05
    return th


def test_bf16_rounding_ties_to_even():
    assert mlp.bf16(1.0) == 1.0
    assert mlp.bf16(1.0 + 2.0 ** -8) == 1.0                       # tie -> even (1.0)
    assert mlp.bf16(1.0 + 3 * 2.0 ** -8) == 1.0 + 2.0 ** -6       # tie -> even (mantissa ...10)
    assert mlp.bf16(1.0 + 2.0 ** -8 + 2.0 ** -20) == 1.0 + 2.0 ** -7
    assert mlp.bf16(-3.0) == -3.0
    x = np.random.default_rng(1).normal(size=1000)
    assert np.all(np.abs(mlp.bf16(x) - x) <= np.abs(x) * 2.0 ** -8 * (1 + 1e-12))


def test_network_gradient_matches_central_differences():
    th = _theta(3)
    rng = np.random.default_rng(4)
    n = 24
    u = rng.uniform(-1, 1, (n, 3))
    t = rng.normal(0.3, 0.5, n)
    mask = (rng.uniform(size=n) < 0.8).astype(float)
    loss, g = mlp.loss_and_grad(th, u, t, mask, n, emulate_bf16=False)
    h = 1e-6
    fd = np.empty_like(g)
    for i in range(mlp.N_PARAMS):
        e = np.zeros_like(th)
        e[i] = h
        fd[i] = (mlp.loss_and_grad(th + e, u, t, mask, n, False)[0]
                 - mlp.loss_and_grad(th - e, u, t, mask, n, False)[0]) / (2 * h)
    # every block of the parameter vector is covered (a dropped or transposed term fails here)
    for lo, hi in ((0, 256), (256, 4352), (4352, 4416), (4416, 8512), (8512, 8576), (8576, 8640), (8640, 8641)):
        err = np.linalg.norm(g[lo:hi] - fd[lo:hi]) / max(np.linalg.norm(fd[lo:hi]), 1e-300)
        assert err < 1e-6, (lo, hi, err)


def test_bf16_emulated_gradient_is_close_to_exact():
    th = _theta(5)
    rng = np.random.default_rng(6)
    u = rng.uniform(-1, 1, (256, 3))
    t = rng.normal(0.3, 0.5, 256)
    mask = np.ones(256)
    _, g = mlp.loss_and_grad(th, u, t, mask, 256, emulate_bf16=False)
    _, gb = mlp.loss_and_grad(th, u, t, mask, 256, emulate_bf16=True)
    assert np.linalg.norm(gb - g) / np.linalg.norm(g) < 2e-2


def test_adam_first_step_closed_form():
    th = np.array([0.5, -0.25, 1.0], dtype=np.float32)
    g = np.array([2.0, -1e-3, 0.0], dtype=np.float32)
    th1, m, v = mlp.adam(th, np.zeros(3), np.zeros(3), g, 1, 1e-3)
    # first step: m_hat = g, v_hat = g^2, so the update is lr * g / (|g| + eps)
    expect = th - 1e-3 * g / (np.abs(g) + 1e-8)
    assert np.allclose(th1, expect, rtol=1e-6, atol=1e-9)
    assert np.allclose(m, 0.1 * g) and np.allclose(v, 0.001 * g * g)


def test_bond_put_closed_form_matches_monte_carlo():
    T, S, X = 2.0, 3.0, 0.975
    spy = 64
    z = oracle.philox_normals(11, 200_000, int(T * spy))
    r, lb = oracle.berm_simulate(MODEL, 1.0 / spy, z)
    pay = np.maximum(X - swaption.zcb(MODEL, r[-1], T, S), 0.0) * np.exp(-lb[-1])
    mc, se = pay.mean(), pay.std() / math.sqrt(len(pay))
    cf = float(swaption.zbp(MODEL, MODEL[3], 0.0, T, S, X))
    assert abs(mc - cf) < 4 * se + 1e-4 * cf, (mc, cf, se)
    # the discounting alone: E[exp(-int r)] = P(0, T)
    d = np.exp(-lb[-1])
    assert abs(d.mean() - swaption.zcb(MODEL, MODEL[3], 0.0, T)) < 4 * d.std() / math.sqrt(len(d))


def test_lsm_single_exercise_price_and_exposure_match_closed_form():
    spy, E, K, N = 4, 2, 0.031, 1e6
    z = oracle.philox_normals(12, 100_000, E * spy + spy)
    r, lb = oracle.berm_simulate(MODEL, 1.0 / spy, z)
    book = dict(notional=[N], strike=[K], first_ex=[E], n_ex=[1])
    pol = oracle.berm_policy(MODEL, book, spy, r, lb)
    cf0 = float(swaption.payer_one_period(MODEL, MODEL[3], 0.0, E, N, K))
    se = pol["y0"][0].std() / math.sqrt(r.shape[1])
    assert abs(pol["price"][0] - cf0) < 4 * se, (pol["price"][0], cf0, se)
    # exposure at an interior date: the cubic regression against the closed-form value on the same paths
    step = 5
    _, n_alive, ee = oracle.berm_judge(MODEL, spy, r, lb, pol["tau"][0], pol["y0"][0], step)
    assert n_alive == r.shape[1]
    truth = float(np.mean(swaption.payer_one_period(MODEL, r[step], step / spy, E, N, K)))
    assert abs(ee - truth) < 0.02 * truth, (ee, truth)
