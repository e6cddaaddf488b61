"""Pins of the oracle's G2++ restatement (models/hull_white_2f.rs, exposure.rs:305-480).

Each check is one that a plausible slip (a dropped variance or covariance term, a sign of the
B(a)x / B(b)y terms, the factor 2 of the cross term, swapped shock lists, the grid-0 row count)
fails:
- eta = rho = 0 reduces G2++ to HW1F with long-run mean r0 at r = r0 + x, bond and whole
  portfolio run (the one-factor oracle is pinned by the reference's golden vectors);
- V(tau) against the variance of int_0^tau (x + y) ds by quadrature of the OU covariances;
- the bond at a non-zero state against a conditional Monte Carlo of exp(-int r);
- the exact scheme's moments (variances and covariance of x_T, y_T) against their closed forms;
- the reference's own tests: unit bond at zero maturity, par swap with zero NPV, b = 50 / eta = 0
  within 1e-6 of HW1F (hull_white_2f.rs:223-245), t = 0 profile equal to the deterministic NPV
  (exposure.rs:606-617), zero CVA at zero hazard (exposure.rs:632-645);
- the portfolio run on a singleton equals the reference-shaped profile on the same draws.
"""
import numpy as np
import pytest
from scipy import integrate

import oracle

M2 = (0.1, 0.2, 0.01, 0.012, -0.7, 0.03)  # hull_white_2f.rs tests::reference_model
PAY5, ACC5 = [1.0, 2.0, 3.0, 4.0, 5.0], [1.0] * 5


def test_degenerate_second_factor_is_hull_white_1f():
    a, b, s, _, _, r0 = M2
    for t, T in ((0.0, 0.5), (0.0, 5.0), (1.0, 3.0), (2.0, 30.0)):
        for x in (-0.012, 0.0, 0.021):
            p2 = oracle.zcb2((a, b, s, 0.0, 0.0, r0), x, 0.0, t, T)
            p1 = oracle.zcb(a, r0, s, r0 + x, t, T)
            assert abs(p2 / p1 - 1.0) < 1e-13, (t, T, x)
    # hull_white_2f.rs:223-245: b = 50, eta = rho = 0 against HW1F(a, r0, sigma) at r0
    for T in (1.0, 5.0, 10.0):
        p2 = oracle.zcb2((0.1, 50.0, 0.01, 0.0, 0.0, 0.03), 0.0, 0.0, 0.0, T)
        assert abs(p2 - oracle.zcb(0.1, 0.03, 0.01, 0.03, 0.0, T)) < 1e-6


@pytest.mark.parametrize("rho", [-0.7, 0.5])
def test_variance_term_is_the_variance_of_the_integrated_rate(rho):
    a, b, s, eta = 0.1, 0.2, 0.01, 0.012
    tau = 7.0

    def cov(u, v, k1, k2, s1, s2, r):  # Cov(X_u, Y_v), v <= u, OU processes from 0
        return s1 * s2 * r * np.exp(-k1 * u - k2 * v) * (np.exp((k1 + k2) * v) - 1.0) / (k1 + k2)

    def f(v, u):  # covariance of r_u and r_v on the triangle v <= u (smooth there)
        return (cov(u, v, a, a, s, s, 1.0) + cov(u, v, b, b, eta, eta, 1.0) + cov(u, v, a, b, s, eta, rho)
                + cov(u, v, b, a, eta, s, rho))

    half, _ = integrate.dblquad(f, 0.0, tau, 0.0, lambda u: u, epsabs=1e-16, epsrel=1e-12)
    v_closed = oracle.hw2_var(a, s, tau) + oracle.hw2_var(b, eta, tau) + oracle.hw2_cross(a, b, s, eta, rho, tau)
    assert abs(2.0 * half / v_closed - 1.0) < 1e-9


def test_bond_at_a_state_matches_conditional_monte_carlo():
    a, b, s, eta, rho, r0 = M2
    s, eta = 0.03, 0.025  # larger vols so that the convexity terms are well above the MC error
    m = (a, b, s, eta, rho, r0)
    x0, y0, t, T = 0.01, -0.015, 1.0, 4.0
    n_paths, n_steps = 40000, 300
    dt = (T - t) / n_steps
    z = np.random.default_rng(3).standard_normal((2, n_steps, n_paths))
    xs, ys = oracle.simulate2(m, dt, z, scheme=1)
    r = (xs + x0 * np.exp(-a * dt * np.arange(n_steps + 1))[:, None]) + \
        (ys + y0 * np.exp(-b * dt * np.arange(n_steps + 1))[:, None]) + r0
    integral = dt * (r[0] / 2 + r[1:-1].sum(axis=0) + r[-1] / 2)
    disc = np.exp(-integral)
    mc, se = disc.mean(), disc.std() / np.sqrt(n_paths)
    p = oracle.zcb2(m, x0, y0, t, T)
    assert abs(mc - p) < 4 * se + 1e-6 * p, (mc, p, se)
    # the state terms matter at this size: moving x0 by 0.01 moves P by far more than the error
    assert abs(oracle.zcb2(m, x0 + 0.01, y0, t, T) - p) > 50 * se


def test_exact_scheme_moments():
    a, b, s, eta, rho, r0 = M2
    T, n_steps, n_paths = 3.0, 12, 200000
    z = np.random.default_rng(5).standard_normal((2, n_steps, n_paths))
    xs, ys = oracle.simulate2(M2, T / n_steps, z, scheme=1)
    x, y = xs[-1], ys[-1]
    vx = s * s * (1 - np.exp(-2 * a * T)) / (2 * a)
    vy = eta * eta * (1 - np.exp(-2 * b * T)) / (2 * b)
    cxy = rho * s * eta * (1 - np.exp(-(a + b) * T)) / (a + b)
    se = np.sqrt(2.0 / n_paths)
    assert abs(x.var() / vx - 1) < 4 * se
    assert abs(y.var() / vy - 1) < 4 * se
    assert abs(np.mean(x * y) / np.sqrt(vx * vy) - cxy / np.sqrt(vx * vy)) < 4 / np.sqrt(n_paths)


def test_euler_scheme_correlates_the_second_shock_list():
    """hull_white_2f.rs:130-137: zy = zy_indep * sqrt(1 - rho^2) + zx * rho, one Euler step each."""
    a, b, s, eta, rho, _ = M2
    dt = 0.25
    z = np.random.default_rng(2).standard_normal((2, 1, 5))
    xs, ys = oracle.simulate2(M2, dt, z, scheme=0)
    assert np.array_equal(xs[0], np.zeros(5)) and np.array_equal(ys[0], np.zeros(5))
    np.testing.assert_allclose(xs[1], s * z[0, 0] * np.sqrt(dt), rtol=1e-15)
    zc = z[1, 0] * np.sqrt(1 - rho * rho) + z[0, 0] * rho
    np.testing.assert_allclose(ys[1], eta * zc * np.sqrt(dt), rtol=1e-15)


def test_philox_stream_one_holds_both_factors_of_a_step():
    z = oracle.philox_normals2(7, 3, 4)
    z1 = oracle.philox_normals(7, 3, 8)
    assert z.shape == (2, 4, 3)
    assert not np.any(np.isin(z.ravel(), z1.ravel()))  # a stream of its own
    for s in range(4):
        for p in range(3):
            o = oracle.philox4x32_10([s, p, 0, 1], [7, 0])
            wa, wb = (o[1] << 32) | o[0], (o[3] << 32) | o[2]
            u1, u2 = ((wa >> 11) + 1) * 2.0 ** -53, (wb >> 11) * 2.0 ** -53
            rad, th = np.sqrt(-2 * np.log(u1)), 2 * np.pi * u2
            assert abs(z[0, s, p] - rad * np.cos(th)) < 1e-14 and abs(z[1, s, p] - rad * np.sin(th)) < 1e-14


def test_reference_par_swap_and_unit_bond():
    assert oracle.zcb2(M2, 0.01, -0.005, 1.0, 1.0) == 1.0  # hull_white_2f.rs:205-209
    p = [oracle.zcb2(M2, 0.0, 0.0, 0.0, T) for T in (1.0, 5.0, 10.0)]
    assert p[0] > p[1] > p[2]
    k = oracle.par_rate2(M2, 0.0, PAY5, ACC5)
    assert abs(oracle.irs_npv_reset2(M2, 0.0, 0.0, 0.0, 1e6, k, 0.0, PAY5, ACC5)) < 1e-6  # irs.rs:298-310


def test_profile_at_time_zero_is_the_deterministic_npv():
    k = oracle.par_rate2(M2, 0.0, PAY5, ACC5) + 0.001
    z = np.zeros((2, 1, 1000))
    t, ee, pfe = oracle.exposure_profile2(M2, 1e6, k, False, 0.0, PAY5, ACC5, [0.0], 1, z)
    v = max(oracle.irs_npv_reset2(M2, 0.0, 0.0, 0.0, 1e6, k, 0.0, PAY5, ACC5), 0.0)
    assert len(t) == 1 and ee[0] == v and pfe[0] == v


def test_cva2_closed_forms():
    t = np.array([0.5, 1.0, 2.0, 3.0])
    ee = np.full(4, 1000.0)
    assert oracle.unilateral_cva2(M2, t, ee, 0.0, 0.4) == 0.0
    h, rec = 0.02, 0.4
    s = np.exp(-h * t)
    dp = np.concatenate([[1.0], s[:-1]]) - s
    disc = np.array([oracle.zcb2(M2, 0.0, 0.0, 0.0, ti) for ti in t])
    assert abs(oracle.unilateral_cva2(M2, t, ee, h, rec) - np.sum((1 - rec) * ee * dp * disc)) < 1e-9


def _singleton(notional, k, start, pay, acc):
    return {"set_id": np.zeros(1, np.int32), "notional": np.array([notional]), "fixed_rate": np.array([k]),
            "start": np.array([start]), "flow_off": np.array([0, len(pay)], np.int32), "pay": np.array(pay),
            "acc": np.array(acc)}


def test_portfolio_run_on_a_singleton_is_the_reference_profile():
    n_paths, n_steps = 600, 52
    z = oracle.philox_normals2(9, n_paths, n_steps)
    k = oracle.par_rate2(M2, 0.0, PAY5, ACC5)
    times = [0.0, 1.0, 2.0, 3.0, 4.0]
    _, ee, pfe = oracle.exposure_profile2(M2, 1e6, k, False, 0.0, PAY5, ACC5, times, n_steps, z)
    out = oracle.portfolio_exposure(M2, _singleton(1e6, k, 0.0, PAY5, ACC5), times, n_steps, z)
    assert np.array_equal(out["set_ee"][0], ee) and np.array_equal(out["set_pfe"][0], pfe)
    assert ee[0] > 0.0 or pfe[0] == 0.0


@pytest.mark.parametrize("scheme", [0, 1])
def test_portfolio_run_reduces_to_one_factor(scheme):
    """eta = rho = 0: every valuation (reset and between-reset dates, forward starts) is the HW1F one
    with long-run mean r0; the Euler and exact recursions of x are those of r - r0."""
    a, _, s, _, _, r0 = M2
    from helpers import small_portfolio
    port = small_portfolio(n_trades=9, n_sets=2, seed=4).as_dict()
    dates = np.arange(0, 12 * 4 + 1) / 12.0
    n_paths, n_steps = 300, 48
    z1 = np.random.default_rng(6).standard_normal((n_steps, n_paths))
    z2 = np.stack([z1, np.random.default_rng(7).standard_normal((n_steps, n_paths))])
    one = oracle.portfolio_exposure((a, r0, s, r0), port, dates, n_steps, z1, scheme=scheme, hazard_rate=0.02,
                                    recovery_rate=0.4, want_trade=True)
    two = oracle.portfolio_exposure((a, 0.7, s, 0.0, 0.0, r0), port, dates, n_steps, z2, scheme=scheme,
                                    hazard_rate=0.02, recovery_rate=0.4, want_trade=True)
    scale = np.abs(port["notional"]).sum()
    for key in ("set_ee", "set_pfe", "set_cva", "trade_ee"):
        assert np.max(np.abs(two[key] - one[key])) < 1e-9 * scale, key
