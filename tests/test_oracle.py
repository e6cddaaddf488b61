"""Pins of the oracle (CPU FP64 restatement of the reference path) — runs without a GPU.

Each check is chosen so that a plausible mistake (dropped term, wrong sign or index, swapped
operand) fails: reference golden values, closed forms, known-answer vectors, invariants, brute
force on tiny inputs and Monte Carlo convergence within stated standard errors.
"""
import math

import numpy as np
import pytest

import oracle as o

A, B, SIG, R0 = 0.10, 0.03, 0.01, 0.02
PAY5, ACC5 = [1.0, 2.0, 3.0, 4.0, 5.0], [1.0] * 5


# ----------------------------------------------------------------------------- exact goldens
def test_config1_pv_and_dv01_match_reference_bit_for_bit():
    # PLAN_API_REFACTOR.md:330-331 (quantdesk run, market pillars [1,2] at 2%)
    assert o.npv_from_market([1.0, 2.0], [0.02, 0.02], 1e6, 0.02, 0.0, PAY5, ACC5) == 948.4537547220389
    assert o.dv01([1.0, 2.0], [0.02, 0.02], 1e6, 0.02, 0.0, PAY5, ACC5) == 480.18800212936185


def test_par_swap_dv01_and_zero_pv_match_reference():
    # clients/excel/README.md:324 and test_registry.cpp:245-260 (market [1] at 2%)
    k = o.par_rate_from_market([1.0], [0.02], 0.0, PAY5, ACC5)
    assert o.dv01([1.0], [0.02], 1e6, k, 0.0, PAY5, ACC5) == 480.4686940754473
    assert abs(o.npv_from_market([1.0], [0.02], 1e6, k, 0.0, PAY5, ACC5)) < 1e-6


def test_bucketed_dv01_sums_to_parallel_up_to_convexity():
    p, z = [1.0, 2.0, 3.0, 5.0], [0.018, 0.019, 0.02, 0.021]
    bk = o.dv01_bucketed(p, z, 1e6, 0.02, 0.0, PAY5, ACC5)
    par = o.dv01(p, z, 1e6, 0.02, 0.0, PAY5, ACC5)
    assert abs(bk.sum() - par) < 1e-3 * abs(par)
    # the 5y pillar only moves the flows beyond 3y
    assert bk[-1] > 0 and all(bk > -1e-9)


def test_pv_at_a_valuation_time():
    """measure.cpp:146-184: every discount factor over DF(v). v = 0 is the t = 0 form bit for bit; on a
    flat curve DF(t)/DF(v) = exp(-z (t - v)), so the value at v of a swap starting at s equals the t = 0
    value of the same swap moved v earlier; a swap already started (v > start) is refused (NaN)."""
    p, z = [1.0, 2.0, 3.0, 5.0], [0.018, 0.019, 0.02, 0.021]
    pay = [2.0, 3.0, 4.0, 5.0, 6.0]
    assert o.npv_from_market_at(p, z, 1e6, 0.02, 1.0, pay, ACC5, 0.0) == o.npv_from_market(p, z, 1e6, 0.02, 1.0, pay, ACC5)
    assert o.dv01_at(p, z, 1e6, 0.02, 1.0, pay, ACC5, 0.0) == o.dv01(p, z, 1e6, 0.02, 1.0, pay, ACC5)
    flat = [0.02] * 4
    for v in (0.25, 0.7, 1.0):
        moved = [t - v for t in pay]
        got = o.npv_from_market_at(p, flat, 1e6, 0.021, 1.0, pay, ACC5, v)
        ref = o.npv_from_market(p, flat, 1e6, 0.021, 1.0 - v, moved, ACC5)
        assert abs(got - ref) <= 1e-9 * 1e6, (v, got, ref)
        d_got = o.dv01_at(p, flat, 1e6, 0.021, 1.0, pay, ACC5, v)
        # the parallel bump moves DF(v) too: on a flat curve the DV01 at v is the moved swap's DV01
        d_ref = o.dv01(p, flat, 1e6, 0.021, 1.0 - v, moved, ACC5)
        assert abs(d_got - d_ref) <= 1e-9 * 1e6
    assert math.isnan(o.npv_from_market_at(p, z, 1e6, 0.02, 1.0, pay, ACC5, 1.5))


def test_curve_interpolation_and_flat_extrapolation():
    p, z = [1.0, 2.0, 4.0], [0.01, 0.02, 0.03]
    assert o.curve_zero_rate(p, z, 0.5) == 0.01
    assert o.curve_zero_rate(p, z, 9.0) == 0.03
    assert abs(o.curve_zero_rate(p, z, 3.0) - 0.025) < 1e-15
    assert abs(o.curve_df(p, z, 3.0) - math.exp(-0.025 * 3.0)) < 1e-15


# ----------------------------------------------------------------------------- Philox
@pytest.mark.parametrize("ctr,key,expected", [
    ([0, 0, 0, 0], [0, 0], [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]),
    ([0xffffffff] * 4, [0xffffffff] * 2, [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]),
    ([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], [0xa4093822, 0x299f31d0],
     [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]),
])
def test_philox4x32_10_known_answers(ctr, key, expected):
    # Random123 known-answer vectors (Salmon et al., SC'11)
    assert o.philox4x32_10(ctr, key) == expected


def test_philox_normals_are_standard_and_counter_based():
    z = o.philox_normals(7, 20000, 6)
    assert abs(z.mean()) < 4 / math.sqrt(z.size)
    assert abs(z.var() - 1.0) < 4 * math.sqrt(2.0 / z.size)
    # path p's draws do not depend on how many paths are generated
    z_small = o.philox_normals(7, 10, 6)
    assert np.array_equal(z_small, z[:, :10])
    u1, u2 = o.philox_uniforms(7, 3, 1)
    assert 0.0 < u1 <= 1.0 and 0.0 <= u2 < 1.0


# ----------------------------------------------------------------------------- model
def test_zcb_unit_at_zero_tau_and_decreasing():
    assert o.zcb(A, B, SIG, R0, 2.0, 2.0) == 1.0
    p = [o.zcb(A, B, SIG, R0, 0.0, T) for T in (1.0, 5.0, 10.0)]
    assert p[0] > p[1] > p[2] > 0


def test_zcb_without_volatility_is_the_deterministic_discount():
    # sigma = 0: r(s) = b + (r0 - b) e^{-as}, P(0,T) = exp(-(bT + (r0 - b)(1 - e^{-aT})/a))
    for T in (0.5, 3.0, 12.0):
        closed = math.exp(-(B * T + (R0 - B) * (1 - math.exp(-A * T)) / A))
        assert abs(o.zcb(A, B, 0.0, R0, 0.0, T) - closed) < 1e-14


def test_zcb_matches_monte_carlo_discount():
    # hull_white.rs:165-203: MC price of a 5y bond within a few standard errors (exact scheme,
    # trapezoid discount, dt = 1/100: discretisation bias ~1e-7, far below the SE)
    n, T, steps = 40000, 5.0, 500
    z = o.philox_normals(3, n, steps)
    r = o.simulate(A, B, 0.03, R0, T / steps, z, scheme=1)
    integral = (r[:-1] + r[1:]).sum(axis=0) * 0.5 * (T / steps)
    disc = np.exp(-integral)
    se = disc.std(ddof=1) / math.sqrt(n)
    assert abs(disc.mean() - o.zcb(A, B, 0.03, R0, 0.0, T)) < 4 * se


def test_euler_and_exact_steps_follow_their_recursions():
    steps, dt = 12, 0.25
    z = np.zeros((steps, 1))
    r = o.simulate(A, B, SIG, R0, dt, z, scheme=0)[:, 0]
    k = np.arange(steps + 1)
    assert np.allclose(r, B + (R0 - B) * (1 - A * dt) ** k, rtol=0, atol=1e-16)
    r = o.simulate(A, B, SIG, R0, dt, z, scheme=1)[:, 0]
    assert np.allclose(r, B + (R0 - B) * np.exp(-A * dt * k), rtol=0, atol=1e-16)
    # a unit shock at step j propagates as sigma sqrt(dt) (1 - a dt)^(k-j-1) (Euler)
    z[4, 0] = 1.0
    r1 = o.simulate(A, B, SIG, R0, dt, z, scheme=0)[:, 0]
    r0 = o.simulate(A, B, SIG, R0, dt, np.zeros_like(z), scheme=0)[:, 0]
    expect = np.where(k > 4, SIG * math.sqrt(dt) * (1 - A * dt) ** np.maximum(k - 5, 0), 0.0)
    assert np.allclose(r1 - r0, expect, rtol=0, atol=1e-16)
    # exact scheme: conditional sd sigma sqrt((1 - e^{-2a dt}) / (2a))
    r1 = o.simulate(A, B, SIG, R0, dt, z, scheme=1)[:, 0]
    r0 = o.simulate(A, B, SIG, R0, dt, np.zeros_like(z), scheme=1)[:, 0]
    sd = SIG * math.sqrt((1 - math.exp(-2 * A * dt)) / (2 * A))
    assert abs((r1 - r0)[5] - sd) < 1e-17


# ----------------------------------------------------------------------------- IRS
def test_par_swap_has_zero_model_npv_and_payer_gains_when_rates_rise():
    k = o.par_rate(A, B, SIG, R0, 0.0, PAY5, ACC5)
    assert abs(o.irs_npv_reset(A, B, SIG, R0, 0.0, 1e6, k, 0.0, PAY5, ACC5)) < 1e-6
    assert o.irs_npv_reset(A, B, SIG, 0.03, 0.0, 1e6, k, 0.0, PAY5, ACC5) > 0


def test_reset_dates_and_remaining_swap():
    assert o.is_reset_date(0.0, PAY5, 0.0) and o.is_reset_date(0.0, PAY5, 4.0)
    assert not o.is_reset_date(0.0, PAY5, 5.0) and not o.is_reset_date(0.0, PAY5, 0.5)
    # at t = 2 the remaining swap is the 3y swap starting at 2: N(1 - P(2,5)) - NK sum_{3,4,5} P(2,Ti)
    r, k = 0.025, 0.02
    p = [o.zcb(A, B, SIG, r, 2.0, T) for T in (3.0, 4.0, 5.0)]
    expect = 1e6 * (1 - p[-1]) - 1e6 * k * sum(p)
    assert abs(o.irs_npv_reset(A, B, SIG, r, 2.0, 1e6, k, 0.0, PAY5, ACC5) - expect) < 1e-8
    assert o.irs_value(A, B, SIG, r, 2.0, 0.0, 1e6, k, 0.0, PAY5, ACC5) == \
        o.irs_npv_reset(A, B, SIG, r, 2.0, 1e6, k, 0.0, PAY5, ACC5)


def test_between_reset_value_without_volatility_is_the_forward_value():
    # sigma = 0: P(t,T | r(t)) = P(0,T) / P(0,t), so the floating leg at t in (T_{k-1}, T_k) is
    # (P(0,T_{k-1}) - P(0,T_n)) / P(0,t) and the fixed leg is K sum_{T_i > t} P(0,T_i) / P(0,t)
    t, k = 1.5, 0.021
    r_t = B + (R0 - B) * math.exp(-A * t)
    r_fix = B + (R0 - B) * math.exp(-A * 1.0)
    P0 = lambda T: o.zcb(A, B, 0.0, R0, 0.0, T)
    fl = (P0(1.0) - P0(5.0)) / P0(t)
    fx = k * sum(P0(T) for T in (2.0, 3.0, 4.0, 5.0)) / P0(t)
    v = o.irs_value(A, B, 0.0, r_t, t, r_fix, 1e6, k, 0.0, PAY5, ACC5)
    assert abs(v - 1e6 * (fl - fx)) < 1e-7
    assert o.fixing_date(t, 0.0, PAY5) == 1.0 and o.fixing_date(2.0, 0.0, PAY5) == -1.0


def test_between_reset_floating_leg_is_a_martingale():
    # E[D(0,t) * P(t,T_k)/P(T_{k-1},T_k | r(T_{k-1}))] = P(0,T_{k-1}) for t in the period
    n, dt = 40000, 0.01
    z = o.philox_normals(5, n, 150)
    r = o.simulate(A, B, 0.02, R0, dt, z, scheme=1)
    t, t_fix, t_pay = 1.5, 1.0, 2.0
    disc = np.exp(-(r[:-1] + r[1:]).sum(axis=0) * 0.5 * dt)
    fl = np.array([o.zcb(A, B, 0.02, r[150, p], t, t_pay) / o.zcb(A, B, 0.02, r[100, p], t_fix, t_pay)
                   for p in range(n)])
    x = disc * fl
    se = x.std(ddof=1) / math.sqrt(n)
    assert abs(x.mean() - o.zcb(A, B, 0.02, R0, 0.0, t_fix)) < 4 * se + 1e-6


# ----------------------------------------------------------------------------- EE / PFE / CVA
@pytest.mark.parametrize("n", [1, 2, 19, 20, 21, 100, 1001])
def test_pfe_is_the_ceil_095_order_statistic(n):
    x = np.random.default_rng(n).exponential(size=n)
    ee, pfe = o.ee_pfe(x)
    q = min(max(math.ceil(0.95 * n) - 1, 0), n - 1)
    assert pfe == np.sort(x)[q]
    assert abs(ee - x.mean()) <= 1e-15 * abs(x).sum()


def test_cva_closed_forms():
    # exposure.rs:729-759 profiles: linear in EE, zero without default risk, one-date closed form
    t = [0.0, 1.0, 2.0]
    c1 = o.unilateral_cva(A, B, SIG, R0, t, [0.0, 12000.0, 9500.0], 0.02, 0.4)
    c2 = o.unilateral_cva(A, B, SIG, R0, t, [0.0, 24000.0, 19000.0], 0.02, 0.4)
    assert abs(c2 - 2 * c1) < 1e-10 and c1 > 0
    assert o.unilateral_cva(A, B, SIG, R0, t, [0.0, 12000.0, 9500.0], 0.0, 0.4) == 0.0
    one = o.unilateral_cva(A, B, SIG, R0, [2.0], [1000.0], 0.03, 0.25)
    assert abs(one - 0.75 * 1000.0 * (1 - math.exp(-0.06)) * o.zcb(A, B, SIG, R0, 0.0, 2.0)) < 1e-12


def _batch_stats(x, n_batches):
    """Mean and the standard error of a len(x)/n_batches-path estimate of EE and PFE95."""
    xb = x.reshape(n_batches, -1)
    ee_b = xb.mean(axis=1)
    q = math.ceil(0.95 * xb.shape[1]) - 1
    pfe_b = np.sort(xb, axis=1)[:, q]
    return ee_b, pfe_b


def test_seed7_goldens_within_three_standard_errors():
    """Statistical anchors of the par 5y swap (seed 7, 5000 paths, 208 steps; burn-ndarray stream,
    clients/excel/README.md:321-335): our estimate from 100k Philox paths against the golden,
    within 3 combined standard errors (golden SE = std of 5000-path batch estimates)."""
    gold_ee = [0.0, 12862.617942080262, 13673.529752568928, 11957.816098484913, 7124.106240024746]
    gold_pfe = [0.0, 51009.920875152675, 53607.17081592724, 46152.44561612198, 27535.557267142714]
    n, steps = 100000, 208
    z = o.philox_normals(2024, n, steps)
    r = o.simulate(A, B, SIG, R0, 4.0 / steps, z)
    k = o.par_rate(A, B, SIG, R0, 0.0, PAY5, ACC5)
    assert abs(max(o.irs_npv_reset(A, B, SIG, R0, 0.0, 1e6, k, 0.0, PAY5, ACC5), 0.0)) < 1e-6
    ee_prof, ee_se = [0.0], [0.0]
    for d, t in enumerate([1.0, 2.0, 3.0, 4.0], start=1):
        x = np.array([max(o.irs_npv_reset(A, B, SIG, rr, t, 1e6, k, 0.0, PAY5, ACC5), 0.0) for rr in r[52 * d]])
        ee_b, pfe_b = _batch_stats(x, 20)
        se_gold_ee, se_gold_pfe = ee_b.std(ddof=1), pfe_b.std(ddof=1)
        se_ee = math.hypot(se_gold_ee, se_gold_ee / math.sqrt(20))
        se_pfe = math.hypot(se_gold_pfe, se_gold_pfe / math.sqrt(20))
        assert abs(x.mean() - gold_ee[d]) < 3 * se_ee, (t, x.mean(), gold_ee[d], se_ee)
        assert abs(np.mean(pfe_b) - gold_pfe[d]) < 3 * se_pfe, (t, np.mean(pfe_b), gold_pfe[d], se_pfe)
        ee_prof.append(x.mean())
        ee_se.append(se_ee)
    cva = o.unilateral_cva(A, B, SIG, R0, [0, 1, 2, 3, 4.0], ee_prof, 0.02, 0.4)
    # CVA is linear in EE: its SE is bounded by the per-date SEs summed with the CVA weights
    cva_se = o.unilateral_cva(A, B, SIG, R0, [0, 1, 2, 3, 4.0], ee_se, 0.02, 0.4)
    assert abs(cva - 503.64194077997536) < 3 * cva_se, (cva, cva_se)


def test_reference_profile_equals_singleton_portfolio_run():
    n, steps = 300, 52
    z = o.philox_normals(1, n, steps)
    times = [0.0, 1.0, 2.0, 3.0, 4.0]
    _, ee, pfe = o.exposure_profile(A, B, SIG, R0, 1e6, 0.021, False, 0.0, PAY5, ACC5, times, steps, z)
    port = dict(set_id=np.array([0], dtype=np.int32), notional=np.array([1e6]), fixed_rate=np.array([0.021]),
                start=np.array([0.0]), flow_off=np.array([0, 5], dtype=np.int32), pay=np.array(PAY5),
                acc=np.array(ACC5))
    res = o.portfolio_exposure((A, B, SIG, R0), port, times, steps, z, want_trade=True)
    assert np.array_equal(res["set_ee"][0], ee) and np.array_equal(res["set_pfe"][0], pfe)
    assert np.array_equal(res["trade_ee"][0], ee)


def test_netting_set_aggregates_the_right_trades():
    rng = np.random.default_rng(0)
    port = dict(set_id=np.array([1, 0, 1], dtype=np.int32), notional=np.array([1e6, -2e6, 5e5]),
                fixed_rate=np.array([0.02, 0.025, 0.015]), start=np.zeros(3),
                flow_off=np.array([0, 5, 8, 10], dtype=np.int32),
                pay=np.array(PAY5 + [1.0, 2.0, 3.0, 0.5, 1.0]), acc=np.array(ACC5 + [1.0] * 3 + [0.5, 0.5]))
    dates = [0.25, 0.5, 1.0, 2.5]
    z = rng.standard_normal((10, 400))
    res = o.portfolio_exposure((A, B, SIG, R0), port, dates, 10, z, want_trade=True, want_path=True)
    v = res["path_mtm"]
    for s, members in ((0, [1]), (1, [0, 2])):
        x = np.maximum(v[:, members, :].sum(axis=1), 0.0)
        assert np.allclose(res["set_ee"][s], x.mean(axis=1), rtol=1e-13, atol=1e-9)
        assert np.allclose(res["set_mtm"][:, s, :], v[:, members, :].sum(axis=1), rtol=1e-13, atol=1e-9)
    # a one-trade set is that trade's own exposure
    assert np.allclose(res["set_ee"][0], res["trade_ee"][1], rtol=0, atol=0)
