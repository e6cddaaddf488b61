"""GPU parity: the HIP path through the C ABI against the oracle on the same inputs.

Tolerances (FP64): per-path MtM within 1e-12 of the legs' magnitude; EE / PFE / CVA within 1e-10
relative on shared host draws; Philox runs within 1e-10 (the GPU and CPU draw bit-identical
uniforms; log/sin/cos may differ in the last ulp).
"""
import numpy as np
import pytest

import oracle
from engine_quant_amd import eqhip
from helpers import MODEL, assert_rel, assert_value_rel, leg_scale, small_portfolio

pytestmark = pytest.mark.gpu

A, B, SIG, R0 = MODEL
PAY5, ACC5 = [1.0, 2.0, 3.0, 4.0, 5.0], [1.0] * 5
RESET5 = [0.0, 1.0, 2.0, 3.0, 4.0]


def test_reference_profile_shared_draws(ctx):
    """exposure.rs:73-138 on the golden case (par 5y swap, 208 steps) with shared draws."""
    n_paths, n_steps = 5000, 208
    z = oracle.philox_normals(7, n_paths, n_steps)
    t, ee, pfe = ctx.exposure_profile(A, B, SIG, R0, 1e6, 0.0, True, 0.0, PAY5, ACC5, RESET5, n_steps, n_paths,
                                      7, rng=eqhip.RNG_HOST_DRAWS, host_draws=z)
    t_o, ee_o, pfe_o = oracle.exposure_profile(A, B, SIG, R0, 1e6, 0.0, True, 0.0, PAY5, ACC5, RESET5, n_steps, z)
    assert np.array_equal(t, t_o)
    assert_rel(ee, ee_o, 1e-10, what="EE")
    assert_rel(pfe, pfe_o, 1e-10, what="PFE95")
    cva = ctx.unilateral_cva(A, B, SIG, R0, t, ee, 0.02, 0.4)
    assert_rel(cva, oracle.unilateral_cva(A, B, SIG, R0, t_o, ee_o, 0.02, 0.4), 1e-10, what="CVA")


def test_reference_profile_philox_matches_oracle_stream(ctx):
    n_paths, n_steps = 4096, 104
    z = oracle.philox_normals(11, n_paths, n_steps)
    times = [0.0, 1.0, 2.0]
    _, ee, pfe = ctx.exposure_profile(A, B, SIG, R0, 1e6, 0.0191, False, 0.0, PAY5, ACC5, times, n_steps, n_paths,
                                      11, rng=eqhip.RNG_PHILOX)
    _, ee_o, pfe_o = oracle.exposure_profile(A, B, SIG, R0, 1e6, 0.0191, False, 0.0, PAY5, ACC5, times, n_steps, z)
    assert_rel(ee, ee_o, 1e-10, what="EE philox")
    assert_rel(pfe, pfe_o, 1e-10, what="PFE philox")


def test_reference_batch_matches_oracle(ctx):
    """Homogeneous batch (exposure.rs:169-235): per-trade profiles on one simulation."""
    n_paths, n_steps = 2000, 104
    z = oracle.philox_normals(7, n_paths, n_steps)
    notionals, rates = [1e6, 2e6, 5e5], [0.018, 0.02, 0.022]
    times = [0.0, 1.0, 2.0]
    _, ee, pfe = ctx.exposure_profile_batch(A, B, SIG, R0, notionals, rates, 0.0, PAY5, ACC5, times, n_steps,
                                            n_paths, 7, rng=eqhip.RNG_HOST_DRAWS, host_draws=z)
    for i, (n, k) in enumerate(zip(notionals, rates)):
        _, ee_o, pfe_o = oracle.exposure_profile(A, B, SIG, R0, n, k, False, 0.0, PAY5, ACC5, times, n_steps, z)
        assert_rel(ee[i], ee_o, 1e-10, what=f"EE trade {i}")
        assert_rel(pfe[i], pfe_o, 1e-10, what=f"PFE trade {i}")


@pytest.mark.parametrize("disc", [eqhip.DISC_EULER, eqhip.DISC_EXACT])
def test_portfolio_netting_shared_draws(ctx, disc):
    """Netting sets, mixed schedules, forward starts, monthly dates between resets."""
    port = small_portfolio(n_trades=14, n_sets=3, seed=5)
    dates = np.arange(0, 12 * 6 + 1) / 12.0
    n_paths, n_steps = 700, 72
    z = np.random.default_rng(1).standard_normal((n_steps, n_paths))
    got = ctx.exposure_run(MODEL, port, dates, n_paths, n_steps, rng=eqhip.RNG_HOST_DRAWS, disc=disc,
                           host_draws=z, hazard_rate=0.02, recovery_rate=0.4, want_trade_ee=True,
                           want_path=True, want_set_mtm=True)
    ref = oracle.portfolio_exposure(MODEL, port.as_dict(), dates, n_steps, z, scheme=disc, hazard_rate=0.02,
                                    recovery_rate=0.4, want_trade=True, want_path=True)
    scale = leg_scale(port, MODEL, dates)
    assert_rel(got["path_mtm"], ref["path_mtm"], 1e-12, scale=scale[None, :, None], what="per-path MtM")
    set_scale = np.array([scale[port.set_id == s].sum() for s in range(port.n_sets)])
    assert_rel(got["set_mtm"], ref["set_mtm"], 1e-12, scale=set_scale[None, :, None], what="set MtM")
    assert_value_rel(got["set_ee"], ref["set_ee"], leg=set_scale, what="set EE")
    assert_value_rel(got["set_pfe"], ref["set_pfe"], leg=set_scale, what="set PFE")
    assert_value_rel(got["set_cva"], ref["set_cva"], leg=set_scale, what="set CVA")
    assert_value_rel(got["trade_ee"], ref["trade_ee"], leg=scale, what="trade EE")
    assert_value_rel(got["trade_cva"], ref["trade_cva"], leg=scale, what="trade CVA")


def test_pfe_is_an_exact_order_statistic(ctx):
    """The radix select returns a sample, bit-identical to the oracle's sorted pick."""
    port = small_portfolio(n_trades=6, n_sets=2, seed=9)
    dates = np.arange(1, 25) / 4.0
    n_paths, n_steps = 1237, 24
    z = np.random.default_rng(2).standard_normal((n_steps, n_paths))
    got = ctx.exposure_run(MODEL, port, dates, n_paths, n_steps, rng=eqhip.RNG_HOST_DRAWS, host_draws=z,
                           want_set_mtm=True)
    for s in range(port.n_sets):
        for k in range(len(dates)):
            x = np.maximum(got["set_mtm"][k, s], 0.0)
            q = min(int(np.ceil(0.95 * n_paths)) - 1, n_paths - 1)
            assert got["set_pfe"][s, k] == np.sort(x)[q]


def test_pv_dv01_matches_oracle(ctx):
    pillars, zeros = [1.0, 2.0], [0.02, 0.02]
    port = eqhip.singleton_portfolio([1e6], [0.02], 0.0, PAY5, ACC5)
    pv, dv, bk = ctx.pv_dv01(pillars, zeros, port)
    assert abs(pv[0] - 948.4537547220389) < 1e-9
    assert abs(dv[0] - 480.18800212936185) < 1e-9
    port2 = small_portfolio(n_trades=9, n_sets=1, seed=4)
    p20 = np.linspace(0.25, 30, 20)
    z20 = np.linspace(0.015, 0.035, 20)
    pv, dv, bk = ctx.pv_dv01(p20, z20, port2)
    for j in range(port2.n_trades):
        f0, f1 = port2.flow_off[j], port2.flow_off[j + 1]
        args = (port2.notional[j], port2.fixed_rate[j], port2.start[j], port2.pay[f0:f1], port2.acc[f0:f1])
        assert_rel(pv[j], oracle.npv_from_market(p20, z20, *args), 1e-12, scale=abs(port2.notional[j]))
        assert_rel(dv[j], oracle.dv01(p20, z20, *args), 1e-12, scale=abs(port2.notional[j]))
        assert_rel(bk[j], oracle.dv01_bucketed(p20, z20, *args), 1e-12, scale=abs(port2.notional[j]))


def test_pv_dv01_at_a_valuation_time(ctx):
    """eqhip_pv_dv01_at: PV and parallel DV01 at valuation time v against the oracle, v = 0 equal to
    eqhip_pv_dv01 bit for bit, the bucketed ladder at t = 0, a trade started before v refused."""
    port = small_portfolio(n_trades=9, n_sets=1, seed=4, start=1.0)
    p20, z20 = np.linspace(0.25, 30, 20), np.linspace(0.015, 0.035, 20)
    base = ctx.pv_dv01(p20, z20, port)
    at0 = ctx.pv_dv01(p20, z20, port, valuation_time=0.0)
    for a, b in zip(base, at0):
        assert np.array_equal(a, b)
    v = 0.6
    pv, dv, bk = ctx.pv_dv01(p20, z20, port, valuation_time=v)
    assert np.array_equal(bk, base[2])
    for j in range(port.n_trades):
        f0, f1 = port.flow_off[j], port.flow_off[j + 1]
        args = (port.notional[j], port.fixed_rate[j], port.start[j], port.pay[f0:f1], port.acc[f0:f1])
        assert_rel(pv[j], oracle.npv_from_market_at(p20, z20, *args, v), 1e-12, scale=abs(port.notional[j]))
        assert_rel(dv[j], oracle.dv01_at(p20, z20, *args, v), 1e-12, scale=abs(port.notional[j]))
    with pytest.raises(eqhip.EqhipError) as e:
        ctx.pv_dv01(p20, z20, port, valuation_time=float(port.start.min()) + 0.1)
    assert e.value.code == eqhip.E_INVALID and "posterior al start" in str(e.value)


def test_deterministic_branch_returns_one_date(ctx):
    """exposure.rs:98-108: t_max <= 0 returns a single deterministic point."""
    t, ee, pfe = ctx.exposure_profile(A, B, SIG, R0, 1e6, 0.025, False, 0.0, PAY5, ACC5, [0.0], 10, 100, 1)
    v = max(oracle.irs_npv_reset(A, B, SIG, R0, 0.0, 1e6, 0.025, 0.0, PAY5, ACC5), 0.0)
    assert len(ee) == 1 and abs(ee[0] - v) <= 1e-9 and ee[0] == pfe[0]


def test_invalid_monitoring_date_is_rejected(ctx):
    with pytest.raises(eqhip.EqhipError) as e:
        ctx.exposure_profile(A, B, SIG, R0, 1e6, 0.02, False, 0.0, PAY5, ACC5, [0.5], 10, 100, 1)
    assert e.value.code == eqhip.E_INVALID


def test_many_tiles_per_unit_shared_draws(ctx):
    """A launch whose work units span several 64-path tiles (the bench-size regime: short rates are
    prefetched a tile ahead and the date's flow list is staged once per unit) against the oracle."""
    port = small_portfolio(n_trades=30, n_sets=2, seed=8)
    dates = np.arange(1, 7) / 3.0
    n_paths, n_steps = 120_000, 24
    z = np.random.default_rng(3).standard_normal((n_steps, n_paths))
    plan = ctx.plan(MODEL, port, dates, n_paths, n_steps, rng=eqhip.RNG_HOST_DRAWS, disc=eqhip.DISC_EXACT,
                    want_trade_ee=True)
    got = plan.execute(z, 0.02, 0.4)
    assert plan.timing()["tiles_per_block"] > 1
    plan.close()
    ref = oracle.portfolio_exposure(MODEL, port.as_dict(), dates, n_steps, z, scheme=1, hazard_rate=0.02,
                                    recovery_rate=0.4, want_trade=True)
    scale = leg_scale(port, MODEL, dates)
    set_scale = np.array([scale[port.set_id == s].sum() for s in range(port.n_sets)])
    assert_value_rel(got["set_ee"], ref["set_ee"], leg=set_scale, what="set EE")
    assert_value_rel(got["set_pfe"], ref["set_pfe"], leg=set_scale, what="set PFE")
    assert_value_rel(got["trade_ee"], ref["trade_ee"], leg=scale, what="trade EE")


def staggered_annual_portfolio(n_sets=12, seed=6):
    """Annual swaps starting one month apart (12 families, one per start month), two per start
    (2y and 3y), trade j in set j mod n_sets. Between resets every family is in a period fixed at
    a different month, so a monitoring date carries 12 fixings."""
    rng = np.random.default_rng(seed)
    pays, accs, off, starts = [], [], [0], []
    for m in range(12):
        for years in (2, 3):
            s0 = m / 12.0
            pays.append(s0 + np.arange(1, years + 1, dtype=np.float64))
            accs.append(np.ones(years))
            off.append(off[-1] + years)
            starts.append(s0)
    n = len(starts)
    notional = np.exp(rng.normal(np.log(1e6), 0.5, n)) * np.where(rng.uniform(size=n) < 0.5, 1, -1)
    return eqhip.Portfolio(set_id=(np.arange(n) % n_sets).astype(np.int32), notional=notional,
                           fixed_rate=0.02 + rng.normal(0, 0.004, n), start=np.array(starts),
                           flow_off=np.array(off, dtype=np.int32), pay=np.concatenate(pays), acc=np.concatenate(accs))


def test_many_sets_and_more_fixings_than_waves(ctx):
    """More netting sets than the split limit (one wave per set) and more fixings at a date than
    the block has waves, against the oracle, with per-path checks."""
    port = staggered_annual_portfolio(n_sets=12)
    assert port.n_sets > 8
    dates = np.concatenate([[0.0], (np.arange(35) + 0.5) / 12.0])  # between the monthly resets
    n_paths, n_steps = 900, 69  # dt = 1/24: every reset and date on the grid
    t = dates[20]
    mid = {round(s0 + np.floor(t - s0), 9) for s0 in port.start if s0 < t}
    assert len(mid) > 8  # fixings at this date
    z = np.random.default_rng(4).standard_normal((n_steps, n_paths))
    got = ctx.exposure_run(MODEL, port, dates, n_paths, n_steps, rng=eqhip.RNG_HOST_DRAWS, disc=eqhip.DISC_EXACT,
                           host_draws=z, hazard_rate=0.02, recovery_rate=0.4, want_trade_ee=True,
                           want_path=True, want_set_mtm=True)
    ref = oracle.portfolio_exposure(MODEL, port.as_dict(), dates, n_steps, z, scheme=1, hazard_rate=0.02,
                                    recovery_rate=0.4, want_trade=True, want_path=True)
    scale = leg_scale(port, MODEL, dates)
    set_scale = np.array([scale[port.set_id == s].sum() for s in range(port.n_sets)])
    assert_rel(got["path_mtm"], ref["path_mtm"], 1e-12, scale=scale[None, :, None], what="per-path MtM")
    assert_rel(got["set_mtm"], ref["set_mtm"], 1e-12, scale=set_scale[None, :, None], what="set MtM")
    assert_value_rel(got["set_ee"], ref["set_ee"], leg=set_scale, what="set EE")
    assert_value_rel(got["set_pfe"], ref["set_pfe"], leg=set_scale, what="set PFE")
    assert_value_rel(got["set_cva"], ref["set_cva"], leg=set_scale, what="set CVA")
    assert_value_rel(got["trade_ee"], ref["trade_ee"], leg=scale, what="trade EE")
