"""G2++ exposure on the GPU (the HullWhite2F arm of exposure.rs:305-480) against the oracle.

Tolerances as for HW1F: per-path MtM within 1e-12 of the legs' magnitude; EE / PFE / CVA within
1e-10 relative on shared draws and on the Philox stream (bit-identical uniforms on both sides).
"""
import numpy as np
import pytest

import engine_quant_amd as engine
import oracle
from engine_quant_amd import eqhip
from helpers import assert_rel, assert_value_rel, leg_scale, small_portfolio

pytestmark = pytest.mark.gpu

M2 = (0.1, 0.2, 0.01, 0.012, -0.7, 0.03)  # hull_white_2f.rs tests::reference_model
PAY5, ACC5 = [1.0, 2.0, 3.0, 4.0, 5.0], [1.0] * 5
RESET5 = [0.0, 1.0, 2.0, 3.0, 4.0]


def test_reference_profile_shared_draws(ctx):
    """exposure.rs:305-370 on the reference's 2F case (par 5y swap, 208 steps), shared draws."""
    n_paths, n_steps = 5000, 208
    z = oracle.philox_normals2(7, n_paths, n_steps)
    t, ee, pfe = ctx.exposure_profile_2f(*M2, 1e6, 0.0, True, 0.0, PAY5, ACC5, RESET5, n_steps, n_paths, 7,
                                         rng=eqhip.RNG_HOST_DRAWS, host_draws=z)
    t_o, ee_o, pfe_o = oracle.exposure_profile2(M2, 1e6, 0.0, True, 0.0, PAY5, ACC5, RESET5, n_steps, z)
    assert np.array_equal(t, t_o)
    assert_rel(ee, ee_o, 1e-10, what="EE")
    assert_rel(pfe, pfe_o, 1e-10, what="PFE95")
    assert np.all(ee >= 0) and np.all(pfe >= ee)
    cva = ctx.unilateral_cva_2f(*M2, t, ee, 0.02, 0.4)
    assert_rel(cva, oracle.unilateral_cva2(M2, t_o, ee_o, 0.02, 0.4), 1e-10, what="CVA")
    assert cva > 0.0 and ctx.unilateral_cva_2f(*M2, t, ee, 0.0, 0.4) == 0.0


def test_reference_profile_philox_matches_oracle_stream(ctx):
    n_paths, n_steps = 4096, 104
    z = oracle.philox_normals2(11, n_paths, n_steps)
    times = [0.0, 1.0, 2.0]
    _, ee, pfe = ctx.exposure_profile_2f(*M2, 1e6, 0.0305, False, 0.0, PAY5, ACC5, times, n_steps, n_paths, 11)
    _, ee_o, pfe_o = oracle.exposure_profile2(M2, 1e6, 0.0305, False, 0.0, PAY5, ACC5, times, n_steps, z)
    assert_rel(ee, ee_o, 1e-10, what="EE philox")
    assert_rel(pfe, pfe_o, 1e-10, what="PFE philox")


def test_time_zero_only_is_the_deterministic_npv(ctx):
    """exposure.rs:606-617: monitoring at t = 0 alone gives max(NPV(0, 0), 0) exactly once."""
    k = oracle.par_rate2(M2, 0.0, PAY5, ACC5) + 0.001
    t, ee, pfe = ctx.exposure_profile_2f(*M2, -1e6, k, False, 0.0, PAY5, ACC5, [0.0], 1, 1000, 11)
    v = max(oracle.irs_npv_reset2(M2, 0.0, 0.0, 0.0, -1e6, k, 0.0, PAY5, ACC5), 0.0)
    assert len(t) == 1
    assert_rel(ee, [v], 1e-12, what="EE t=0")
    assert_rel(pfe, [v], 1e-12, what="PFE t=0")


def test_reference_batch_and_cva_batch(ctx):
    """exposure.rs:378-447 and 468-480: per-trade profiles on one simulation."""
    n_paths, n_steps = 2000, 104
    z = oracle.philox_normals2(7, n_paths, n_steps)
    notionals, rates = [1e6, 2.5e6, -5e5], [0.02, 0.015, 0.031]
    times = [0.0, 1.0, 2.0, 3.0]
    t, ee, pfe = ctx.exposure_profile_2f_batch(*M2, notionals, rates, 0.0, PAY5, ACC5, times, n_steps, n_paths, 7,
                                               rng=eqhip.RNG_HOST_DRAWS, host_draws=z)
    cvas = ctx.unilateral_cva_2f_batch(*M2, t, ee, 0.02, 0.4)
    for i, (n, k) in enumerate(zip(notionals, rates)):
        _, ee_o, pfe_o = oracle.exposure_profile2(M2, n, k, False, 0.0, PAY5, ACC5, times, n_steps, z)
        assert_rel(ee[i], ee_o, 1e-10, what=f"EE trade {i}")
        assert_rel(pfe[i], pfe_o, 1e-10, what=f"PFE trade {i}")
        assert_rel(cvas[i], oracle.unilateral_cva2(M2, times, ee_o, 0.02, 0.4), 1e-10, what=f"CVA trade {i}")


@pytest.mark.parametrize("disc", [eqhip.DISC_EULER, eqhip.DISC_EXACT])
def test_portfolio_netting_shared_draws(ctx, disc):
    """Netting sets, mixed schedules, forward starts and monthly dates between resets under G2++."""
    port = small_portfolio(n_trades=14, n_sets=3, seed=5)
    dates = np.arange(0, 12 * 6 + 1) / 12.0
    n_paths, n_steps = 700, 72
    z = np.random.default_rng(1).standard_normal((2, n_steps, n_paths))
    got = ctx.exposure_run(M2, port, dates, n_paths, n_steps, rng=eqhip.RNG_HOST_DRAWS, disc=disc, host_draws=z,
                           hazard_rate=0.02, recovery_rate=0.4, want_trade_ee=True, want_path=True,
                           want_set_mtm=True)
    ref = oracle.portfolio_exposure(M2, port.as_dict(), dates, n_steps, z, scheme=disc, hazard_rate=0.02,
                                    recovery_rate=0.4, want_trade=True, want_path=True)
    scale = leg_scale(port, (0.1, 0.03, 0.01, 0.03), dates)
    assert_rel(got["path_mtm"], ref["path_mtm"], 1e-12, scale=scale[None, :, None], what="per-path MtM")
    set_scale = np.array([scale[port.set_id == s].sum() for s in range(port.n_sets)])
    assert_rel(got["set_mtm"], ref["set_mtm"], 1e-12, scale=set_scale[None, :, None], what="set MtM")
    assert_value_rel(got["set_ee"], ref["set_ee"], leg=set_scale, what="set EE")
    assert_value_rel(got["set_pfe"], ref["set_pfe"], leg=set_scale, what="set PFE")
    assert_value_rel(got["set_cva"], ref["set_cva"], leg=set_scale, what="set CVA")
    assert_value_rel(got["trade_ee"], ref["trade_ee"], leg=scale, what="trade EE")
    assert_value_rel(got["trade_cva"], ref["trade_cva"], leg=scale, what="trade CVA")


def test_portfolio_philox_more_fixings_than_waves(ctx):
    """Quarterly and monthly schedules starting at staggered dates: many fixings per date."""
    rng = np.random.default_rng(8)
    pays, accs, off, starts = [], [], [0], []
    for j in range(40):
        f = (4, 12)[j % 2]
        s0 = (j % 23) / 24.0
        pay = s0 + np.arange(1, 3 * f + 1) / f
        pays.append(pay)
        accs.append(np.full(len(pay), 1.0 / f))
        off.append(off[-1] + len(pay))
        starts.append(s0)
    port = eqhip.Portfolio(set_id=(np.arange(40) % 5).astype(np.int32),
                           notional=rng.normal(0, 1e6, 40), fixed_rate=0.03 + rng.normal(0, 0.004, 40),
                           start=np.array(starts), flow_off=np.array(off, np.int32), pay=np.concatenate(pays),
                           acc=np.concatenate(accs))
    dates = np.arange(1, 4 * 24) / 24.0
    n_paths, n_steps = 1500, 190
    got = ctx.exposure_run(M2, port, dates, n_paths, n_steps, seed=21, hazard_rate=0.02, recovery_rate=0.4)
    z = oracle.philox_normals2(21, n_paths, n_steps)
    ref = oracle.portfolio_exposure(M2, port.as_dict(), dates, n_steps, z, hazard_rate=0.02, recovery_rate=0.4)
    set_scale = np.array([np.abs(port.notional[port.set_id == s]).sum() * 1.2 for s in range(5)])
    assert_value_rel(got["set_ee"], ref["set_ee"], leg=set_scale, what="set EE")
    assert_value_rel(got["set_pfe"], ref["set_pfe"], leg=set_scale, what="set PFE")
    assert_value_rel(got["set_cva"], ref["set_cva"], leg=set_scale, what="set CVA")


def test_g2pp_refusals(ctx):
    port = small_portfolio(n_trades=3, n_sets=1, seed=2)
    plan = ctx.plan(M2, port, [0.5, 1.0], 256, 8)
    try:
        with pytest.raises(eqhip.EqhipError, match="HullWhite1F"):
            plan.cva_adjoint(0, 0.02, 0.4)
    finally:
        plan.close()
    with pytest.raises(eqhip.EqhipError, match="rho"):
        ctx.exposure_profile_2f(0.1, 0.2, 0.01, 0.012, -1.5, 0.03, 1e6, 0.03, False, 0.0, PAY5, ACC5, RESET5, 8,
                                100, 1)


def test_measures_dispatch_on_hull_white_2f(ctx):
    """measure.cpp:53-80 / 257-282: the same measures serve HullWhite2F (test_registry.cpp:285-405)."""
    eng = engine.Engine()
    model = eng.create_model("HullWhite2F", {"a": 0.1, "b": 0.2, "sigma": 0.01, "eta": 0.012, "rho": -0.7,
                                             "r0": 0.03})
    swap = eng.create_product("IRSwap", {"notional": 1e6, "payment_times": PAY5, "accruals": ACC5})
    market = engine.MarketSnapshot(pillars=[1.0, 5.0], zero_rates=[0.03, 0.03], hazard_rate=0.02,
                                   recovery_rate=0.4)
    pricing = engine.PricingContext({"n_paths": 5000.0, "n_steps": 208.0, "seed": 7.0})
    gpu = engine.ExecutionContext({"backend": "gpu"})
    res = eng.price(swap, ["PV", "DV01", "ExpectedExposure", "PFE95", "UnilateralCVA"], model, market, pricing, gpu)
    ee, pfe = np.array(res["ExpectedExposure"].primary), np.array(res["PFE95"].primary)
    assert len(ee) == 5 and np.all(ee >= 0) and np.all(pfe >= ee)
    assert res["UnilateralCVA"].scalar > 0.0
    t, ee_c, pfe_c = ctx.exposure_profile_2f(*M2, 1e6, 0.0, True, 0.0, PAY5, ACC5, RESET5, 208, 5000, 7)
    assert np.array_equal(ee, ee_c) and np.array_equal(pfe, pfe_c)
    assert res["UnilateralCVA"].scalar == ctx.unilateral_cva_2f(*M2, t, ee_c, 0.02, 0.4)
    zero = engine.MarketSnapshot(pillars=[1.0, 5.0], zero_rates=[0.03, 0.03], hazard_rate=0.0, recovery_rate=0.4)
    assert abs(eng.price(swap, ["UnilateralCVA"], model, zero, pricing, gpu)["UnilateralCVA"].scalar) < 1e-9
    # the Greek measure bumps the two-factor parameters by name
    g = eng.price(swap, [("Greek", {"metric": "UnilateralCVA", "risk_factor": "model.eta"})], model, market,
                  pricing, gpu)["Greek"]
    assert g.has_scalar and np.isfinite(g.scalar) and g.bump_used == max(1e-2 * 0.012, 1e-4)
