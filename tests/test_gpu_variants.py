"""Valuation variants (paths per lane 1 / 2 / 4, trades per work item 2 / 4 / 8): every variant that fits a
plan writes the same netting-set values bit for bit (class legs and set sums are formed with explicit
fused multiply-adds in a fixed order), so set EE / PFE95 / CVA do not depend on the variant; per-trade EE
sums over path slices that follow the tile width, so it agrees to rounding. Each variant is also checked
against the oracle on shared host draws, with path counts that leave a partial last tile, a date on grid
index 0, dates between resets and forward starts."""
import numpy as np
import pytest

import oracle
from engine_quant_amd import eqhip
from engine_quant_amd.portfolios import MODEL2, monthly_dates, swap_portfolio
from helpers import MODEL, assert_rel, assert_value_rel, leg_scale, small_portfolio

pytestmark = pytest.mark.gpu

VARIANTS = [(1, 2), (1, 4), (1, 8), (2, 2), (2, 4), (2, 8), (4, 2)]


def _run_variants(ctx, model, port, dates, n_paths, n_steps, z, disc, want_path=False):
    out = {}
    for ppl, tpi in VARIANTS:
        plan = ctx.plan(model, port, dates, n_paths, n_steps, rng=eqhip.RNG_HOST_DRAWS, disc=disc, want_trade_ee=True)
        try:
            plan.set_variant(ppl, tpi)
        except eqhip.EqhipError as e:
            assert e.code == eqhip.E_CAPACITY
            plan.close()
            continue
        print(f"variant paths/lane={ppl} trades/item={tpi}", flush=True)
        out[(ppl, tpi)] = plan.execute(z, 0.02, 0.4, want_path=want_path, want_set_mtm=True)
        t = plan.timing()
        assert (t["paths_per_tile"], t["trades_per_item"]) == (64 * ppl, tpi)
        plan.close()
    return out


@pytest.mark.parametrize("n_paths", [1237, 4000])
def test_variants_agree_and_match_the_oracle(ctx, n_paths):
    port = small_portfolio(n_trades=30, n_sets=3, seed=12)
    dates = np.concatenate([[0.0], np.arange(1, 12 * 5 + 1) / 12.0 + 1.0 / 48.0])  # grid 0, between resets
    n_steps = 241  # dt = 1/48: every reset and date on the grid
    z = np.random.default_rng(6).standard_normal((n_steps, n_paths))
    runs = _run_variants(ctx, MODEL, port, dates, n_paths, n_steps, z, eqhip.DISC_EXACT, want_path=(n_paths == 1237))
    assert {(1, 2), (1, 4), (1, 8), (2, 2), (2, 4), (2, 8)} <= set(runs)
    base = runs[(1, 4)]
    for v, got in runs.items():
        for key in ("set_ee", "set_pfe", "set_cva", "set_mtm"):
            assert np.array_equal(got[key], base[key]), (v, key)
        assert_value_rel(got["trade_ee"], base["trade_ee"], rel=1e-12, what=f"{v} trade EE")
        if "path_mtm" in got:
            assert np.array_equal(got["path_mtm"], base["path_mtm"]), v
    ref = oracle.portfolio_exposure(MODEL, port.as_dict(), dates, n_steps, z, scheme=1, hazard_rate=0.02,
                                    recovery_rate=0.4, want_trade=True)
    leg = leg_scale(port, MODEL, dates)
    set_leg = np.array([leg[port.set_id == s].sum() for s in range(port.n_sets)])
    for v, got in runs.items():
        assert_value_rel(got["set_ee"], ref["set_ee"], leg=set_leg, what=f"{v} set EE")
        assert_value_rel(got["set_pfe"], ref["set_pfe"], leg=set_leg, what=f"{v} set PFE")
        assert_value_rel(got["trade_ee"], ref["trade_ee"], leg=leg, what=f"{v} trade EE")


def test_four_paths_per_lane_on_a_small_book(ctx):
    """A book small enough for four paths per lane, with more netting sets than the split limit (one wave
    per set sums the same chunks as the split path): every variant against the oracle."""
    port = small_portfolio(n_trades=12, n_sets=10, seed=21, max_years=3)
    dates = np.arange(1, 13) / 4.0
    n_paths, n_steps = 1500, 12
    z = np.random.default_rng(8).standard_normal((n_steps, n_paths))
    runs = _run_variants(ctx, MODEL, port, dates, n_paths, n_steps, z, eqhip.DISC_EULER)
    assert set(runs) == set(VARIANTS)
    ref = oracle.portfolio_exposure(MODEL, port.as_dict(), dates, n_steps, z, scheme=0, hazard_rate=0.02,
                                    recovery_rate=0.4, want_trade=True)
    leg = leg_scale(port, MODEL, dates)
    set_leg = np.array([leg[port.set_id == s].sum() for s in range(port.n_sets)])
    for v, got in runs.items():
        for key in ("set_ee", "set_pfe", "set_cva", "set_mtm"):
            assert np.array_equal(got[key], runs[(1, 4)][key]), (v, key)
        assert_value_rel(got["set_ee"], ref["set_ee"], leg=set_leg, what=f"{v} set EE")
        assert_value_rel(got["trade_ee"], ref["trade_ee"], leg=leg, what=f"{v} trade EE")


def test_headline_book_variants_that_fit(ctx):
    """The 1000-swap book on 360 monthly dates: its LDS leaves one path per lane; every trades-per-item
    choice gives the same set values."""
    port = swap_portfolio(1000, seed=7)
    dates = monthly_dates(30)
    n_paths, n_steps = 384, 360
    z = oracle.philox_normals(7, n_paths, n_steps)
    runs = _run_variants(ctx, MODEL, port, dates, n_paths, n_steps, z, eqhip.DISC_EXACT)
    assert set(runs) == {(1, 2), (1, 4), (1, 8)}
    for v, got in runs.items():
        for key in ("set_ee", "set_pfe", "set_cva", "set_mtm"):
            assert np.array_equal(got[key], runs[(1, 4)][key]), (v, key)
        assert_value_rel(got["trade_ee"], runs[(1, 4)]["trade_ee"], rel=1e-12, what=f"{v} trade EE")


def test_two_factor_plans_take_one_path_per_lane(ctx):
    port = small_portfolio(n_trades=10, n_sets=2, seed=3)
    dates = np.arange(1, 25) / 4.0
    n_paths, n_steps = 700, 24
    z = oracle.philox_normals2(9, n_paths, n_steps)
    runs = _run_variants(ctx, MODEL2, port, dates, n_paths, n_steps, z, eqhip.DISC_EXACT)
    assert set(runs) == {(1, 2), (1, 4), (1, 8)}
    ref = oracle.portfolio_exposure(MODEL2, port.as_dict(), dates, n_steps, z, scheme=1, hazard_rate=0.02,
                                    recovery_rate=0.4, want_trade=True, want_path=True)
    leg = leg_scale(port, MODEL, dates)
    set_leg = np.array([leg[port.set_id == s].sum() for s in range(port.n_sets)])
    for v, got in runs.items():
        assert np.array_equal(got["set_mtm"], runs[(1, 4)]["set_mtm"]), v
        assert_rel(got["set_mtm"], ref["set_mtm"], 1e-12, scale=set_leg[None, :, None], what=f"{v} set MtM")
        assert_value_rel(got["set_ee"], ref["set_ee"], leg=set_leg, what=f"{v} set EE")


def test_plan_autotune_and_batch_tune_log(ctx):
    port = small_portfolio(n_trades=20, n_sets=2, seed=4)
    dates = np.arange(1, 41) / 4.0
    plan = ctx.plan(MODEL, port, dates, 20000, 40, seed=3, rng=eqhip.RNG_PHILOX, disc=eqhip.DISC_EXACT,
                    want_trade_ee=True, autotune=True)
    tuned = plan.execute(None, 0.02, 0.4)
    again = plan.execute(None, 0.02, 0.4)
    plan.close()
    fixed = ctx.exposure_run(MODEL, port, dates, 20000, 40, seed=3, rng=eqhip.RNG_PHILOX, disc=eqhip.DISC_EXACT,
                             hazard_rate=0.02, recovery_rate=0.4, want_trade_ee=True)
    for key in ("set_ee", "set_pfe", "set_cva"):
        assert np.array_equal(tuned[key], fixed[key]) and np.array_equal(again[key], fixed[key]), key
    jobs = [(small_portfolio(n_trades=n, n_sets=1, seed=n), p) for n, p in ((5, 3000), (40, 9000), (120, 5000))]
    batch = eqhip.Batch(ctx, MODEL, jobs, dates, 40, seed=3, disc=eqhip.DISC_EXACT, want_trade_ee=True)
    batch.execute(0.02, 0.4)
    log = batch.timing()["tune_log"]
    batch.close()
    for ppl, tpi in VARIANTS:
        if (ppl, tpi) != (1, 4):
            assert f"paths/lane<={ppl} trades/item={tpi}" in log, (ppl, tpi, log)
    assert "chosen" in log
