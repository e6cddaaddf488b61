"""Per-trade PFE95 from portfolio plans (ee_pfe_per_trade, exposure.rs:151-159): the rank
ceil(0.95 n) - 1 order statistic of max(V_trade, 0) over the paths, at any date (between resets,
before a forward start, on grid index 0 with its single row, after maturity). Checked against the
oracle's per-trade restatement on shared host draws to 1e-10 of the value (floored at 1e-6 of the
legs' magnitude, the rounding scale of a value that is a difference of its legs), and against an exact
sort of the GPU's own per-path trade values."""
import ctypes

import numpy as np
import pytest

import oracle
from engine_quant_amd import eqhip
from engine_quant_amd.portfolios import MODEL2, monthly_dates, swap_portfolio
from helpers import MODEL, assert_rel, assert_value_rel, diverse_book, leg_scale, small_portfolio

pytestmark = pytest.mark.gpu

H, RR = 0.02, 0.4


def _rank(n):
    return min(max(int(np.ceil(0.95 * n)) - 1, 0), n - 1)


def _check_against_oracle(got, ref, port, dates):
    leg = leg_scale(port, MODEL, dates)
    assert_value_rel(got["trade_pfe"], ref["trade_pfe"], leg=leg, what="trade PFE95")
    assert_value_rel(got["trade_ee"], ref["trade_ee"], leg=leg, what="trade EE")


def test_small_book_against_oracle_and_own_sort(ctx):
    port = small_portfolio(n_trades=24, n_sets=3, seed=31, max_years=6, start=0.0)
    # forward starts on some trades
    port.start[::4] += 0.5
    for j in range(0, port.n_trades, 4):
        f0, f1 = port.flow_off[j], port.flow_off[j + 1]
        port.pay[f0:f1] += 0.5
    dates = np.concatenate([[0.0], np.arange(1, 8 * 4 + 1) / 4.0 + 1.0 / 24.0])  # grid 0, between resets
    n_paths, n_steps = 1237, 8 * 24
    z = np.random.default_rng(5).standard_normal((n_steps, n_paths))
    plan = ctx.plan(MODEL, port, dates, n_paths, n_steps, rng=eqhip.RNG_HOST_DRAWS, disc=eqhip.DISC_EXACT,
                    want_trade_ee=True, want_trade_pfe=True)
    got = plan.execute(z, H, RR, want_path=True)
    t = plan.timing()
    plan.close()
    assert t["trade_pfe_groups"] == 1 and t["trade_pfe_ms"] > 0.0
    ref = oracle.portfolio_exposure(MODEL, port.as_dict(), dates, n_steps, z, scheme=1, hazard_rate=H,
                                    recovery_rate=RR, want_trade=True)
    _check_against_oracle(got, ref, port, dates)
    # the GPU's own per-path values, sorted: the same order statistic to the rounding of the two forms
    leg = leg_scale(port, MODEL, dates)
    pm = got["path_mtm"]  # [date][trade][path]
    for k in range(len(dates)):
        n = 1 if k == 0 else n_paths
        srt = np.sort(np.maximum(pm[k, :, :n], 0.0), axis=1)
        assert_rel(got["trade_pfe"][:, k], srt[:, _rank(n)], 1e-12, scale=leg, what=f"date {k}")
    # matured trades have exactly zero PFE
    mat = np.array([port.pay[port.flow_off[j + 1] - 1] for j in range(port.n_trades)])
    assert np.all(got["trade_pfe"][mat[:, None] <= dates[None, :] + 1e-9] == 0.0)


def test_headline_book_on_the_philox_stream(ctx):
    """The 1000-swap book on 360 monthly dates, on the first 256 paths of the Philox stream."""
    port = swap_portfolio(1000, seed=7)
    dates = monthly_dates(30)
    n_paths, n_steps = 256, 360
    print(f"1000 swaps x {n_paths} Philox paths x 360 dates, per-trade PFE95", flush=True)
    plan = ctx.plan(MODEL, port, dates, n_paths, n_steps, seed=7, rng=eqhip.RNG_PHILOX, disc=eqhip.DISC_EXACT,
                    want_trade_ee=True, want_trade_pfe=True)
    got = plan.execute(None, H, RR)
    plan.close()
    z = oracle.philox_normals(7, n_paths, n_steps)
    ref = oracle.portfolio_exposure(MODEL, port.as_dict(), dates, n_steps, z, scheme=1, hazard_rate=H,
                                    recovery_rate=RR, want_trade=True)
    _check_against_oracle(got, ref, port, dates)


def test_book_in_parts_and_in_date_groups(ctx, monkeypatch):
    """A 500-schedule book valued in parts, with a class-leg buffer capped to a few dates, so that the
    valuation and the selects run in several date groups."""
    port = diverse_book(720, n_sets=2, seed=11)
    dates = np.arange(1, 12 * 10 + 1) / 12.0
    n_paths, n_steps = 600, 120
    z = oracle.philox_normals(5, n_paths, n_steps)
    n_cls = len({(port.start[j], tuple(port.pay[port.flow_off[j]:port.flow_off[j + 1]]),
                  tuple(port.acc[port.flow_off[j]:port.flow_off[j + 1]])) for j in range(port.n_trades)})
    monkeypatch.setenv("EQHIP_TRADE_PFE_BYTES", str(16 * n_paths * (n_cls + 40) * 7))
    plan = ctx.plan(MODEL, port, dates, n_paths, n_steps, rng=eqhip.RNG_HOST_DRAWS, disc=eqhip.DISC_EXACT,
                    want_trade_ee=True, want_trade_pfe=True)
    got = plan.execute(z, H, RR)
    t = plan.timing()
    plan.close()
    assert t["n_parts"] > 1 and t["trade_pfe_groups"] > 10, t
    ref = oracle.portfolio_exposure(MODEL, port.as_dict(), dates, n_steps, z, scheme=1, hazard_rate=H,
                                    recovery_rate=RR, want_trade=True)
    _check_against_oracle(got, ref, port, dates)


def test_set_results_do_not_change_and_variants(ctx):
    """A plan with per-trade PFE gives the same set results as one without; it values one path per lane
    (any trades per item), and a run without the trade_pfe output skips the class legs."""
    port = small_portfolio(n_trades=30, n_sets=2, seed=9)
    dates = np.arange(1, 21) / 4.0
    n_paths, n_steps = 3000, 20
    base = ctx.exposure_run(MODEL, port, dates, n_paths, n_steps, seed=3, disc=eqhip.DISC_EXACT, hazard_rate=H,
                            recovery_rate=RR, want_trade_ee=True)
    plan = ctx.plan(MODEL, port, dates, n_paths, n_steps, seed=3, disc=eqhip.DISC_EXACT, want_trade_ee=True,
                    want_trade_pfe=True)
    with pytest.raises(eqhip.EqhipError, match="one path per lane"):
        plan.set_variant(2, 4)
    runs = []
    for tpi in (2, 4, 8):
        plan.set_variant(1, tpi)
        runs.append(plan.execute(None, H, RR))
    plan.close()
    for got in runs:
        for key in ("set_ee", "set_pfe", "set_cva"):
            assert np.array_equal(got[key], base[key]), key
        assert np.array_equal(got["trade_pfe"], runs[0]["trade_pfe"])
    one = ctx.exposure_run(MODEL, port, dates, n_paths, n_steps, seed=3, disc=eqhip.DISC_EXACT, hazard_rate=H,
                           recovery_rate=RR, want_trade_pfe=True)
    assert np.array_equal(one["trade_pfe"], runs[0]["trade_pfe"])


def test_what_per_trade_pfe_refuses(ctx):
    port = small_portfolio(n_trades=4, n_sets=1, seed=2)
    with pytest.raises(eqhip.EqhipError, match="HullWhite1F plans only") as e:
        ctx.plan(MODEL2, port, [0.5, 1.0], 64, 4, want_trade_pfe=True)
    assert e.value.code == eqhip.E_INVALID
    plan = ctx.plan(MODEL, port, [0.5, 1.0], 64, 4)
    nd, nt = 2, port.n_trades
    ee, pfe, tp = np.zeros(nd), np.zeros(nd), np.zeros((nt, nd))
    out = eqhip.OutputsC(eqhip._dp(ee), eqhip._dp(pfe), None, None, None, None, None, eqhip._dp(tp))
    rc = eqhip.lib().eqhip_plan_execute(plan.handle, None, None, ctypes.byref(out))
    plan.close()
    assert rc == eqhip.E_INVALID and "EQHIP_PLAN_TRADE_PFE" in eqhip.last_error()
