"""Portfolios too diverse for one block: more schedule classes and cash-flow dates than a block's LDS
holds, or more per-trade work items than a block has slots. The plan values them in parts (groups of
classes, a class with many trades in chunks of its trades) over the same paths, one after another into
the same netting-set rows, and the results are checked against the oracle on shared host draws:
per-path netting-set values to 1e-12 of the legs' magnitude, EE / PFE95 / CVA to 1e-10 of the value."""
import numpy as np
import pytest

import oracle
from engine_quant_amd import eqhip
from engine_quant_amd.portfolios import MODEL2
from helpers import MODEL, assert_rel, assert_value_rel, diverse_book, leg_scale

pytestmark = pytest.mark.gpu

H, RR = 0.02, 0.4


def _check(got, ref, port, dates):
    leg = leg_scale(port, MODEL, dates)
    set_leg = np.array([leg[port.set_id == s].sum() for s in range(port.n_sets)])
    assert_rel(got["set_mtm"], ref["set_mtm"], 1e-12, scale=set_leg[None, :, None], what="per-path set MtM")
    assert_value_rel(got["set_ee"], ref["set_ee"], leg=set_leg, what="set EE")
    assert_value_rel(got["set_pfe"], ref["set_pfe"], leg=set_leg, what="set PFE95")
    assert_value_rel(got["set_cva"], ref["set_cva"], leg=set_leg, what="set CVA")
    if "trade_ee" in got:
        assert_value_rel(got["trade_ee"], ref["trade_ee"], leg=leg, what="trade EE")
        assert_value_rel(got["trade_cva"], ref["trade_cva"], leg=leg, what="trade CVA")


def test_500_schedule_book_in_parts(ctx):
    """Over 500 distinct schedules: forward starts, trades already running (a short first stub), spot
    starts; monthly dates, so that most dates fall between resets."""
    port = diverse_book(720, n_sets=2, seed=11)
    n_sched = len({(port.start[j], tuple(port.pay[port.flow_off[j]:port.flow_off[j + 1]]))
                   for j in range(port.n_trades)})
    assert n_sched >= 500
    dates = np.arange(1, 12 * 10 + 1) / 12.0
    n_paths, n_steps = 600, 120
    z = oracle.philox_normals(5, n_paths, n_steps)
    print(f"{n_sched} schedules, {port.n_trades} trades x {n_paths} paths x {len(dates)} dates", flush=True)
    plan = ctx.plan(MODEL, port, dates, n_paths, n_steps, rng=eqhip.RNG_HOST_DRAWS, disc=eqhip.DISC_EXACT,
                    want_trade_ee=True)
    got = plan.execute(z, H, RR, want_set_mtm=True)
    n_parts = plan.timing()["n_parts"]
    plan.close()
    assert n_parts > 1
    ref = oracle.portfolio_exposure(MODEL, port.as_dict(), dates, n_steps, z, scheme=1, hazard_rate=H,
                                    recovery_rate=RR, want_trade=True, want_path=True)
    _check(got, ref, port, dates)


def test_distinct_start_book_in_parts_philox(ctx):
    """200 schedules with distinct starts (600 cash-flow dates), on the Philox stream, a date on grid
    index 0 and every date between resets."""
    n = 200
    start = np.arange(n) / 1000.0
    pay = np.stack([start + 1.0, start + 2.0], axis=1).ravel()
    port = eqhip.Portfolio(set_id=(np.arange(n) % 3).astype(np.int32),
                           notional=np.where(np.arange(n) % 2 == 0, 1e6, -7e5), fixed_rate=np.full(n, 0.02),
                           start=start, flow_off=np.arange(0, 2 * n + 1, 2, dtype=np.int32), pay=pay,
                           acc=np.ones(2 * n))
    dates = np.array([0.0, 0.5, 1.5])
    n_paths, n_steps = 3000, 6
    plan = ctx.plan(MODEL, port, dates, n_paths, n_steps, seed=4, rng=eqhip.RNG_PHILOX, disc=eqhip.DISC_EULER,
                    want_trade_ee=True)
    got = plan.execute(None, H, RR, want_set_mtm=True)
    assert plan.timing()["n_parts"] > 1
    plan.close()
    z = oracle.philox_normals(4, n_paths, n_steps)
    ref = oracle.portfolio_exposure(MODEL, port.as_dict(), dates, n_steps, z, scheme=0, hazard_rate=H,
                                    recovery_rate=RR, want_trade=True, want_path=True)
    _check(got, ref, port, dates)


def test_one_class_with_more_work_items_than_slots(ctx):
    """8193 same-signed trades of one schedule: more work items of eight trades than a block has threads,
    so the class's trades are valued in chunks by separate parts."""
    n = 8 * 1024 + 1
    rng = np.random.default_rng(2)
    port = eqhip.Portfolio(set_id=np.zeros(n, np.int32), notional=rng.uniform(1e5, 1e6, n),
                           fixed_rate=rng.uniform(0.01, 0.03, n), start=np.zeros(n),
                           flow_off=np.arange(0, 2 * n + 1, 2, dtype=np.int32), pay=np.tile([1.0, 2.0], n),
                           acc=np.ones(2 * n))
    dates = np.array([0.5, 1.0, 1.5])
    n_paths, n_steps = 256, 6
    z = np.random.default_rng(3).standard_normal((n_steps, n_paths))
    plan = ctx.plan(MODEL, port, dates, n_paths, n_steps, rng=eqhip.RNG_HOST_DRAWS, want_trade_ee=True)
    got = plan.execute(z, H, RR, want_set_mtm=True)
    assert plan.timing()["n_parts"] > 1
    plan.close()
    ref = oracle.portfolio_exposure(MODEL, port.as_dict(), dates, n_steps, z, hazard_rate=H, recovery_rate=RR,
                                    want_trade=True, want_path=True)
    _check(got, ref, port, dates)


def test_cube_and_adjoint_of_a_plan_in_parts(ctx):
    port = diverse_book(300, n_sets=1, seed=5, max_years=4)
    dates = np.arange(1, 17) / 4.0
    n_paths, n_steps = 2000, 48
    z = oracle.philox_normals(8, n_paths, n_steps)
    plan = ctx.plan(MODEL, port, dates, n_paths, n_steps, rng=eqhip.RNG_HOST_DRAWS, disc=eqhip.DISC_EXACT,
                    keep_cube=True)
    got = plan.execute(z, H, RR)
    assert plan.timing()["n_parts"] > 1
    cube, _ = plan.exposure_cube(0)
    assert_value_rel(np.maximum(cube, 0.0).mean(axis=1), got["set_ee"][0], rel=1e-12, what="EE of the cube")
    cva, grad, check = plan.cva_adjoint(0, H, RR, host_draws=z)
    plan.close()
    base = oracle.portfolio_exposure_fused(MODEL, port.as_dict(), dates, n_steps, z=z, scheme=1, hazard_rate=H,
                                           recovery_rate=RR, want_set_mtm=True)
    assert_rel(cube, base["set_mtm"][:, 0, :], 1e-12, scale=leg_scale(port, MODEL, dates).sum(), what="cube")
    assert_value_rel(cva, base["set_cva"][0], what="adjoint's CVA")
    assert_value_rel(check, base["set_cva"][0], what="adjoint's forward CVA")
    for i, (name, h) in enumerate((("a", 1e-4), ("b", 1e-4), ("sigma", 1e-5), ("r0", 1e-4))):
        up, dn = list(MODEL), list(MODEL)
        up[i] += h
        dn[i] -= h
        fd = (oracle.portfolio_exposure_fused(tuple(up), port.as_dict(), dates, n_steps, z=z, scheme=1, hazard_rate=H,
                                              recovery_rate=RR)["set_cva"][0]
              - oracle.portfolio_exposure_fused(tuple(dn), port.as_dict(), dates, n_steps, z=z, scheme=1,
                                                hazard_rate=H, recovery_rate=RR)["set_cva"][0]) / (2 * h)
        assert abs(grad[name] - fd) <= 2e-3 * abs(fd) + 1e-9 * abs(cva), (name, grad[name], fd)


def test_what_parts_cannot_serve(ctx):
    # one schedule with more cash-flow dates than a block's LDS holds, with per-trade EE (without it,
    # path mode values the plan)
    pay = np.arange(1, 401) / 12.0
    port = eqhip.Portfolio(set_id=np.zeros(1, np.int32), notional=np.full(1, 1e6), fixed_rate=np.full(1, 0.02),
                           start=np.zeros(1), flow_off=np.array([0, 400], np.int32), pay=pay, acc=np.full(400, 1 / 12))
    with pytest.raises(eqhip.EqhipError, match="more LDS than a block has") as e:
        ctx.plan(MODEL, port, [0.5, 1.0], 64, 4, want_trade_ee=True)
    assert e.value.code == eqhip.E_CAPACITY
    # G2++ plans are not valued in parts (this book has more schedule families than path mode holds)
    port = diverse_book(600, n_sets=1, seed=1)
    with pytest.raises(eqhip.EqhipError, match="too diverse") as e:
        ctx.plan(MODEL2, port, np.arange(1, 13) / 12.0, 64, 12)
    assert e.value.code == eqhip.E_CAPACITY
