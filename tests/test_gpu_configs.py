"""The benchmark workloads on their own generators, on the GPU, against the oracle.

Counter-based Philox makes path p's shocks independent of the number of paths, so the oracle checks
the very paths the GPU runs, at a path count it finishes in seconds; the full-size runs are checked
through properties that hold for any path count and against the oracle's fused run.

  config 2  1000 swaps (swap_portfolio seed 7), 360 monthly dates, exact scheme: shared host draws
            on 256 paths (per-path netting-set MtM, set EE / PFE95 / CVA, trade EE / CVA), the Philox
            stream on 256 paths, and the full 100k paths.
  config 3  batched_sets(64, seed 7), 120 quarterly dates: batched equals unbatched bit for bit for all
            64 sets; three sets against the oracle on their first paths.
  config 4  the 1000-swap book on 120 quarterly dates: the CVA adjoint against oracle central
            differences on 2048 shared-draw paths; PV pillar and model NPV sensitivities of every trade.
  config 5  bermudan_book(200, seed 7), quarterly to 30y: LSM policy, its application and the
            polynomial judge against the oracle on 4096 training / 2048 inference paths.

Tolerances: per-path MtM within 1e-12 of the legs' magnitude; EE / PFE / CVA within 1e-10 of the
value itself (floored at the row's largest value, see helpers.assert_value_rel).
"""
import numpy as np
import pytest

import oracle
from engine_quant_amd import eqhip
from engine_quant_amd.berm import BermudanBook
from engine_quant_amd.portfolios import (MODEL, batched_sets, bermudan_book, monthly_dates, quarterly_dates,
                                         swap_portfolio, zero_curve)
from helpers import assert_rel, assert_value_rel, leg_scale

pytestmark = pytest.mark.gpu

H, RR = 0.02, 0.4
EXACT, HOST, PHILOX = eqhip.DISC_EXACT, eqhip.RNG_HOST_DRAWS, eqhip.RNG_PHILOX


@pytest.fixture(scope="module")
def book():
    return swap_portfolio(1000, seed=7)


def _cva_leg(leg, dates):
    # the CVA of a trade whose exposure is at most 1e-6 of its legs is at most this much of the legs
    return leg * (1.0 - RR) * (1.0 - np.exp(-H * float(np.max(dates))))


def _check_book(got, ref, port, dates, what):
    leg = leg_scale(port, MODEL, dates)
    set_leg = np.array([leg[port.set_id == s].sum() for s in range(port.n_sets)])
    assert_value_rel(got["set_ee"], ref["set_ee"], leg=set_leg, what=f"{what} set EE")
    assert_value_rel(got["set_pfe"], ref["set_pfe"], leg=set_leg, what=f"{what} set PFE95")
    assert_value_rel(got["set_cva"], ref["set_cva"], leg=_cva_leg(set_leg, dates), what=f"{what} set CVA")
    if "trade_ee" in got:
        assert_value_rel(got["trade_ee"], ref["trade_ee"], leg=leg, what=f"{what} trade EE")
        assert_value_rel(got["trade_cva"], ref["trade_cva"], leg=_cva_leg(leg, dates), what=f"{what} trade CVA")


# ------------------------------------------------------------------------------------------- config 2
def test_config2_headline_book_shared_draws(ctx, book):
    dates = monthly_dates(30)
    n_paths, n_steps = 256, 360
    z = oracle.philox_normals(7, n_paths, n_steps)
    print(f"config 2: 1000 swaps x {n_paths} host-draw paths x 360 dates", flush=True)
    plan = ctx.plan(MODEL, book, dates, n_paths, n_steps, rng=HOST, disc=EXACT, want_trade_ee=True)
    got = plan.execute(z, H, RR, want_set_mtm=True)
    plan.close()
    ref = oracle.portfolio_exposure_fused(MODEL, book.as_dict(), dates, n_steps, z=z, scheme=1, hazard_rate=H,
                                          recovery_rate=RR, want_trade=True, want_set_mtm=True)
    set_leg = leg_scale(book, MODEL, dates).sum()
    assert_rel(got["set_mtm"], ref["set_mtm"], 1e-12, scale=set_leg, what="per-path netting-set MtM")
    _check_book(got, ref, book, dates, "config 2 shared draws")


def test_config2_headline_book_philox_stream(ctx, book):
    dates = monthly_dates(30)
    n_paths, n_steps = 256, 360
    got = ctx.exposure_run(MODEL, book, dates, n_paths, n_steps, seed=7, rng=PHILOX, disc=EXACT, hazard_rate=H,
                           recovery_rate=RR, want_trade_ee=True)
    ref = oracle.portfolio_exposure_fused(MODEL, book.as_dict(), dates, n_steps, n_paths=n_paths, seed=7, scheme=1,
                                          hazard_rate=H, recovery_rate=RR, want_trade=True)
    _check_book(got, ref, book, dates, "config 2 Philox")


def test_config2_full_size(ctx, book):
    """100k paths: properties of any path count, bitwise repeatability, and the oracle's fused run on
    the same Philox paths (schedule classes shared there, so equal to rounding)."""
    dates = monthly_dates(30)
    n_paths, n_steps = 100_000, 360
    print("config 2: 1000 swaps x 100k Philox paths x 360 dates", flush=True)
    plan = ctx.plan(MODEL, book, dates, n_paths, n_steps, seed=7, rng=PHILOX, disc=EXACT, want_trade_ee=True)
    got = plan.execute(None, H, RR)
    again = plan.execute(None, H, RR)
    plan.close()
    for k in got:
        assert np.array_equal(got[k], again[k]), k
    ee, pfe = got["set_ee"][0], got["set_pfe"][0]
    assert np.all(ee >= 0) and np.all(got["trade_ee"] >= 0) and np.all(got["trade_cva"] >= 0)
    # PFE95 >= EE is no law: late dates with under 5% of paths exposed have PFE95 = 0 < EE (the oracle
    # comparison below pins PFE95 itself)
    assert np.all(pfe >= 0)
    assert np.all(ee <= got["trade_ee"].sum(axis=0) * (1 + 1e-12))  # max(sum V, 0) <= sum max(V, 0)
    assert got["set_cva"][0] <= got["trade_cva"].sum() * (1 + 1e-12)
    print("config 2: oracle fused run on 100k paths", flush=True)
    ref = oracle.portfolio_exposure_fused(MODEL, book.as_dict(), dates, n_steps, n_paths=n_paths, seed=7, scheme=1,
                                          hazard_rate=H, recovery_rate=RR, want_trade=True, share_classes=True,
                                          n_threads=16)
    _check_book(got, ref, book, dates, "config 2 full size")


# ------------------------------------------------------------------------------------------- config 3
def test_config3_64_sets_batched_equal_unbatched(ctx):
    jobs = batched_sets(64, seed=7)
    dates = quarterly_dates(30)
    print(f"config 3: {len(jobs)} sets, {sum(p.n_trades for p, _ in jobs)} trades, "
          f"{sum(n for _, n in jobs)} paths", flush=True)
    batch = eqhip.Batch(ctx, MODEL, jobs, dates, len(dates), seed=7, disc=EXACT, want_trade_ee=True)
    got = batch.execute(H, RR)
    assert "chosen" in batch.timing()["tune_log"]
    batch.close()
    for j, (port, n_paths) in enumerate(jobs):
        solo = ctx.exposure_run(MODEL, port, dates, n_paths, len(dates), seed=7, rng=PHILOX, disc=EXACT,
                                hazard_rate=H, recovery_rate=RR, want_trade_ee=True)
        for key in ("set_ee", "set_pfe", "set_cva"):
            assert np.array_equal(got[j][key], solo[key]), (j, key)
        # per-trade sums follow each plan's path chunks: equal to rounding
        assert_value_rel(got[j]["trade_ee"], solo["trade_ee"], rel=1e-12, what=f"set {j} trade EE")


def test_config3_sets_against_oracle(ctx):
    jobs = batched_sets(64, seed=7)
    dates = quarterly_dates(30)
    sub = [(jobs[s][0], n) for s, n in ((0, 512), (1, 300), (2, 700))]
    batch = eqhip.Batch(ctx, MODEL, sub, dates, len(dates), seed=7, disc=EXACT, want_trade_ee=True)
    got = batch.execute(H, RR)
    batch.close()
    for j, (port, n_paths) in enumerate(sub):
        ref = oracle.portfolio_exposure_fused(MODEL, port.as_dict(), dates, len(dates), n_paths=n_paths, seed=7,
                                              scheme=1, hazard_rate=H, recovery_rate=RR, want_trade=True)
        _check_book(got[j], ref, port, dates, f"config 3 set {j}")


# ------------------------------------------------------------------------------------------- config 4
BUMPS = {"a": 1e-4, "b": 1e-4, "sigma": 1e-5, "r0": 1e-4}


def test_config4_adjoint_on_headline_book(ctx, book):
    dates = quarterly_dates(30)
    n_paths, n_steps = 2048, 120
    z = oracle.philox_normals(7, n_paths, n_steps)
    print(f"config 4: CVA adjoint, 1000 swaps x {n_paths} host-draw paths x 120 dates", flush=True)
    plan = ctx.plan(MODEL, book, dates, n_paths, n_steps, rng=HOST, disc=EXACT)
    cva, grad, check = plan.cva_adjoint(0, H, RR, host_draws=z)
    plan.close()

    def oracle_cva(model, h=H, rr=RR):
        return oracle.portfolio_exposure_fused(model, book.as_dict(), dates, n_steps, z=z, scheme=1, hazard_rate=h,
                                               recovery_rate=rr)

    base = oracle_cva(MODEL)
    assert_value_rel(cva, base["set_cva"][0], what="adjoint's CVA")
    assert_value_rel(check, base["set_cva"][0], what="adjoint's forward CVA")
    for i, name in enumerate(("a", "b", "sigma", "r0")):
        m_up, m_dn = list(MODEL), list(MODEL)
        m_up[i] += BUMPS[name]
        m_dn[i] -= BUMPS[name]
        fd = (oracle_cva(tuple(m_up))["set_cva"][0] - oracle_cva(tuple(m_dn))["set_cva"][0]) / (2 * BUMPS[name])
        assert abs(grad[name] - fd) <= 2e-3 * abs(fd) + 1e-9 * abs(cva), (name, grad[name], fd)
    # credit: exact derivatives of the CVA sum over the same EE profile
    ee = base["set_ee"][0]
    for i, name in enumerate(("hazard_rate", "recovery_rate")):
        e = 1e-6
        up, dn = [H, RR], [H, RR]
        up[i] += e
        dn[i] -= e
        fd = (oracle.unilateral_cva(*MODEL, dates, ee, *up) - oracle.unilateral_cva(*MODEL, dates, ee, *dn)) / (2 * e)
        assert abs(grad[name] - fd) <= 1e-6 * abs(fd), (name, grad[name], fd)


def test_config4_pv_sensitivities_of_every_trade(ctx, book):
    """Pillar deltas of all 1000 trades against the one-sided bucketed DV01 (measure.cpp:211-222) and, on
    every 25th trade, against oracle central differences; the model NPV Greeks (the reference's AAD,
    api.rs:394-415,519-552) against oracle central differences on every trade."""
    pillars, zeros = zero_curve(7)
    dz, npv, greeks = ctx.pv_sensitivities(pillars, zeros, book, MODEL)
    _, _, bk = ctx.pv_dv01(pillars, zeros, book, bump=1e-4, bucketed=True)
    off = book.flow_off
    N = np.abs(book.notional)
    assert np.all(np.abs(bk - 1e-4 * dz) <= 5e-3 * np.abs(bk) + 1e-6 * N[:, None])
    h = 1e-6
    for j in range(book.n_trades):
        pay, acc = book.pay[off[j]:off[j + 1]], book.acc[off[j]:off[j + 1]]
        n_j, k_j = book.notional[j], book.fixed_rate[j]
        if j % 25 == 0:
            for p in range(len(pillars)):
                up, dn = zeros.copy(), zeros.copy()
                up[p] += h
                dn[p] -= h
                fd = (oracle.npv_from_market(pillars, up, n_j, k_j, 0.0, pay, acc)
                      - oracle.npv_from_market(pillars, dn, n_j, k_j, 0.0, pay, acc)) / (2 * h)
                assert abs(dz[j, p] - fd) <= 1e-6 * N[j] * 30, (j, p, dz[j, p], fd)
        assert abs(npv[j] - oracle.irs_npv_reset(*MODEL, 0.0, n_j, k_j, 0.0, pay, acc)) <= 1e-9 * N[j]
        for i in range(4):
            mu, md = list(MODEL), list(MODEL)
            mu[i] += h
            md[i] -= h
            fd = (oracle.irs_npv_reset(*mu, 0.0, n_j, k_j, 0.0, pay, acc)
                  - oracle.irs_npv_reset(*md, 0.0, n_j, k_j, 0.0, pay, acc)) / (2 * h)
            assert abs(greeks[j, i] - fd) <= 1e-6 * abs(fd) + 1e-8 * N[j], (j, i, greeks[j, i], fd)


# ------------------------------------------------------------------------------------------- config 5
def test_config5_bermudan_book_policy_and_judge(ctx):
    bk = bermudan_book(200, seed=7)
    spy, hz = 4, 30
    n_tr, n_in = 4096, 2048
    zt = oracle.philox_normals(707, n_tr, spy * hz)
    zi = oracle.philox_normals(1477, n_in, spy * hz)
    print(f"config 5: 200 Bermudan swaptions, {n_tr} / {n_in} host-draw paths x 120 dates", flush=True)
    B = BermudanBook(ctx, MODEL, bk["notional"], bk["strike"], bk["first_ex"], bk["n_ex"], n_tr, n_in, spy, hz,
                     z_train=zt, z_infer=zi)
    pol = B.fit_policy()
    tau_t, y0_t = B.pathwise(0)
    tau_i, y0_i = B.pathwise(1)
    ee, coef = B.exposure_lsm(1)
    B.close()
    rt, lbt = oracle.berm_simulate(MODEL, 1.0 / spy, zt)
    ri, lbi = oracle.berm_simulate(MODEL, 1.0 / spy, zi)
    ref = oracle.berm_policy(MODEL, bk, spy, rt, lbt)
    assert np.array_equal(tau_t, ref["tau"]), int((tau_t != ref["tau"]).sum())
    assert np.allclose(y0_t, ref["y0"], rtol=1e-12, atol=1e-9)
    scale = np.abs(ref["ex_coef"]).max(axis=2, keepdims=True)
    assert np.all(np.abs(pol["ex_coef"] - ref["ex_coef"]) <= 1e-8 * np.maximum(scale, 1e-300) + 1e-12)
    assert_value_rel(pol["price_train"], ref["price"], rel=1e-11, what="in-sample prices")
    t_i, y_i = oracle.berm_apply(MODEL, bk, spy, ri, lbi, ref["ex_coef"])
    assert np.array_equal(tau_i, t_i), int((tau_i != t_i).sum())
    assert np.allclose(y0_i, y_i, rtol=1e-12, atol=1e-9)
    assert_value_rel(pol["price_infer"], y_i.mean(axis=1), rel=1e-11, what="out-of-sample prices")
    kmax = (bk["first_ex"] + bk["n_ex"] - 1) * spy
    for j in range(len(bk["notional"])):
        assert np.all(ee[j, kmax[j]:] == 0.0)
        ref_ee = np.zeros(kmax[j])
        for k in range(1, kmax[j] + 1):
            c, _, ref_ee[k - 1] = oracle.berm_judge(MODEL, spy, ri, lbi, tau_i[j], y0_i[j], k)
            assert np.allclose(coef[j, k - 1], c, rtol=1e-8, atol=1e-7 * max(1.0, np.abs(c).max())), (j, k)
        assert_value_rel(ee[j, :kmax[j]], ref_ee, rel=1e-9, what=f"judge EE of swaption {j}")
