"""The "Greek" measure and the Greek sweep (cpp/engine/src/greeks.cpp) on the GPU measures.

Pins: a bump-and-reval Greek equals the central difference of two priced runs; CVA is linear in
the recovery rate, so its recovery derivative is -CVA / (1 - R) up to rounding; its hazard-rate
derivative agrees with the analytic derivative of the CVA sum (exposure.rs:242-275) to O(h^2); the
exposure profile does not depend on the market curve (the model's own bond discounts it); the
pathwise route (CVA adjoint) agrees with common-random-number central differences.
"""
import numpy as np
import pytest

import engine_quant_amd as engine
import oracle

pytestmark = pytest.mark.gpu

A, B, SIG, R0 = 0.1, 0.03, 0.01, 0.02
PAY = [1.0, 2.0, 3.0, 4.0, 5.0]


def setup(fixed_rate=0.021, n_paths=4000, n_steps=104, hazard=0.02, recovery=0.4):
    eng = engine.Engine()
    model = eng.create_model("HullWhite1F", {"a": A, "b": B, "sigma": SIG, "r0": R0})
    params = {"notional": 1e6, "payment_times": PAY, "accruals": [1.0] * 5}
    if fixed_rate is not None:
        params["fixed_rate"] = fixed_rate
    product = eng.create_product("IRSwap", params)
    market = engine.MarketSnapshot(pillars=[1.0, 2.0, 5.0], zero_rates=[0.02, 0.022, 0.025], hazard_rate=hazard,
                                   recovery_rate=recovery)
    pricing = engine.PricingContext({"n_paths": float(n_paths), "n_steps": float(n_steps), "seed": 7.0})
    execution = engine.ExecutionContext({"backend": "gpu", "discretisation": "exact"})
    return eng, model, product, market, pricing, execution


def greek(eng, product, model, market, pricing, execution, **params):
    out = eng.price(product, [("Greek", params)], model, market, pricing, execution)
    return out["Greek"]


def cva(eng, product, model, market, pricing, execution):
    return eng.price(product, ["UnilateralCVA"], model, market, pricing, execution)["UnilateralCVA"].scalar


def test_bump_and_reval_is_the_central_difference_of_two_runs():
    eng, model, product, market, pricing, execution = setup()
    g = greek(eng, product, model, market, pricing, execution, metric="UnilateralCVA", risk_factor="model.sigma")
    h = max(1e-2 * SIG, 1e-4)  # greeks.cpp:160-164
    assert g.bump_used == h
    up = eng.create_model("HullWhite1F", {"a": A, "b": B, "sigma": SIG + h, "r0": R0})
    dn = eng.create_model("HullWhite1F", {"a": A, "b": B, "sigma": SIG - h, "r0": R0})
    fd = (cva(eng, product, up, market, pricing, execution) - cva(eng, product, dn, market, pricing, execution)) / (2 * h)
    assert g.has_scalar and g.scalar == fd
    # the profile part is the same stencil on EE (primary) and PFE95 (secondary)
    assert len(g.times) == 5 and len(g.primary) == 5 and len(g.secondary) == 5


def test_recovery_and_hazard_derivatives_match_the_cva_formula():
    eng, model, product, market, pricing, execution = setup()
    base = eng.price(product, ["UnilateralCVA"], model, market, pricing, execution)["UnilateralCVA"]
    c, t, ee = base.scalar, np.array(base.times), np.array(base.primary)
    g_r = greek(eng, product, model, market, pricing, execution, metric="UnilateralCVA",
                risk_factor="credit.recovery_rate")
    assert abs(g_r.scalar - (-c / (1 - 0.4))) <= 1e-9 * abs(c)
    g_h = greek(eng, product, model, market, pricing, execution, metric="UnilateralCVA",
                risk_factor="credit.hazard_rate")
    # d/dh sum_k (1-R) EE_k (S_{k-1} - S_k) P(0,t_k), S = exp(-h t); P from the model (hull_white.rs:50-80)
    h = 0.02
    bf = (1 - np.exp(-A * t)) / A
    p0 = np.exp((bf - t) * (A * A * B - SIG * SIG / 2) / (A * A) - SIG * SIG * bf * bf / (4 * A)) * np.exp(-bf * R0)
    tp = np.concatenate([[0.0], t[:-1]])
    dS = -tp * np.exp(-h * tp) + t * np.exp(-h * t)
    analytic = np.sum((1 - 0.4) * ee * dS * p0)
    assert abs(g_h.scalar - analytic) <= 1e-6 * abs(analytic)


def test_model_greeks_match_oracle_differences_on_the_same_paths():
    """The bump-and-reval model Greeks (default bumps, greeks.cpp:160-164) of UnilateralCVA and of the EE /
    PFE95 profile against central differences of the oracle's runs on the same Philox paths, exact scheme:
    the GPU's bumped runs and the oracle's see identical shocks (the reference's test of the same
    stencil, test_greeks.cpp:923-983, compares with its own bumped runs)."""
    n_paths, n_steps = 4000, 104
    eng, model, product, market, pricing, execution = setup(n_paths=n_paths, n_steps=n_steps)
    z = oracle.philox_normals(7, n_paths, n_steps)
    port = dict(set_id=[0], notional=[1e6], fixed_rate=[0.021], start=[0.0], flow_off=[0, 5], pay=PAY,
                acc=[1.0] * 5)
    times = [0.0, 1.0, 2.0, 3.0, 4.0]
    base = dict(a=A, b=B, sigma=SIG, r0=R0)
    for name in ("a", "b", "sigma", "r0"):
        g = greek(eng, product, model, market, pricing, execution, metric="UnilateralCVA",
                  risk_factor=f"model.{name}")
        h = g.bump_used
        assert h == max(1e-2 * abs(base[name]), 1e-4)
        runs = []
        for sgn in (1.0, -1.0):
            m = dict(base)
            m[name] += sgn * h
            runs.append(oracle.portfolio_exposure((m["a"], m["b"], m["sigma"], m["r0"]), port, times, n_steps, z,
                                                  scheme=1, hazard_rate=0.02, recovery_rate=0.4))
        fd = (runs[0]["set_cva"][0] - runs[1]["set_cva"][0]) / (2 * h)
        assert abs(g.scalar - fd) <= 1e-8 * abs(fd), (name, g.scalar, fd)
        ee_fd = (runs[0]["set_ee"][0] - runs[1]["set_ee"][0]) / (2 * h)
        pfe_fd = (runs[0]["set_pfe"][0] - runs[1]["set_pfe"][0]) / (2 * h)
        assert np.all(np.abs(np.array(g.primary) - ee_fd) <= 1e-8 * np.abs(ee_fd).max()), name
        assert np.all(np.abs(np.array(g.secondary) - pfe_fd) <= 1e-8 * np.abs(pfe_fd).max()), name


def test_pv_at_the_pricing_date_and_its_theta():
    """PV and DV01 honour PricingContext::pricing_date (measure.cpp:146-200,366-380,425-445): every
    discount factor over DF(pricing_date), the bucketed ladder at t = 0; a pricing date after the start is
    refused with the reference's message; time.theta of PV is PV(t + 1/365) - PV(t), not divided."""
    eng = engine.Engine()
    model = eng.create_model("HullWhite1F", {"a": A, "b": B, "sigma": SIG, "r0": R0})
    pay = [2.0, 3.0, 4.0, 5.0, 6.0]
    product = eng.create_product("IRSwap", {"notional": 1e6, "fixed_rate": 0.021, "start": 1.0, "payment_times": pay,
                                            "accruals": [1.0] * 5})
    pillars, zeros = [1.0, 2.0, 5.0], [0.02, 0.022, 0.025]
    market = engine.MarketSnapshot(pillars=pillars, zero_rates=zeros, hazard_rate=0.02, recovery_rate=0.4)
    execution = engine.ExecutionContext({"backend": "gpu"})
    args = (1e6, 0.021, 1.0, pay, [1.0] * 5)
    for t in (0.0, 0.4, 1.0):
        pricing = engine.PricingContext({"pricing_date": t, "n_paths": 100.0, "n_steps": 10.0, "seed": 7.0})
        out = eng.price(product, ["PV", "DV01"], model, market, pricing, execution)
        assert abs(out["PV"].scalar - oracle.npv_from_market_at(pillars, zeros, *args, t)) <= 1e-9
        assert abs(out["DV01"].scalar - oracle.dv01_at(pillars, zeros, *args, t)) <= 1e-9
    pricing = engine.PricingContext({"pricing_date": 0.4, "n_paths": 100.0, "n_steps": 10.0, "seed": 7.0})
    g = greek(eng, product, model, market, pricing, execution, metric="PV", risk_factor="time.theta")
    ref = (oracle.npv_from_market_at(pillars, zeros, *args, 0.4 + 1.0 / 365.0)
           - oracle.npv_from_market_at(pillars, zeros, *args, 0.4))
    assert g.bump_used == 1.0 / 365.0 and abs(g.scalar - ref) <= 1e-9
    late = engine.PricingContext({"pricing_date": 1.5, "n_paths": 100.0, "n_steps": 10.0, "seed": 7.0})
    with pytest.raises(ValueError, match="posterior al start"):
        eng.price(product, ["PV"], model, market, late, execution)
    # the bucketed ladder stays at t = 0, as the reference's does, so it is served after the start too
    bk = eng.price(product, [("DV01", {"bucketed": True})], model, market, late, execution)["DV01"]
    ref_bk = oracle.dv01_bucketed(pillars, zeros, *args)
    assert np.all(np.abs(np.array(bk.primary) - ref_bk) <= 1e-9)
    rep = eng.all_greeks(product, "PV", model, market, pricing, execution)
    assert "time.theta" in [x.risk_factor for x in rep.greeks]


def test_exposure_does_not_depend_on_the_market_curve():
    eng, model, product, market, pricing, execution = setup()
    for rf in ("curve.parallel", "curve.pillar:1"):
        g = greek(eng, product, model, market, pricing, execution, metric="ExposureProfile", risk_factor=rf)
        assert np.all(np.array(g.primary) == 0.0) and np.all(np.array(g.secondary) == 0.0)


def test_pv_pillar_greek_is_the_central_version_of_the_bucketed_dv01():
    eng, model, product, market, pricing, execution = setup()
    g = greek(eng, product, model, market, pricing, execution, metric="PV", risk_factor="curve.pillar:2")
    dv = eng.price(product, [("DV01", {"bucketed": True})], model, market, pricing, execution)["DV01"]
    one_sided = dv.primary[2] / 1e-4  # DV01 is bumped - base, not divided (measure.cpp:195-222)
    assert abs(g.scalar - one_sided) <= 1e-3 * abs(one_sided)


def test_pathwise_cva_greeks_match_crn_central_differences():
    eng, model, product, market, pricing, execution = setup(n_paths=20000)
    for name, tol in (("a", 2e-3), ("b", 2e-3), ("r0", 2e-3), ("sigma", 2e-3)):
        pw = greek(eng, product, model, market, pricing, execution, metric="UnilateralCVA",
                   risk_factor=f"model.{name}", method="pathwise")
        fd = greek(eng, product, model, market, pricing, execution, metric="UnilateralCVA",
                   risk_factor=f"model.{name}", bump=1e-5)
        assert pw.bump_used is None
        assert abs(pw.scalar - fd.scalar) <= tol * abs(fd.scalar), (name, pw.scalar, fd.scalar)


def test_pathwise_refuses_what_it_cannot_serve():
    eng, model, product, market, pricing, execution = setup(fixed_rate=None)  # par rate: K moves with the model
    with pytest.raises(ValueError, match="fixed_rate"):
        greek(eng, product, model, market, pricing, execution, metric="UnilateralCVA", risk_factor="model.a",
              method="pathwise")
    eng, model, product, market, pricing, execution = setup()
    with pytest.raises(ValueError, match="pathwise"):
        greek(eng, product, model, market, pricing, execution, metric="ExposureProfile", risk_factor="model.a",
              method="pathwise")
    with pytest.raises(ValueError, match="ModelParameter"):
        greek(eng, product, model, market, pricing, execution, metric="UnilateralCVA",
              risk_factor="credit.hazard_rate", method="pathwise")


def test_second_order_and_cross_stencils():
    eng, model, product, market, pricing, execution = setup()
    g2 = greek(eng, product, model, market, pricing, execution, metric="UnilateralCVA", risk_factor="model.r0",
               order=2.0)
    h = max(1e-2 * R0, 1e-4)  # greeks.cpp:160-164
    assert g2.bump_used == h
    ups = [eng.create_model("HullWhite1F", {"a": A, "b": B, "sigma": SIG, "r0": R0 + s * h}) for s in (1, 0, -1)]
    v = [cva(eng, product, m, market, pricing, execution) for m in ups]
    assert g2.scalar == (v[0] - 2 * v[1] + v[2]) / (h * h)
    # CVA is linear in the recovery rate: its cross derivative with any factor is -dCVA/dx / (1 - R)
    gx = greek(eng, product, model, market, pricing, execution, metric="UnilateralCVA", risk_factor="model.b",
               cross_factor="credit.recovery_rate")
    gb = greek(eng, product, model, market, pricing, execution, metric="UnilateralCVA", risk_factor="model.b")
    assert abs(gx.scalar - (-gb.scalar / (1 - 0.4))) <= 1e-6 * abs(gb.scalar)


def test_all_greeks_sweep():
    eng, model, product, market, pricing, execution = setup()
    rep = eng.all_greeks(product, "UnilateralCVA", model, market, pricing, execution, include_curve_buckets=True)
    names = [g.risk_factor for g in rep.greeks]
    assert names[:4] == ["model.a", "model.b", "model.r0", "model.sigma"]
    assert "curve.parallel" in names and "curve.pillar:2" in names
    assert "credit.hazard_rate" in names and "credit.recovery_rate" in names
    assert all(g.method_used == "bump_and_reval" for g in rep.greeks)
    assert any(s.startswith("time.theta") for s in rep.skipped)
    # the CVA discounts with the model bond, so its curve Greeks are exactly zero
    for g in rep.greeks:
        if g.risk_factor.startswith("curve."):
            assert g.value == 0.0
