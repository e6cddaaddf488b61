"""CPU-only checks: the C-ABI library loads and exports every declared symbol, the host mirror of
the reference API validates its inputs like the reference does, and requesting the GPU without a
device fails loudly instead of running anything on the CPU."""
import os
import re

import numpy as np
import pytest

import engine_quant_amd as engine
from engine_quant_amd import eqhip
from engine_quant_amd.portfolios import batched_sets, monthly_dates, swap_portfolio, zero_curve

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_every_declared_symbol():
    header = open(os.path.join(ROOT, "include", "eqhip.h")).read()
    declared = set(re.findall(r"^(?:int|size_t)\s+(eqhip_\w+)\(", header, flags=re.M))
    assert declared == set(eqhip.EXPORTED)
    lib = eqhip.lib()
    for name in declared:
        assert hasattr(lib, name), name


def test_gpu_request_without_device_fails_loudly():
    if eqhip.is_gpu_backend_available():
        pytest.skip("a GPU is visible")
    with pytest.raises(eqhip.EqhipError) as e:
        eqhip.Context(0)
    assert e.value.code == eqhip.E_NO_DEVICE


def test_host_contexts_validate_like_the_reference():
    with pytest.raises(ValueError):
        engine.ExecutionContext({"backend": "tpu"})
    with pytest.raises(ValueError):
        engine.ExecutionContext({"backend": "gpu", "precision": "fp32"})
    with pytest.raises(ValueError):
        engine.PricingContext({"n_paths": 0.0, "n_steps": 10.0, "seed": 1.0})
    with pytest.raises(ValueError):
        engine.MarketSnapshot(pillars=[2.0, 1.0], zero_rates=[0.02, 0.02])
    eng = engine.Engine()
    with pytest.raises(KeyError):
        eng.create_model("NoExiste", {})
    with pytest.raises(ValueError):
        eng.create_product("IRSwap", {"notional": 1.0, "payment_times": [1.0, 2.0], "accruals": [1.0]})
    assert sorted(eng.list_measures()) == ["DV01", "ExpectedExposure", "ExposureProfile", "Greek", "PFE95", "PV",
                                           "UnilateralCVA"]
    # model.cpp:39-71: both short-rate models, every parameter required
    assert sorted(eng.list_models()) == ["HullWhite1F", "HullWhite2F"]
    m2 = eng.create_model("HullWhite2F", {"a": 0.1, "b": 0.2, "sigma": 0.01, "eta": 0.012, "rho": -0.7, "r0": 0.03})
    assert m2.type_name == "HullWhite2F"
    with pytest.raises((KeyError, ValueError)):
        eng.create_model("HullWhite2F", {"a": 0.1, "b": 0.2, "sigma": 0.01, "rho": -0.7, "r0": 0.03})


def test_cpu_backend_is_the_reference_engine_not_this_library():
    eng = engine.Engine()
    model = eng.create_model("HullWhite1F", {"a": 0.1, "b": 0.03, "sigma": 0.01, "r0": 0.02})
    product = eng.create_product("IRSwap", {"notional": 1e6, "payment_times": [1.0, 2.0], "accruals": [1.0, 1.0]})
    market = engine.MarketSnapshot(pillars=[1.0], zero_rates=[0.02])
    pricing = engine.PricingContext({"n_paths": 10.0, "n_steps": 10.0, "seed": 1.0})
    with pytest.raises(ValueError, match="engine-core"):
        eng.price(product, ["ExpectedExposure"], model, market, pricing, engine.ExecutionContext({"backend": "cpu"}))
    with pytest.raises(ValueError, match="medida desconocida"):
        eng.price(product, ["NoSuchMeasure"], model, market, pricing, engine.ExecutionContext({"backend": "gpu"}))


def test_price_many_reports_the_error_a_serial_loop_would():
    """price_many runs its calendar groups over several GPU streams, largest first, and still reports
    the error of the first failing group in calendar order: here the first group's (a trade without
    fixed_rate, refused before any device work), though that group is claimed last and the others
    fail on the missing device (or succeed on a GPU) first. Twice: the workers outlive a failed call."""
    eng = engine.Engine()
    model = eng.create_model("HullWhite1F", {"a": 0.1, "b": 0.03, "sigma": 0.01, "r0": 0.02})
    par = eng.create_product("IRSwap", {"notional": 1e6, "payment_times": [1.0, 2.0], "accruals": [1.0, 1.0]})
    longer = [eng.create_product("IRSwap", {"notional": 1e6, "fixed_rate": 0.02,
                                            "payment_times": [float(t) for t in range(1, n + 1)],
                                            "accruals": [1.0] * n}) for n in (10, 20, 30)]
    market = engine.MarketSnapshot(pillars=[1.0], zero_rates=[0.02])
    pricing = engine.PricingContext({"n_paths": 64.0, "n_steps": 10.0, "seed": 1.0})
    for _ in range(2):
        with pytest.raises(ValueError, match="fixed_rate"):
            eng.price_many([par] + longer, ["PV"], model, market, pricing, engine.ExecutionContext({"backend": "gpu"}))


def test_synthetic_portfolio_generator_shape():
    pillars, zeros = zero_curve(7)
    assert len(pillars) == 20 and pillars[0] == 0.25 and pillars[-1] == 30.0
    port = swap_portfolio(1000, seed=7, pillars=pillars, zeros=zeros)
    assert port.n_trades == 1000 and port.n_sets == 1
    mats = port.pay[port.flow_off[1:] - 1]
    assert mats.min() >= 1.0 and mats.max() <= 30.0
    assert 0.4 < np.mean(port.notional > 0) < 0.6
    assert len(monthly_dates()) == 360
    sets = batched_sets(4, seed=1)
    assert all(10 <= p.n_trades <= 2000 and 10_000 <= n <= 200_000 for p, n in sets)


def test_greek_measure_validates_like_the_reference():
    """greeks.cpp:954-1010, 1062-1120: malformed risk factors, orders and methods are refused before
    anything runs."""
    eng = engine.Engine()
    model = eng.create_model("HullWhite1F", {"a": 0.1, "b": 0.03, "sigma": 0.01, "r0": 0.02})
    product = eng.create_product("IRSwap", {"notional": 1e6, "fixed_rate": 0.02, "payment_times": [1.0, 2.0],
                                            "accruals": [1.0, 1.0]})
    market = engine.MarketSnapshot(pillars=[1.0], zero_rates=[0.02], hazard_rate=0.02, recovery_rate=0.4)
    pricing = engine.PricingContext({"n_paths": 10.0, "n_steps": 10.0, "seed": 1.0})
    gpu = engine.ExecutionContext({"backend": "gpu"})

    def price(**params):
        return eng.price(product, [("Greek", params)], model, market, pricing, gpu)

    for rf, msg in (("a", "formato"), ("model.", "requiere"), ("curve.pillar:-1", "entero"),
                    ("curve.pillar:+1", "entero"), ("curve.pillar: 1", "entero"), ("curve.pillar:", "entero"),
                    ("curve.pillar:1x", "entero"), ("curve.flat", "desconocido"),
                    ("credit.spread", "desconocido"), ("time.rho", "desconocido"), ("fx.spot", "scope")):
        with pytest.raises(ValueError, match=msg):
            price(metric="UnilateralCVA", risk_factor=rf)
    with pytest.raises(ValueError, match="order debe ser 1 o 2"):
        price(metric="UnilateralCVA", risk_factor="model.a", order=3.0)
    with pytest.raises(ValueError, match="tercera derivada"):
        price(metric="UnilateralCVA", risk_factor="model.a", order=2.0, cross_factor="model.b")
    with pytest.raises(ValueError, match="time.theta"):
        price(metric="UnilateralCVA", risk_factor="time.theta")
    with pytest.raises(ValueError, match="ModelParameter"):
        price(metric="UnilateralCVA", risk_factor="credit.hazard_rate", method="pathwise")
    with pytest.raises(ValueError, match="GreekMethod"):
        price(metric="UnilateralCVA", risk_factor="model.a", method="finite")
    with pytest.raises(ValueError, match="no disponible"):
        price(metric="UnilateralCVA", risk_factor="model.a", method="aad")
    with pytest.raises(ValueError, match="no tiene el parametro"):
        price(metric="UnilateralCVA", risk_factor="model.kappa")
    with pytest.raises(ValueError, match="fuera de rango"):
        price(metric="UnilateralCVA", risk_factor="curve.pillar:3")
    with pytest.raises(ValueError, match="derivadas cruzadas"):
        price(metric="UnilateralCVA", risk_factor="model.a", cross_factor="time.theta")
    for unserved in ("likelihood_ratio_hessian", "aad_forward_over_forward"):
        with pytest.raises(ValueError, match="no disponible en este motor"):
            price(metric="UnilateralCVA", risk_factor="model.a", method=unserved)
