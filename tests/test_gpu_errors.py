"""Error behaviour of the C ABI: every invalid input is refused on the host, before any kernel runs,
with EQHIP_E_INVALID or EQHIP_E_CAPACITY and a message (the reference's int status + last-error
convention, cpp/engine/src/abi.cpp:43-45, and its assert messages, exposure.rs:83-90).
"""
import ctypes

import numpy as np
import pytest

from engine_quant_amd import eqhip
from helpers import MODEL, small_portfolio

pytestmark = pytest.mark.gpu

PAY5, ACC5 = [1.0, 2.0, 3.0, 4.0, 5.0], [1.0] * 5


def _port(set_id, flow_off, pay, start=None, n=None):
    n = len(flow_off) - 1 if n is None else n
    return eqhip.Portfolio(set_id=np.array(set_id, np.int32), notional=np.full(n, 1e6),
                           fixed_rate=np.full(n, 0.02), start=np.zeros(n) if start is None else np.array(start),
                           flow_off=np.array(flow_off, np.int32), pay=np.array(pay, float),
                           acc=np.ones(len(pay)))


def _raises(code, match, fn, *a, **k):
    with pytest.raises(eqhip.EqhipError, match=match) as e:
        fn(*a, **k)
    assert e.value.code == code


def test_portfolio_validation(ctx):
    dates = [0.5, 1.0]
    bad = [
        (_port([0, 0], [1, 2, 3], [1.0, 2.0, 3.0]), "flow_offset\\[0\\] must be 0"),
        (_port([0, 0], [0, 2, 2], [1.0, 2.0]), "at least one period"),
        (_port([0], [0, 2], [2.0, 1.0]), "increase strictly"),
        (_port([0], [0, 1], [1.0], start=[1.0]), "increase strictly"),
    ]
    for port, msg in bad:
        _raises(eqhip.E_INVALID, msg, ctx.plan, MODEL, port, dates, 64, 4)
    # set ids outside [0, n_sets): built by hand (the Python class derives n_sets from them)
    port = _port([0, 3], [0, 1, 2], [1.0, 1.0])
    pc = port._c()
    pc.n_sets = 2
    mc = eqhip.McC(64, 4, 0, eqhip.RNG_PHILOX, eqhip.DISC_EULER)
    m = eqhip.ModelHW1F(*MODEL)
    d = np.array(dates)
    h = ctypes.c_void_p()
    code = eqhip.lib().eqhip_plan_create(ctx.handle, ctypes.byref(m), ctypes.byref(pc),
                                         d.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), 2, ctypes.byref(mc), 0,
                                         ctypes.byref(h))
    assert code == eqhip.E_INVALID and "set_id out of range" in eqhip.last_error()


def test_plan_settings_validation(ctx):
    port = small_portfolio(n_trades=3, n_sets=1, seed=2)
    _raises(eqhip.E_INVALID, "n_paths", ctx.plan, MODEL, port, [1.0], 0, 4)
    _raises(eqhip.E_INVALID, "n_steps", ctx.plan, MODEL, port, [1.0], 64, 0)
    _raises(eqhip.E_INVALID, "unknown rng", ctx.plan, MODEL, port, [1.0], 64, 4, rng=7)
    _raises(eqhip.E_INVALID, "unknown discretisation", ctx.plan, MODEL, port, [1.0], 64, 4, disc=9)
    _raises(eqhip.E_INVALID, "at least one date", ctx.plan, MODEL, port, [], 64, 4)
    _raises(eqhip.E_INVALID, "a must be non-zero", ctx.plan, (0.0, 0.03, 0.01, 0.02), port, [1.0], 64, 4)
    _raises(eqhip.E_INVALID, "a, b and a \\+ b", ctx.plan, (0.1, -0.1, 0.01, 0.01, 0.0, 0.02), port, [1.0], 64, 4)


def test_execute_validation(ctx):
    port = small_portfolio(n_trades=3, n_sets=1, seed=2)
    plan = ctx.plan(MODEL, port, [0.5, 1.0], 64, 4, rng=eqhip.RNG_HOST_DRAWS)
    try:
        with pytest.raises(ValueError, match="host_draws must have shape"):
            plan.execute(np.zeros((4, 65)))
        code = eqhip.lib().eqhip_plan_execute(plan.handle, None, None, None)
        assert code == eqhip.E_INVALID and "set_ee / set_pfe" in eqhip.last_error()
        ee, pfe = np.zeros(2), np.zeros(2)
        out = eqhip.OutputsC(eqhip._dp(ee), eqhip._dp(pfe), None, None, None, None, None)
        code = eqhip.lib().eqhip_plan_execute(plan.handle, None, None, ctypes.byref(out))
        assert code == eqhip.E_INVALID and "host_draws required" in eqhip.last_error()
        tee = np.zeros((3, 2))
        out = eqhip.OutputsC(eqhip._dp(ee), eqhip._dp(pfe), None, eqhip._dp(tee), None, None, None)
        z = np.zeros((4, 64))
        code = eqhip.lib().eqhip_plan_execute(plan.handle, eqhip._dp(z), None, ctypes.byref(out))
        assert code == eqhip.E_INVALID and "want_trade_ee" in eqhip.last_error()
        cva = np.zeros(1)
        out = eqhip.OutputsC(eqhip._dp(ee), eqhip._dp(pfe), eqhip._dp(cva), None, None, None, None)
        code = eqhip.lib().eqhip_plan_execute(plan.handle, eqhip._dp(z), None, ctypes.byref(out))
        assert code == eqhip.E_INVALID and "credit" in eqhip.last_error()
    finally:
        plan.close()
    assert eqhip.lib().eqhip_plan_execute(None, None, None, None) == eqhip.E_INVALID


def test_capacity_limits(ctx):
    # G2++: 200 distinct schedules (400 cash-flow dates) need more LDS than a block has, and two-factor
    # plans are not valued in parts
    from engine_quant_amd.portfolios import MODEL2
    n = 200
    start = np.arange(n) / 1000.0
    pay = np.stack([start + 1.0, start + 2.0], axis=1).ravel()
    port = eqhip.Portfolio(set_id=np.zeros(n, np.int32), notional=np.full(n, 1e6), fixed_rate=np.full(n, 0.02),
                           start=start, flow_off=np.arange(0, 2 * n + 1, 2, dtype=np.int32), pay=pay,
                           acc=np.ones(2 * n))
    _raises(eqhip.E_CAPACITY, "too diverse", ctx.plan, MODEL2, port, [0.5, 1.5], 64, 4)
    # a batch job must fit one block as well
    _raises(eqhip.E_CAPACITY, "too diverse", eqhip.Batch, ctx, MODEL, [(port, 64)], [0.5, 1.5], 4)
    # per-trade EE of a G2++ plan: more work items (8 same-signed trades of a class each) than threads
    n = 8 * 1024 + 1
    port = eqhip.Portfolio(set_id=np.zeros(n, np.int32), notional=np.full(n, 1e6), fixed_rate=np.full(n, 0.02),
                           start=np.zeros(n), flow_off=np.arange(n + 1, dtype=np.int32), pay=np.ones(n),
                           acc=np.ones(n))
    _raises(eqhip.E_CAPACITY, "per-trade EE supports at most", ctx.plan, MODEL2, port, [0.5], 64, 4,
            want_trade_ee=True)
    ctx.plan(MODEL2, port, [0.5], 64, 4).close()  # the same portfolio without per-trade EE is fine


def test_reference_shaped_entry_point_validation(ctx):
    a, b, s, r0 = MODEL
    _raises(eqhip.E_INVALID, "n_steps debe ser al menos 1", ctx.exposure_profile, a, b, s, r0, 1e6, 0.02, False,
            0.0, PAY5, ACC5, [0.0, 1.0], 0, 100, 1)
    _raises(eqhip.E_INVALID, "monitoring_times", ctx.exposure_profile, a, b, s, r0, 1e6, 0.02, False, 0.0, PAY5,
            ACC5, [], 8, 100, 1)
    _raises(eqhip.E_INVALID, "no es una fecha de reseteo", ctx.exposure_profile, a, b, s, r0, 1e6, 0.02, False,
            0.0, PAY5, ACC5, [0.5], 8, 100, 1)
    _raises(eqhip.E_INVALID, "al menos un periodo", ctx.exposure_profile, a, b, s, r0, 1e6, 0.02, False, 0.0,
            [], [], [0.0], 8, 100, 1)
    _raises(eqhip.E_INVALID, "al menos un trade", ctx.exposure_profile_batch, a, b, s, r0, [], [], 0.0, PAY5,
            ACC5, [0.0], 8, 100, 1)
    _raises(eqhip.E_INVALID, "rho", ctx.unilateral_cva_2f, 0.1, 0.2, 0.01, 0.01, 2.0, 0.03, [1.0], [1.0], 0.02,
            0.4)


def test_adjoint_and_batch_validation(ctx):
    port = small_portfolio(n_trades=4, n_sets=2, seed=3)
    plan = ctx.plan(MODEL, port, [1.0, 1.1, 2.0], 128, 5)  # dt = 0.4: 1.0 and 1.1 share grid step 3
    try:
        _raises(eqhip.E_INVALID, "set index", plan.cva_adjoint, 7, 0.02, 0.4)
        _raises(eqhip.E_INVALID, "one grid step per monitoring date", plan.cva_adjoint, 0, 0.02, 0.4)
    finally:
        plan.close()
    h = ctypes.c_void_p()
    pcs = (eqhip.PortfolioC * 1)(port._c())
    npaths = (ctypes.c_uint64 * 1)(128)
    d = np.array([1.0, 2.0])
    m = eqhip.ModelHW1F(*MODEL)
    code = eqhip.lib().eqhip_batch_create(ctx.handle, ctypes.byref(m), 1, pcs, npaths, eqhip._dp(d), 2, 2, 0,
                                          eqhip.RNG_HOST_DRAWS, eqhip.DISC_EULER, 0, ctypes.byref(h))
    assert code == eqhip.E_INVALID and "Philox" in eqhip.last_error()


def test_two_contexts_from_two_threads_match_serial_runs(ctx):
    """One context per host thread (its own stream): a plan executing on one while the PV/DV01 ladder and
    a second plan run on another gives the results of running them one after the other."""
    from concurrent.futures import ThreadPoolExecutor

    port = small_portfolio(n_trades=20, n_sets=3, seed=11)
    pillars, zeros = np.array([0.5, 1.0, 2.0, 5.0, 10.0]), np.array([0.02, 0.021, 0.023, 0.026, 0.028])
    dates = np.arange(1, 12 * 4 + 1) / 12.0
    ctx2 = eqhip.Context(0)
    plan_a = ctx.plan(MODEL, port, dates, 20000, 48, seed=3, want_trade_ee=True)
    plan_b = ctx2.plan((0.1, 0.2, 0.01, 0.012, -0.7, 0.03), port, dates, 20000, 48, seed=4)
    try:
        serial_a = plan_a.execute(None, 0.02, 0.4)
        serial_b = plan_b.execute(None, 0.02, 0.4)
        serial_pv = ctx2.pv_dv01(pillars, zeros, port)

        def other():
            out = []
            for _ in range(3):
                out.append((plan_b.execute(None, 0.02, 0.4), ctx2.pv_dv01(pillars, zeros, port)))
            return out

        with ThreadPoolExecutor(max_workers=1) as pool:
            fut = pool.submit(other)
            mine = [plan_a.execute(None, 0.02, 0.4) for _ in range(3)]
            theirs = fut.result()
        for res in mine:
            for k in serial_a:
                assert np.array_equal(res[k], serial_a[k]), k
        for res_b, pv in theirs:
            for k in serial_b:
                assert np.array_equal(res_b[k], serial_b[k]), k
            for x, y in zip(pv, serial_pv):
                assert np.array_equal(x, y)
    finally:
        plan_a.close()
        plan_b.close()
        ctx2.close()
