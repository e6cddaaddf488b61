"""Pathwise adjoint of netting-set CVA (reverse sweep with short-rate checkpoints every 12 steps)
against central finite differences with common random numbers: through the GPU plan on the Philox
stream, and through the oracle on shared host draws. The reference computes these Greeks by
bump-and-revalue with CRN (cpp/engine/src/greeks.cpp:160-169)."""
import numpy as np
import pytest

import oracle
from engine_quant_amd import eqhip
from helpers import MODEL, diverse_book, small_portfolio

pytestmark = pytest.mark.gpu

H, RR = 0.02, 0.4
BUMPS = {"a": 1e-4, "b": 1e-4, "sigma": 1e-5, "r0": 1e-4}


def _model(**kw):
    m = dict(zip(("a", "b", "sigma", "r0"), MODEL))
    m.update(kw)
    return (m["a"], m["b"], m["sigma"], m["r0"])


def _cva_gpu(ctx, port, dates, n_paths, model, disc, set_index):
    res = ctx.exposure_run(model, port, dates, n_paths, len(dates), seed=9, rng=eqhip.RNG_PHILOX, disc=disc,
                           hazard_rate=H, recovery_rate=RR)
    return res["set_cva"][set_index]


@pytest.mark.parametrize("disc", [eqhip.DISC_EXACT, eqhip.DISC_EULER])
def test_adjoint_matches_crn_central_differences(ctx, disc):
    port = small_portfolio(n_trades=16, n_sets=2, seed=21)
    dates = np.arange(1, 25) / 4.0
    n_paths = 20000
    plan = ctx.plan(MODEL, port, dates, n_paths, len(dates), seed=9, rng=eqhip.RNG_PHILOX, disc=disc)
    for s in range(port.n_sets):
        cva, grad, check = plan.cva_adjoint(s, H, RR)
        assert abs(check - cva) <= 1e-10 * abs(cva)
        assert abs(cva - _cva_gpu(ctx, port, dates, n_paths, MODEL, disc, s)) <= 1e-10 * abs(cva)
        for name, h in BUMPS.items():
            up = _cva_gpu(ctx, port, dates, n_paths, _model(**{name: dict(zip("a b sigma r0".split(), MODEL))[name] + h}),
                          disc, s)
            dn = _cva_gpu(ctx, port, dates, n_paths, _model(**{name: dict(zip("a b sigma r0".split(), MODEL))[name] - h}),
                          disc, s)
            fd = (up - dn) / (2 * h)
            assert abs(grad[name] - fd) <= 2e-3 * abs(fd) + 1e-6 * abs(cva) / h * 1e-4, (s, name, grad[name], fd)
        # credit sensitivities are exact derivatives of the CVA sum
        cva_h = [plan.cva_adjoint(s, H + e, RR)[0] for e in (1e-6, -1e-6)]
        assert abs(grad["hazard_rate"] - (cva_h[0] - cva_h[1]) / 2e-6) <= 1e-6 * abs(grad["hazard_rate"])
        cva_r = [plan.cva_adjoint(s, H, RR + e)[0] for e in (1e-6, -1e-6)]
        assert abs(grad["recovery_rate"] - (cva_r[0] - cva_r[1]) / 2e-6) <= 1e-6 * abs(grad["recovery_rate"])
    plan.close()


def test_adjoint_matches_oracle_differences_on_shared_draws(ctx):
    port = small_portfolio(n_trades=8, n_sets=1, seed=5)
    dates = np.arange(1, 17) / 4.0
    n_paths, n_steps = 3000, len(dates)
    z = np.random.default_rng(4).standard_normal((n_steps, n_paths))
    plan = ctx.plan(MODEL, port, dates, n_paths, n_steps, rng=eqhip.RNG_HOST_DRAWS, disc=eqhip.DISC_EXACT)
    cva, grad, _ = plan.cva_adjoint(0, H, RR, host_draws=z)
    base = oracle.portfolio_exposure(MODEL, port.as_dict(), dates, n_steps, z, scheme=1, hazard_rate=H,
                                     recovery_rate=RR)["set_cva"][0]
    assert abs(cva - base) <= 1e-10 * abs(base)
    for i, name in enumerate(("a", "b", "sigma", "r0")):
        h = BUMPS[name]
        m_up, m_dn = list(MODEL), list(MODEL)
        m_up[i] += h
        m_dn[i] -= h
        up = oracle.portfolio_exposure(tuple(m_up), port.as_dict(), dates, n_steps, z, scheme=1, hazard_rate=H,
                                       recovery_rate=RR)["set_cva"][0]
        dn = oracle.portfolio_exposure(tuple(m_dn), port.as_dict(), dates, n_steps, z, scheme=1, hazard_rate=H,
                                       recovery_rate=RR)["set_cva"][0]
        fd = (up - dn) / (2 * h)
        assert abs(grad[name] - fd) <= 2e-3 * abs(fd) + 1e-9, (name, grad[name], fd)
    plan.close()


def test_adjoint_with_many_periods_in_progress(ctx):
    """Forward starts, running stubs (periods fixed at t = 0) and monthly dates: dozens of families in the
    middle of a period at every date, each a term P(t, T_next) / P(t_fix, T_next) whose adjoint reaches
    the short rate at its fixing step. The adjoint's own forward value and its gradient against oracle
    central differences on shared draws."""
    port = diverse_book(60, n_sets=1, seed=5, max_years=4)
    dates = np.arange(1, 49) / 12.0
    n_paths, n_steps = 2000, 48
    z = oracle.philox_normals(8, n_paths, n_steps)
    plan = ctx.plan(MODEL, port, dates, n_paths, n_steps, rng=eqhip.RNG_HOST_DRAWS, disc=eqhip.DISC_EXACT)
    cva, grad, check = plan.cva_adjoint(0, H, RR, host_draws=z)
    plan.close()

    def ocva(model):
        return oracle.portfolio_exposure_fused(model, port.as_dict(), dates, n_steps, z=z, scheme=1, hazard_rate=H,
                                               recovery_rate=RR)["set_cva"][0]

    base = ocva(MODEL)
    assert abs(cva - base) <= 1e-10 * abs(base)
    assert abs(check - base) <= 1e-10 * abs(base)
    for i, name in enumerate(("a", "b", "sigma", "r0")):
        m_up, m_dn = list(MODEL), list(MODEL)
        m_up[i] += BUMPS[name]
        m_dn[i] -= BUMPS[name]
        fd = (ocva(tuple(m_up)) - ocva(tuple(m_dn))) / (2 * BUMPS[name])
        assert abs(grad[name] - fd) <= 2e-3 * abs(fd) + 1e-9 * abs(cva), (name, grad[name], fd)


def test_pv_pillar_and_model_sensitivities(ctx):
    """Analytic dPV/dz (all 20 pillars) against central differences of the oracle's market PV and
    against the one-sided bucketed DV01 (measure.cpp:211-222, bump 1bp, not divided by h); the model
    NPV Greeks against central differences of the oracle's model NPV at t = 0 (api.rs:394-415)."""
    from engine_quant_amd.portfolios import swap_portfolio, zero_curve
    pillars, zeros = zero_curve(7)
    port = swap_portfolio(40, seed=9, pillars=pillars, zeros=zeros)
    dz, npv, greeks = ctx.pv_sensitivities(pillars, zeros, port, MODEL)
    _, _, bk = ctx.pv_dv01(pillars, zeros, port, bump=1e-4, bucketed=True)
    off = port.flow_off
    h = 1e-6
    for j in range(port.n_trades):
        pay, acc = port.pay[off[j]:off[j + 1]], port.acc[off[j]:off[j + 1]]
        N, K = port.notional[j], port.fixed_rate[j]
        for p in range(len(pillars)):
            up, dn = zeros.copy(), zeros.copy()
            up[p] += h
            dn[p] -= h
            fd = (oracle.npv_from_market(pillars, up, N, K, 0.0, pay, acc)
                  - oracle.npv_from_market(pillars, dn, N, K, 0.0, pay, acc)) / (2 * h)
            assert abs(dz[j, p] - fd) <= 1e-6 * abs(N) * 30, (j, p, dz[j, p], fd)
            # the reference's bucketed DV01 is the one-sided difference: bump * dPV/dz + O(bump^2)
            assert abs(bk[j, p] - 1e-4 * dz[j, p]) <= 5e-3 * abs(bk[j, p]) + 1e-6 * abs(N), (j, p)
        base = oracle.irs_npv_reset(*MODEL, 0.0, N, K, 0.0, pay, acc)
        assert abs(npv[j] - base) <= 1e-9 * abs(N)
        for i in range(4):
            mu, md = list(MODEL), list(MODEL)
            mu[i] += h
            md[i] -= h
            a_u, b_u, s_u, r_u = mu
            a_d, b_d, s_d, r_d = md
            fd = (oracle.irs_npv_reset(a_u, b_u, s_u, r_u, 0.0, N, K, 0.0, pay, acc)
                  - oracle.irs_npv_reset(a_d, b_d, s_d, r_d, 0.0, N, K, 0.0, pay, acc)) / (2 * h)
            assert abs(greeks[j, i] - fd) <= 1e-6 * abs(fd) + 1e-4 * abs(N) * 1e-4, (j, i, greeks[j, i], fd)
