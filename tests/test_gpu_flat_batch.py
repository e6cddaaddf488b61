"""Homogeneous batches (one schedule, each trade its own netting set) through the reference-shaped
batch entry point (exposure.rs:169-235, the batch form of npv_batch_over_paths, irs.rs:87-114).
Their values come from the flat kernel when every date is a reset date, precedes the start or follows
maturity, and from the general valuation when a date falls inside a period. Both are checked against
the oracle on the same Philox paths and against a plan over the same trades (the general valuation).

Tolerances: the flat kernel sums the fixed leg as (N K) * sum(P tau) where the oracle sums
(N K) * (P tau) term by term, so values agree to rounding; EE (a sum of 3000 such values) and PFE
(an order statistic of them) within 1e-12 of the larger of |value| and 1.
"""
import numpy as np
import pytest

import oracle
from engine_quant_amd import eqhip
from helpers import assert_rel

pytestmark = pytest.mark.gpu

MODEL = (0.1, 0.03, 0.01, 0.02)
START, PAY, ACC = 1.0, [1.5, 2.0, 2.5, 3.0, 3.5], [0.5] * 5
CASES = {
    "reset dates": ([1.0, 1.5, 2.0, 2.5, 3.0], False),
    "before the start and after maturity": ([0.3, 0.5, 1.0, 2.0, 3.5, 4.2], True),
    "inside a period": ([1.0, 1.75, 2.0], True),
}


def _trades(n, seed):
    rng = np.random.default_rng(seed)
    return rng.choice([-1.0, 1.0], n) * rng.uniform(5e5, 5e6, n), rng.uniform(0.01, 0.04, n)


@pytest.mark.parametrize("case", list(CASES))
@pytest.mark.parametrize("disc", [eqhip.DISC_EULER, eqhip.DISC_EXACT])
def test_homogeneous_batch_matches_oracle_and_general_valuation(ctx, case, disc):
    times, any_date = CASES[case]
    notionals, fixed = _trades(7, 3)
    n_paths, n_steps, seed = 3000, 40, 9
    a, b, sigma, r0 = MODEL
    t, ee, pfe = ctx.exposure_profile_batch(a, b, sigma, r0, notionals, fixed, START, PAY, ACC, times, n_steps,
                                            n_paths, seed, disc=disc, allow_any_date=any_date)
    assert np.array_equal(t, times)
    n = len(notionals)
    port = dict(set_id=np.arange(n, dtype=np.int32), notional=notionals, fixed_rate=fixed, start=np.full(n, START),
                flow_off=np.arange(0, 5 * n + 1, 5, dtype=np.int32), pay=np.tile(PAY, n), acc=np.tile(ACC, n))
    ref = oracle.portfolio_exposure(MODEL, port, times, n_steps, oracle.philox_normals(seed, n_paths, n_steps),
                                    scheme=disc)
    assert_rel(ee, ref["set_ee"], 1e-12, what=f"{case} EE vs oracle")
    assert_rel(pfe, ref["set_pfe"], 1e-12, what=f"{case} PFE vs oracle")
    general = ctx.plan(MODEL, eqhip.Portfolio(**port), times, n_paths, n_steps, seed=seed, disc=disc)
    try:
        out = general.execute()
    finally:
        general.close()
    assert_rel(ee, out["set_ee"], 1e-12, what=f"{case} EE vs general valuation")
    assert_rel(pfe, out["set_pfe"], 1e-12, what=f"{case} PFE vs general valuation")
    if case == "before the start and after maturity":
        assert np.all(ee[:, -2:] == 0.0) and np.all(pfe[:, -2:] == 0.0)  # matured: worth zero


def test_homogeneous_batch_on_host_draws_and_at_t0(ctx):
    """Shared host draws (the parity mode) and a first date on grid step 0 (one value, r0)."""
    notionals, fixed = _trades(4, 8)
    n_paths, n_steps = 512, 16
    z = np.random.default_rng(1).normal(size=(n_steps, n_paths))
    a, b, sigma, r0 = MODEL
    times = [0.0, 1.0, 2.0, 3.0, 4.0]
    pay, acc = [1.0, 2.0, 3.0, 4.0, 5.0], [1.0] * 5
    t, ee, pfe = ctx.exposure_profile_batch(a, b, sigma, r0, notionals, fixed, 0.0, pay, acc, times, n_steps,
                                            n_paths, 0, rng=eqhip.RNG_HOST_DRAWS, host_draws=z)
    for j in range(len(notionals)):
        _, ree, rpfe = oracle.exposure_profile(a, b, sigma, r0, notionals[j], fixed[j], False, 0.0, pay, acc, times,
                                               n_steps, z)
        assert_rel(ee[j], ree, 1e-12, what=f"trade {j} EE")
        assert_rel(pfe[j], rpfe, 1e-12, what=f"trade {j} PFE")
    assert np.array_equal(ee[:, 0], pfe[:, 0])  # grid step 0: one value per trade, EE = PFE


@pytest.mark.parametrize("disc", [eqhip.DISC_EULER, eqhip.DISC_EXACT])
def test_batch_on_a_scenario_set_equals_its_own_simulation(ctx, disc):
    """The batch entry point on a scenario set simulated once equals, bit for bit, the batch call that
    simulates the same Philox paths itself: reset dates and dates past maturity (flat kernel), a date
    inside a period (general valuation), valued from the creating context and from another one."""
    notionals, fixed = _trades(5, 4)
    n_paths, n_steps, seed, t_max = 2000, 48, 13, 4.2
    sc = ctx.scenarios(MODEL, t_max, n_paths, n_steps, seed=seed, disc=disc)
    ctx2 = eqhip.Context(0)
    try:
        for times in ([1.0, 1.5, 2.0, 3.5, 4.2], [1.0, 1.75, 4.2]):
            own = ctx.exposure_profile_batch(*MODEL, notionals, fixed, START, PAY, ACC, times, n_steps, n_paths, seed,
                                             disc=disc, allow_any_date=True)
            for c in (ctx, ctx2):
                on = c.exposure_profile_batch_on(sc, notionals, fixed, START, PAY, ACC, times)
                for x, y in zip(on, own):
                    assert np.array_equal(x, y), times
        with pytest.raises(eqhip.EqhipError, match="t_max"):
            ctx.exposure_profile_batch_on(sc, notionals, fixed, START, PAY, ACC, [1.0, 2.0])
    finally:
        sc.close()
        ctx2.close()
    with pytest.raises(eqhip.EqhipError, match="t_max must be positive"):
        ctx.scenarios(MODEL, 0.0, n_paths, n_steps)
