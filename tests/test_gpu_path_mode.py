"""Path mode: plans without per-trade outputs, with at most 4 schedule families and 8 netting sets, are
valued one thread per path (value_path_kernel) with the family sums and set values in registers. Checked
against the oracle on shared host draws (per-path netting-set values to 1e-12 of the legs' magnitude, EE /
PFE95 / CVA to 1e-10 of the value), against the tiled variants of the same plan (to rounding: the set sums
run in another order), bit for bit against itself batched and unbatched, and on books whose schedules
exceed a tiled block's LDS. The books mix every family state a date can meet: a period starting at the
date itself, a period in progress (the fixing of its reset), and a forward start."""
import numpy as np
import pytest

import oracle
from engine_quant_amd import eqhip
from engine_quant_amd.portfolios import MODEL2
from helpers import MODEL, assert_rel, assert_value_rel, leg_scale

pytestmark = pytest.mark.gpu

H, RR = 0.02, 0.4
# (start, payments per year): spot annual, spot semiannual, forward quarterly, forward semiannual; the
# last two only in books made for the tiled kernels (more families than path mode holds)
FAMILIES = [(0.0, 1), (0.0, 2), (0.25, 4), (1.5, 2), (0.0, 4), (0.75, 1)]


def family_book(n_trades, n_sets, n_fam, seed, max_years=5):
    """Trades drawn from the first n_fam schedule families: each trade's schedule is a prefix of its
    family's, so the plan has exactly n_fam families; every set id below n_sets is used."""
    rng = np.random.default_rng(seed)
    pays, accs, off, starts = [], [], [0], []
    for j in range(n_trades):
        s0, f = FAMILIES[j % n_fam]
        n_per = int(rng.integers(1, max_years * f + 1))
        pays.append(s0 + np.arange(1, n_per + 1) / f)
        accs.append(np.full(n_per, 1.0 / f))
        starts.append(s0)
        off.append(off[-1] + n_per)
    notional = np.exp(rng.normal(np.log(1e6), 0.7, n_trades)) * np.where(rng.uniform(size=n_trades) < 0.5, 1, -1)
    set_id = (np.arange(n_trades) % n_sets).astype(np.int32)
    rng.shuffle(set_id)
    return eqhip.Portfolio(set_id=set_id, notional=notional, fixed_rate=0.02 + rng.normal(0, 0.005, n_trades),
                           start=np.array(starts), flow_off=np.array(off, dtype=np.int32), pay=np.concatenate(pays),
                           acc=np.concatenate(accs))


def _dates():
    # grid index 0, reset dates of every family, and dates between resets
    return np.concatenate([[0.0, 0.5, 1.0, 1.5], np.arange(1, 12 * 4 + 1) / 12.0 + 1.0 / 48.0])


def _set_leg(port, dates):
    leg = leg_scale(port, MODEL, dates)
    return np.array([leg[port.set_id == s].sum() for s in range(port.n_sets)])


def _check_oracle(got, ref, port, dates):
    set_leg = _set_leg(port, dates)
    assert_rel(got["set_mtm"], ref["set_mtm"], 1e-12, scale=set_leg[None, :, None], what="per-path set MtM")
    assert_value_rel(got["set_ee"], ref["set_ee"], leg=set_leg, what="set EE")
    assert_value_rel(got["set_pfe"], ref["set_pfe"], leg=set_leg, what="set PFE95")
    assert_value_rel(got["set_cva"], ref["set_cva"], leg=set_leg, what="set CVA")


def _in_path_mode(plan):
    t = plan.timing()
    return t["trades_per_item"] == 0 and t["paths_per_tile"] in (512, 1024)


@pytest.mark.parametrize("n_sets,n_fam", [(1, 2), (3, 4), (8, 3), (5, 1)])
def test_path_mode_against_the_oracle(ctx, n_sets, n_fam):
    port = family_book(40, n_sets, n_fam, seed=10 + n_sets)
    dates = _dates()
    n_paths = 1237  # a partial last tile
    n_steps = int(round(dates.max() * 48))  # dt = 1/48 puts every date on the grid
    z = np.random.default_rng(n_sets).standard_normal((n_steps, n_paths))
    print(f"path mode: {n_sets} sets, {n_fam} families, {n_paths} paths", flush=True)
    plan = ctx.plan(MODEL, port, dates, n_paths, n_steps, rng=eqhip.RNG_HOST_DRAWS, disc=eqhip.DISC_EXACT)
    got = plan.execute(z, H, RR, want_set_mtm=True)
    assert _in_path_mode(plan)
    plan.close()
    ref = oracle.portfolio_exposure(MODEL, port.as_dict(), dates, n_steps, z, scheme=1, hazard_rate=H,
                                    recovery_rate=RR, want_path=True)
    _check_oracle(got, ref, port, dates)


def test_path_mode_two_factor(ctx):
    port = family_book(30, 2, 4, seed=5)
    dates = _dates()
    n_paths = 900
    n_steps = int(round(dates.max() * 48))
    z = oracle.philox_normals2(3, n_paths, n_steps)
    plan = ctx.plan(MODEL2, port, dates, n_paths, n_steps, rng=eqhip.RNG_HOST_DRAWS, disc=eqhip.DISC_EXACT)
    got = plan.execute(z, H, RR, want_set_mtm=True)
    assert _in_path_mode(plan)
    plan.close()
    ref = oracle.portfolio_exposure(MODEL2, port.as_dict(), dates, n_steps, z, scheme=1, hazard_rate=H,
                                    recovery_rate=RR, want_path=True)
    _check_oracle(got, ref, port, dates)


def test_path_and_tiled_modes_agree(ctx):
    """The same plan in path mode, then in the tiled variants, then in path mode again: the tiled values
    agree to rounding, the second path-mode run equals the first bit for bit."""
    port = family_book(60, 3, 4, seed=2)
    dates = _dates()
    n_paths, n_steps = 5000, int(round(dates.max() * 48))
    plan = ctx.plan(MODEL, port, dates, n_paths, n_steps, seed=9, rng=eqhip.RNG_PHILOX, disc=eqhip.DISC_EULER)
    first = plan.execute(None, H, RR, want_set_mtm=True)
    assert _in_path_mode(plan)
    set_leg = _set_leg(port, dates)
    for ppl, tpi in ((1, 4), (2, 4)):
        plan.set_variant(ppl, tpi)
        tiled = plan.execute(None, H, RR, want_set_mtm=True)
        assert not _in_path_mode(plan)
        assert_rel(tiled["set_mtm"], first["set_mtm"], 1e-13, scale=set_leg[None, :, None], what=f"{ppl} set MtM")
        assert_value_rel(tiled["set_ee"], first["set_ee"], rel=1e-12, leg=set_leg, what=f"{ppl} set EE")
    plan.set_variant(0, 4)
    again = plan.execute(None, H, RR, want_set_mtm=True)
    assert _in_path_mode(plan)
    plan.close()
    for key in ("set_ee", "set_pfe", "set_cva", "set_mtm"):
        assert np.array_equal(again[key], first[key]), key


def test_path_mode_serves_schedules_beyond_a_blocks_lds(ctx):
    """One 400-payment schedule (more cash-flow dates than a tiled block's LDS holds): path mode values it
    without parts; with per-trade EE asked for, the plan is refused as before."""
    pay = np.arange(1, 401) / 12.0
    port = eqhip.Portfolio(set_id=np.array([0, 0, 1], np.int32), notional=np.array([1e6, -4e5, 7e5]),
                           fixed_rate=np.array([0.02, 0.025, 0.018]), start=np.zeros(3),
                           flow_off=np.array([0, 400, 520, 800], np.int32),
                           pay=np.concatenate([pay, pay[:120], pay[:280]]), acc=np.full(800, 1 / 12))
    dates = np.array([0.5, 1.0 + 1.0 / 48.0, 7.25, 20.0, 33.3])
    n_paths, n_steps = 700, 200
    z = np.random.default_rng(4).standard_normal((n_steps, n_paths))
    plan = ctx.plan(MODEL, port, dates, n_paths, n_steps, rng=eqhip.RNG_HOST_DRAWS)
    got = plan.execute(z, H, RR, want_set_mtm=True)
    assert _in_path_mode(plan) and plan.timing()["n_parts"] == 1
    with pytest.raises(eqhip.EqhipError, match="does not fit") as e:
        plan.set_variant(1, 4)
    assert e.value.code == eqhip.E_CAPACITY
    plan.close()
    ref = oracle.portfolio_exposure(MODEL, port.as_dict(), dates, n_steps, z, scheme=0, hazard_rate=H,
                                    recovery_rate=RR, want_path=True)
    _check_oracle(got, ref, port, dates)
    with pytest.raises(eqhip.EqhipError, match="more LDS than a block has") as e:
        ctx.plan(MODEL, port, dates, n_paths, n_steps, want_trade_ee=True)
    assert e.value.code == eqhip.E_CAPACITY


def test_path_mode_refused_where_it_does_not_apply(ctx):
    port = family_book(20, 2, 2, seed=1)
    dates = _dates()
    plan = ctx.plan(MODEL, port, dates, 256, 200, want_trade_ee=True)
    with pytest.raises(eqhip.EqhipError, match="path mode serves") as e:
        plan.set_variant(0, 4)
    assert e.value.code == eqhip.E_CAPACITY
    plan.close()
    nine = family_book(20, 9, 2, seed=1)  # nine netting sets
    plan = ctx.plan(MODEL, nine, dates, 256, 200)
    plan.execute(None, H, RR)
    assert not _in_path_mode(plan)
    with pytest.raises(eqhip.EqhipError, match="path mode serves"):
        plan.set_variant(0, 4)
    plan.close()


def test_keep_cube_in_path_mode(ctx):
    port = family_book(25, 2, 4, seed=8)
    dates = _dates()
    n_paths, n_steps = 3000, int(round(dates.max() * 48))
    z = oracle.philox_normals(6, n_paths, n_steps)
    plan = ctx.plan(MODEL, port, dates, n_paths, n_steps, rng=eqhip.RNG_HOST_DRAWS, disc=eqhip.DISC_EXACT,
                    keep_cube=True)
    got = plan.execute(z, H, RR)
    assert _in_path_mode(plan)
    cube, _ = plan.exposure_cube(1)
    plan.close()
    assert_value_rel(np.maximum(cube, 0.0).mean(axis=1), got["set_ee"][1], rel=1e-12, what="EE of the cube")
    ref = oracle.portfolio_exposure(MODEL, port.as_dict(), dates, n_steps, z, scheme=1, want_path=True)
    assert_rel(cube, ref["set_mtm"][:, 1, :], 1e-12, scale=_set_leg(port, dates)[1], what="cube")


def test_batch_mixes_path_mode_and_tiled_jobs_bit_for_bit(ctx):
    """Jobs in path mode (at most 4 families and 8 sets) beside jobs that take the tiled kernels (more
    families): the batch runs both launches and every job equals its unbatched plan bit for bit."""
    dates = np.arange(1, 4 * 6 + 1) / 4.0
    jobs = [(family_book(30, 1, 2, seed=1), 2000), (family_book(50, 3, 4, seed=2), 700),
            (family_book(40, 2, 6, seed=4), 1500), (family_book(9, 8, 1, seed=3), 3100),
            (family_book(25, 1, 5, seed=9), 900)]
    batch = eqhip.Batch(ctx, MODEL, jobs, dates, len(dates), seed=5, disc=eqhip.DISC_EXACT)
    got = batch.execute(H, RR)
    log = batch.timing()["tune_log"]
    batch.close()
    n_path = int(log.split("path mode: ")[1].split(" of")[0])
    assert 0 < n_path < len(jobs), log
    modes = []
    for j, (port, n_paths) in enumerate(jobs):
        plan = ctx.plan(MODEL, port, dates, n_paths, len(dates), seed=5, rng=eqhip.RNG_PHILOX, disc=eqhip.DISC_EXACT)
        solo = plan.execute(None, H, RR)
        modes.append(_in_path_mode(plan))
        plan.close()
        for key in ("set_ee", "set_pfe", "set_cva"):
            assert np.array_equal(got[j][key], solo[key]), (j, key)
    assert sum(modes) == n_path
