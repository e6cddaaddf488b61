"""Batched netting sets: one simulation + one (autotuned, persistent) valuation launch for many jobs
with uneven trade and path counts; set-level EE / PFE95 / CVA must equal unbatched plans bit for bit
(same seed, grid and discretisation; job j uses the first n_j paths of the counter-based stream)."""
import numpy as np
import pytest

from engine_quant_amd import eqhip
from engine_quant_amd.portfolios import MODEL, quarterly_dates, swap_portfolio, zero_curve

pytestmark = pytest.mark.gpu


def test_batched_sets_equal_unbatched_bit_for_bit(ctx):
    pillars, zeros = zero_curve(11)
    rng = np.random.default_rng(11)
    jobs = []
    for j in range(6):
        n_tr = int(rng.integers(5, 120))
        port = swap_portfolio(n_tr, seed=100 + j, pillars=pillars, zeros=zeros)
        port.set_id = (np.arange(n_tr) % (1 + j % 3)).astype(np.int32)  # 1-3 netting sets per job
        jobs.append((port, int(rng.integers(300, 3000))))
    dates = quarterly_dates(10)
    batch = eqhip.Batch(ctx, MODEL, jobs, dates, len(dates), seed=5, disc=eqhip.DISC_EXACT, want_trade_ee=True)
    got = batch.execute(0.02, 0.4)
    tm = batch.timing()
    assert "chosen" in tm["tune_log"]
    again = batch.execute(0.02, 0.4)  # the tuned launch reproduces itself
    for j, (port, n_paths) in enumerate(jobs):
        for key in ("set_ee", "set_pfe", "set_cva"):
            assert np.array_equal(got[j][key], again[j][key]), (j, key)
        solo = ctx.exposure_run(MODEL, port, dates, n_paths, len(dates), seed=5, rng=eqhip.RNG_PHILOX,
                                disc=eqhip.DISC_EXACT, hazard_rate=0.02, recovery_rate=0.4, want_trade_ee=True)
        for key in ("set_ee", "set_pfe", "set_cva"):
            assert np.array_equal(got[j][key], solo[key]), (j, key)
        scale = np.abs(solo["trade_ee"]).max() + 1.0
        assert np.max(np.abs(got[j]["trade_ee"] - solo["trade_ee"])) <= 1e-12 * scale
    batch.close()
