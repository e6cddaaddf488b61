"""Shared test inputs (small portfolios with mixed schedules) and comparison helpers."""
import numpy as np

from engine_quant_amd.eqhip import Portfolio

MODEL = (0.10, 0.03, 0.01, 0.02)


def small_portfolio(n_trades=12, n_sets=3, seed=3, max_years=6, start=0.0):
    rng = np.random.default_rng(seed)
    pays, accs, off, fixed = [], [], [0], []
    for j in range(n_trades):
        f = int(rng.choice([1, 2, 4]))
        n_per = int(rng.integers(1, max_years * f + 1))
        s0 = start if j % 5 else start + (1.0 / f) * int(rng.integers(0, 3))
        pay = s0 + np.arange(1, n_per + 1) / f
        pays.append(pay)
        accs.append(np.full(n_per, 1.0 / f))
        off.append(off[-1] + n_per)
        fixed.append(0.02 + rng.normal(0, 0.005))
    starts = np.array([pays[j][0] - accs[j][0] for j in range(n_trades)])
    notional = np.exp(rng.normal(np.log(1e6), 0.7, n_trades)) * np.where(rng.uniform(size=n_trades) < 0.5, 1, -1)
    return Portfolio(set_id=rng.integers(0, n_sets, n_trades).astype(np.int32), notional=notional,
                     fixed_rate=np.array(fixed), start=starts, flow_off=np.array(off, dtype=np.int32),
                     pay=np.concatenate(pays), acc=np.concatenate(accs))


def diverse_book(n_trades=600, n_sets=2, seed=11, max_years=10):
    """A book of mostly distinct schedules on the monthly grid: forward starts up to 3y ahead, trades
    already running whose first period is a short stub (the rest of a period that began before today,
    paid at a random month), and spot starts; frequencies 1, 2, 4; tenors up to max_years."""
    rng = np.random.default_rng(seed)
    pays, accs, off, starts = [], [], [0], []
    for _ in range(n_trades):
        f = int(rng.choice([1, 2, 4]))
        n_per = int(rng.integers(1, max_years * f + 1))
        kind = rng.uniform()
        if kind < 0.4:
            s0 = int(rng.integers(1, 37)) / 12.0
            pay = s0 + np.arange(1, n_per + 1) / f
        elif kind < 0.8:
            s0 = 0.0
            stub = int(rng.integers(1, 12 // f)) / 12.0 if f < 12 else 1.0 / 12.0
            pay = stub + np.arange(0, n_per) / f
        else:
            s0 = 0.0
            pay = np.arange(1, n_per + 1) / f
        pay = np.round(pay * 12.0) / 12.0
        acc = np.diff(np.concatenate([[s0], pay]))
        pays.append(pay)
        accs.append(acc)
        starts.append(s0)
        off.append(off[-1] + n_per)
    notional = np.exp(rng.normal(np.log(1e6), 0.7, n_trades)) * np.where(rng.uniform(size=n_trades) < 0.5, 1, -1)
    return Portfolio(set_id=rng.integers(0, n_sets, n_trades).astype(np.int32), notional=notional,
                     fixed_rate=0.02 + rng.normal(0, 0.005, n_trades), start=np.array(starts),
                     flow_off=np.array(off, dtype=np.int32), pay=np.concatenate(pays), acc=np.concatenate(accs))


def leg_scale(port, model, dates):
    """|notional| * (1 + sum of accruals * |K|) per trade: the magnitude of the swap's legs, the
    scale of its mark-to-market rounding."""
    scale = np.zeros(port.n_trades)
    for j in range(port.n_trades):
        f0, f1 = port.flow_off[j], port.flow_off[j + 1]
        scale[j] = abs(port.notional[j]) * (2.0 + np.sum(port.acc[f0:f1]) * abs(port.fixed_rate[j]))
    return scale


def assert_value_rel(actual, expected, rel=1e-10, leg=None, what=""):
    """EE / PFE / CVA against the oracle relative to the quantity itself: |a - e| <= rel * max(|e|, m, f),
    with m the largest |e| of the row (the last axis: a profile over dates; a scalar or a vector of
    CVAs is its own row) and f = 1e-6 * leg, the legs' magnitude of the row (the floor at which the
    rounding of a value that is a difference of its legs is the whole error, for trades whose
    exposure is almost nil). Rows the oracle leaves at exactly zero must be exactly zero."""
    a, e = np.asarray(actual, dtype=np.float64), np.asarray(expected, dtype=np.float64)
    assert a.shape == e.shape, f"{what}: shape {a.shape} != {e.shape}"
    if e.size == 0:
        return
    row = np.abs(e).max(axis=-1, keepdims=True) if e.ndim >= 2 else np.abs(e)
    s = np.maximum(np.abs(e), row)
    if leg is not None:
        lg = np.asarray(leg, dtype=np.float64)
        s = np.maximum(s, 1e-6 * (lg.reshape(lg.shape + (1,) * (e.ndim - lg.ndim)) if e.ndim > lg.ndim else lg))
    err = np.abs(a - e)
    bad = err > rel * s
    if bad.any():
        i = np.unravel_index(int(np.argmax(np.where(bad, err / np.where(s > 0, s, 1e-300), 0.0))), e.shape)
        raise AssertionError(f"{what}: {int(bad.sum())} of {e.size} values off; worst at {i}: {a[i]!r} vs {e[i]!r} "
                             f"(scale {s[i]:.3e}, bound {rel:.0e})")


def assert_rel(actual, expected, rel, scale=None, what=""):
    actual, expected = np.asarray(actual), np.asarray(expected)
    s = np.maximum(np.abs(expected), 1.0) if scale is None else scale
    err = np.max(np.abs(actual - expected) / s) if actual.size else 0.0
    assert err <= rel, f"{what}: max scaled error {err:.3e} > {rel:.1e}"
