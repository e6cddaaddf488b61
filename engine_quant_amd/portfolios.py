"""Seeded synthetic workloads of the exposure benchmarks (pure host-side data generation).

* ``swap_portfolio``: N vanilla IRS drawn as in the one-netting-set GPU workload: maturity
  U(1, 30)y snapped to the payment grid round(T f)/f, frequency f in {1, 2} with equal
  probability, notional lognormal(ln 1e7, 1), payer/receiver 50/50 (receiver = negative notional,
  the reference has no direction flag, cpp/engine/src/product.cpp:7-18), fixed rate = market par
  rate (cpp/engine/src/measure.cpp:124-134) + N(0, 50bp), spot start.
* ``zero_curve``: 20 pillars linspace(0.25, 30)y, zero rates linspace(1.5%, 3.5%) + N(0, 10bp).
* ``batched_sets``: netting sets with log-uniform trade counts and path counts.
"""
from __future__ import annotations

import numpy as np

from .eqhip import Portfolio

MODEL = (0.10, 0.03, 0.01, 0.02)  # HullWhite1F a, b, sigma, r0
MODEL2 = (0.10, 0.20, 0.01, 0.012, -0.7, 0.03)  # HullWhite2F a, b, sigma, eta, rho, r0 (G2++)


def zero_curve(seed=7, n_pillars=20):
    rng = np.random.default_rng(seed + 1000)
    pillars = np.linspace(0.25, 30.0, n_pillars)
    zeros = np.linspace(0.015, 0.035, n_pillars) + rng.normal(0.0, 0.0010, n_pillars)
    return pillars, zeros


def _df(pillars, zeros, t):
    # linear zero-rate interpolation with flat extrapolation, DF = exp(-z t) (market.cpp:26-41)
    z = np.interp(t, pillars, zeros)
    return np.exp(-z * t)


def swap_portfolio(n_trades=1000, seed=7, pillars=None, zeros=None, set_ids=None):
    rng = np.random.default_rng(seed)
    if pillars is None:
        pillars, zeros = zero_curve(seed)
    freq = rng.choice([1, 2], size=n_trades)
    mat = rng.uniform(1.0, 30.0, size=n_trades)
    notional = np.exp(rng.normal(np.log(1e7), 1.0, size=n_trades))
    sign = np.where(rng.uniform(size=n_trades) < 0.5, 1.0, -1.0)
    shock = rng.normal(0.0, 0.0050, size=n_trades)
    pays, accs, off = [], [], [0]
    fixed = np.empty(n_trades)
    for j in range(n_trades):
        f = int(freq[j])
        n_per = max(1, int(round(mat[j] * f)))
        pay = np.arange(1, n_per + 1, dtype=np.float64) / f
        acc = np.full(n_per, 1.0 / f)
        dfs = _df(pillars, zeros, pay)
        par = (1.0 - dfs[-1]) / np.sum(acc * dfs)
        fixed[j] = par + shock[j]
        pays.append(pay)
        accs.append(acc)
        off.append(off[-1] + n_per)
    sid = np.zeros(n_trades, dtype=np.int32) if set_ids is None else np.asarray(set_ids, dtype=np.int32)
    return Portfolio(set_id=sid, notional=notional * sign, fixed_rate=fixed, start=np.zeros(n_trades),
                     flow_off=np.asarray(off, dtype=np.int32), pay=np.concatenate(pays), acc=np.concatenate(accs))


def monthly_dates(years=30):
    return np.arange(1, 12 * years + 1, dtype=np.float64) / 12.0


def quarterly_dates(years=30):
    return np.arange(1, 4 * years + 1, dtype=np.float64) / 4.0


def batched_sets(n_sets=64, seed=7, trade_range=(10, 2000), path_range=(10_000, 200_000)):
    """Per-set (portfolio, n_paths) with log-uniform trade and path counts; per-set seeds derive
    from the master seed."""
    rng = np.random.default_rng(seed)
    n_tr = np.exp(rng.uniform(np.log(trade_range[0]), np.log(trade_range[1]), n_sets)).astype(int)
    n_pa = np.exp(rng.uniform(np.log(path_range[0]), np.log(path_range[1]), n_sets)).astype(int)
    pillars, zeros = zero_curve(seed)
    sets = []
    for s in range(n_sets):
        sets.append((swap_portfolio(int(n_tr[s]), seed=seed * 1000 + s, pillars=pillars, zeros=zeros),
                     int(n_pa[s])))
    return sets


def _model_zcb(model, tau):
    a, b, sigma, r0 = model
    tau = np.asarray(tau, dtype=np.float64)
    bf = (1.0 - np.exp(-a * tau)) / a
    af = np.exp(((bf - tau) * (a * a * b - sigma * sigma / 2.0)) / (a * a) - sigma * sigma * bf * bf / (4.0 * a))
    return af * np.exp(-bf * r0)


def bermudan_book(n_trades=200, seed=7, model=MODEL):
    """Bermudan payer swaptions: first exercise E ~ U{1..10}y, underlying tenor L ~ U{5..20}y (annual
    exercise E..E+L-1 into the co-terminal swap ending at E+L, annual fixed payments), notional
    lognormal(ln 1e7, 1), strike = forward swap rate under the model at r0 + N(0, 50bp)."""
    rng = np.random.default_rng(seed + 500)
    first = rng.integers(1, 11, size=n_trades).astype(np.int32)
    tenor = rng.integers(5, 21, size=n_trades).astype(np.int32)
    notional = np.exp(rng.normal(np.log(1e7), 1.0, size=n_trades))
    shock = rng.normal(0.0, 0.0050, size=n_trades)
    strike = np.empty(n_trades)
    for j in range(n_trades):
        E, L = int(first[j]), int(tenor[j])
        p = _model_zcb(model, np.arange(E, E + L + 1, dtype=np.float64))
        strike[j] = (p[0] - p[-1]) / np.sum(p[1:]) + shock[j]
    return {"notional": notional, "strike": strike, "first_ex": first, "n_ex": tenor}
