"""Closed forms under HW1F with constant mean level (Vasicek), used to pin the Bermudan LSM chain.

TEST INFRASTRUCTURE ONLY (imported by ``tests/``). Textbook results, Brigo & Mercurio, "Interest
Rate Models - Theory and Practice", 2nd ed., eq. (3.40)-(3.41) (zero-coupon bond option, Vasicek):

  ZBP(t, T, S, X) = X P(t,T) N(-h + s_p) - P(t,S) N(-h)
  s_p = sigma sqrt((1 - exp(-2 a (T-t))) / (2a)) B(T,S),   h = ln(P(t,S) / (P(t,T) X)) / s_p + s_p / 2

A payer swaption exercisable once at T into a one-period swap paying fixed K at T+1 (accrual 1)
pays N [1 - (1+K) P(T,T+1)]^+ = N (1+K) [1/(1+K) - P(T,T+1)]^+ at T: a put on the bond.
"""
from __future__ import annotations

import math

import numpy as np

from . import hw_a, hw_b


def _ncdf(x):
    return 0.5 * np.vectorize(math.erfc)(-np.asarray(x, dtype=np.float64) / math.sqrt(2.0))


def zcb(model, r, t, T):
    a, b, sigma, _ = model
    tau = T - t
    bf = hw_b(a, tau)
    return hw_a(a, b, sigma, tau) * np.exp(-bf * np.asarray(r, dtype=np.float64))


def zbp(model, r, t, T, S, X):
    """Put on the S-maturity bond, expiry T, strike X, seen at t < T with short rate r."""
    a, _, sigma, _ = model
    sp = sigma * math.sqrt((1.0 - math.exp(-2.0 * a * (T - t))) / (2.0 * a)) * hw_b(a, S - T)
    pT, pS = zcb(model, r, t, T), zcb(model, r, t, S)
    h = np.log(pS / (pT * X)) / sp + 0.5 * sp
    return X * pT * _ncdf(-h + sp) - pS * _ncdf(-h)


def payer_one_period(model, r, t, T, notional, strike):
    """Value at t of the one-period payer swaption above (t < T)."""
    return notional * (1.0 + strike) * zbp(model, r, t, T, T + 1.0, 1.0 / (1.0 + strike))
