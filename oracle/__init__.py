"""Python view of the oracle ``libeqref.so`` (CPU FP64 restatement of the reference path).

TEST INFRASTRUCTURE ONLY: imported by ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` as the checker / timed CPU baseline. The product package
``engine_quant_amd`` never imports it.

Every wrapper documents the reference file:line its C function restates (see ``eq_ref.c``).
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libeqref.so")
_lib = None

_D = ctypes.c_double
_I = ctypes.c_int
_U32 = ctypes.c_uint32
_U64 = ctypes.c_uint64
_PD = ctypes.POINTER(ctypes.c_double)
_PI32 = ctypes.POINTER(ctypes.c_int32)
_PU32 = ctypes.POINTER(ctypes.c_uint32)


def build() -> str:
    """Compile the oracle with its committed Makefile (gcc, no GPU needed)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.eqr_hw_b.restype = _D
        L.eqr_hw_b.argtypes = [_D, _D]
        L.eqr_hw_a.restype = _D
        L.eqr_hw_a.argtypes = [_D, _D, _D, _D, _D]
        L.eqr_zcb.restype = _D
        L.eqr_zcb.argtypes = [_D] * 6
        L.eqr_philox4x32_10.argtypes = [_PU32, _PU32, _PU32]
        L.eqr_philox_normals.argtypes = [_U64, _U64, _U32, _PD]
        L.eqr_philox_uniforms.argtypes = [_U64, _U64, _U32, _PD, _PD]
        L.eqr_simulate.argtypes = [_D, _D, _D, _D, _D, _U32, _U64, _I, _PD, _PD]
        L.eqr_is_reset_date.restype = _I
        L.eqr_is_reset_date.argtypes = [_D, _PD, _I, _D]
        L.eqr_par_rate.restype = _D
        L.eqr_par_rate.argtypes = [_D, _D, _D, _D, _D, _PD, _PD, _I]
        L.eqr_irs_npv_reset.restype = _D
        L.eqr_irs_npv_reset.argtypes = [_D] * 8 + [_PD, _PD, _I]
        L.eqr_irs_value.restype = _D
        L.eqr_irs_value.argtypes = [_D] * 9 + [_PD, _PD, _I]
        L.eqr_fixing_date.restype = _D
        L.eqr_fixing_date.argtypes = [_D, _D, _PD, _I]
        L.eqr_ee_pfe.argtypes = [_PD, _U64, _PD, _PD]
        L.eqr_unilateral_cva.restype = _D
        L.eqr_unilateral_cva.argtypes = [_D, _D, _D, _D, _PD, _PD, _I, _D, _D]
        L.eqr_exposure_profile.restype = _I
        L.eqr_exposure_profile.argtypes = [_D, _D, _D, _D, _D, _D, _I, _D, _PD, _PD, _I, _PD, _I,
                                           _U32, _U64, _PD, _PD, _PD]
        L.eqr_curve_zero_rate.restype = _D
        L.eqr_curve_zero_rate.argtypes = [_PD, _PD, _I, _D]
        L.eqr_curve_df.restype = _D
        L.eqr_curve_df.argtypes = [_PD, _PD, _I, _D]
        L.eqr_par_rate_from_market.restype = _D
        L.eqr_par_rate_from_market.argtypes = [_PD, _PD, _I, _D, _PD, _PD, _I]
        L.eqr_npv_from_market.restype = _D
        L.eqr_npv_from_market.argtypes = [_PD, _PD, _I, _D, _D, _D, _PD, _PD, _I]
        L.eqr_dv01.restype = _D
        L.eqr_dv01.argtypes = [_PD, _PD, _I, _D, _D, _D, _PD, _PD, _I, _D]
        L.eqr_npv_from_market_at.restype = _D
        L.eqr_npv_from_market_at.argtypes = [_PD, _PD, _I, _D, _D, _D, _PD, _PD, _I, _D]
        L.eqr_dv01_at.restype = _D
        L.eqr_dv01_at.argtypes = [_PD, _PD, _I, _D, _D, _D, _PD, _PD, _I, _D, _D]
        L.eqr_dv01_bucketed.argtypes = [_PD, _PD, _I, _D, _D, _D, _PD, _PD, _I, _D, _PD]
        L.eqr_portfolio_exposure.restype = _I
        L.eqr_portfolio_exposure.argtypes = [
            _D, _D, _D, _D, _I, _I, _PI32, _PD, _PD, _PD, _PI32, _PD, _PD, _PD, _I, _U32, _U64, _I,
            _PD, _D, _D, _PD, _PD, _PD, _PD, _PD, _PD, _PD, _PD, _I]
        L.eqr_max_threads.restype = _I
        L.eqr_portfolio_exposure_fused.restype = _I
        L.eqr_portfolio_exposure_fused.argtypes = [
            _D, _D, _D, _D, _I, _I, _PI32, _PD, _PD, _PD, _PI32, _PD, _PD, _PD, _I, _U32, _U64, _I, _PD, _U64,
            _D, _D, _I, _PD, _PD, _PD, _PD, _PD, _PD, _I]
        L.eqr_collateral_adjust.restype = _D
        L.eqr_collateral_adjust.argtypes = [_D, _I, _I, _D, _D, _D]
        L.eqr_xva_calculate.argtypes = [_PD, _I, _U64, _PD, _PD, _PD, _D, _D, _PD, _I, _D, _D, _PD, _I, _D, _D, _D,
                                        _D, _I, _I, _D, _D, _D, _D, _PD, _PD, _PD]
        L.eqr_hw2_var.restype = _D
        L.eqr_hw2_var.argtypes = [_D] * 3
        L.eqr_hw2_cross.restype = _D
        L.eqr_hw2_cross.argtypes = [_D] * 6
        L.eqr_zcb2.restype = _D
        L.eqr_zcb2.argtypes = [_D] * 10
        L.eqr_philox_normals2.argtypes = [_U64, _U64, _U32, _PD]
        L.eqr_simulate2.argtypes = [_D] * 6 + [_U32, _U64, _I, _PD, _PD, _PD]
        L.eqr_par_rate2.restype = _D
        L.eqr_par_rate2.argtypes = [_D] * 7 + [_PD, _PD, _I]
        L.eqr_irs_npv_reset2.restype = _D
        L.eqr_irs_npv_reset2.argtypes = [_D] * 12 + [_PD, _PD, _I]
        L.eqr_irs_value2.restype = _D
        L.eqr_irs_value2.argtypes = [_D] * 14 + [_PD, _PD, _I]
        L.eqr_unilateral_cva2.restype = _D
        L.eqr_unilateral_cva2.argtypes = [_D] * 6 + [_PD, _PD, _I, _D, _D]
        L.eqr_exposure_profile2.restype = _I
        L.eqr_exposure_profile2.argtypes = [_D] * 8 + [_I, _D, _PD, _PD, _I, _PD, _I, _U32, _U64, _PD, _PD, _PD]
        L.eqr_portfolio_exposure2.restype = _I
        L.eqr_portfolio_exposure2.argtypes = [
            _D, _D, _D, _D, _D, _D, _I, _I, _PI32, _PD, _PD, _PD, _PI32, _PD, _PD, _PD, _I, _U32, _U64, _I,
            _PD, _D, _D, _PD, _PD, _PD, _PD, _PD, _PD, _PD, _PD, _I]
        L.eqr_berm_simulate.argtypes = [_D, _D, _D, _D, _D, _U32, _U64, _PD, _PD, _PD]
        L.eqr_berm_x.restype = _D
        L.eqr_berm_x.argtypes = [_D] * 6
        L.eqr_berm_policy.restype = _I
        L.eqr_berm_policy.argtypes = [_D, _D, _D, _D, _I, _PD, _PD, _PI32, _PI32, _I, _U32, _U64, _PD, _PD,
                                      _PD, _PI32, _PD, _PD]
        L.eqr_berm_apply.restype = _I
        L.eqr_berm_apply.argtypes = [_D, _D, _D, _D, _I, _PD, _PD, _PI32, _PI32, _I, _U32, _U64, _PD, _PD,
                                     _PD, _PI32, _PD]
        L.eqr_berm_judge_fit.restype = _U64
        L.eqr_berm_judge_fit.argtypes = [_D, _D, _D, _D, _U64, _U64, _D, _PD, _PD, _PI32, _PD, _PD]
        L.eqr_berm_ee.restype = _D
        L.eqr_berm_ee.argtypes = [_D, _D, _D, _D, _U64, _U64, _D, _PD, _PI32, _PD]
        _lib = L
    return _lib


def _dp(x):
    if x is None:
        return None
    return x.ctypes.data_as(_PD)


def _f64(x):
    return np.ascontiguousarray(np.asarray(x, dtype=np.float64))


# --------------------------------------------------------------------------- model
def hw_b(a, tau):
    """B(tau), models/hull_white.rs:51-55."""
    return lib().eqr_hw_b(a, tau)


def hw_a(a, b, sigma, tau):
    """A(tau), models/hull_white.rs:59-69."""
    return lib().eqr_hw_a(a, b, sigma, tau, hw_b(a, tau))


def zcb(a, b, sigma, r, t, maturity):
    """P(t,T|r), models/hull_white.rs:75-80."""
    return lib().eqr_zcb(a, b, sigma, r, t, maturity)


# --------------------------------------------------------------------------- RNG
def philox4x32_10(ctr, key):
    c = (ctypes.c_uint32 * 4)(*[int(v) & 0xFFFFFFFF for v in ctr])
    k = (ctypes.c_uint32 * 2)(*[int(v) & 0xFFFFFFFF for v in key])
    o = (ctypes.c_uint32 * 4)()
    lib().eqr_philox4x32_10(c, k, o)
    return [int(v) for v in o]


def philox_uniforms(seed, path, pair):
    u1, u2 = _D(), _D()
    lib().eqr_philox_uniforms(seed, path, pair, ctypes.byref(u1), ctypes.byref(u2))
    return u1.value, u2.value


def philox_normals(seed, n_paths, n_steps):
    """Normals Z[step][path] of the counter-based stream shared with the GPU."""
    z = np.empty((n_steps, n_paths), dtype=np.float64)
    lib().eqr_philox_normals(seed, n_paths, n_steps, _dp(z))
    return z


# --------------------------------------------------------------------------- simulation
def simulate(a, b, sigma, r0, dt, z, scheme=0):
    """Short-rate paths r[step][path] (kernel.rs:41-50, hull_white.rs:87-97; scheme 1 = exact)."""
    z = _f64(z)
    n_steps, n_paths = z.shape
    r = np.empty((n_steps + 1, n_paths), dtype=np.float64)
    lib().eqr_simulate(a, b, sigma, r0, dt, n_steps, n_paths, scheme, _dp(z), _dp(r))
    return r


# --------------------------------------------------------------------------- IRS
def is_reset_date(start, pay, t):
    p = _f64(pay)
    return bool(lib().eqr_is_reset_date(start, _dp(p), len(p), t))


def par_rate(a, b, sigma, r, start, pay, acc):
    """irs.rs:126-145."""
    p, q = _f64(pay), _f64(acc)
    return lib().eqr_par_rate(a, b, sigma, r, start, _dp(p), _dp(q), len(p))


def irs_npv_reset(a, b, sigma, r, t, notional, fixed_rate, start, pay, acc):
    """remaining_from + npv, irs.rs:46-68,163-187."""
    p, q = _f64(pay), _f64(acc)
    return lib().eqr_irs_npv_reset(a, b, sigma, r, t, notional, fixed_rate, start, _dp(p), _dp(q),
                                   len(p))


def irs_value(a, b, sigma, r, t, r_fix, notional, fixed_rate, start, pay, acc):
    p, q = _f64(pay), _f64(acc)
    return lib().eqr_irs_value(a, b, sigma, r, t, r_fix, notional, fixed_rate, start, _dp(p),
                               _dp(q), len(p))


def fixing_date(t, start, pay):
    p = _f64(pay)
    return lib().eqr_fixing_date(t, start, _dp(p), len(p))


def ee_pfe(x):
    """exposure.rs:42-49."""
    x = _f64(x).copy()
    e, q = _D(), _D()
    lib().eqr_ee_pfe(_dp(x), len(x), ctypes.byref(e), ctypes.byref(q))
    return e.value, q.value


def unilateral_cva(a, b, sigma, r0, times, ee, hazard_rate, recovery_rate):
    """exposure.rs:242-275."""
    t, e = _f64(times), _f64(ee)
    return lib().eqr_unilateral_cva(a, b, sigma, r0, _dp(t), _dp(e), len(t), hazard_rate,
                                    recovery_rate)


def exposure_profile(a, b, sigma, r0, notional, fixed_rate, use_par_rate, start, pay, acc,
                     times, n_steps, z):
    """exposure.rs:73-138 on explicit draws z[step][path]; returns (times, ee, pfe)."""
    p, q, t = _f64(pay), _f64(acc), _f64(times)
    z = _f64(z)
    n_paths = z.shape[1]
    ee = np.zeros(len(t))
    pfe = np.zeros(len(t))
    n = lib().eqr_exposure_profile(a, b, sigma, r0, notional, fixed_rate, int(use_par_rate), start,
                                   _dp(p), _dp(q), len(p), _dp(t), len(t), n_steps, n_paths,
                                   _dp(z), _dp(ee), _dp(pfe))
    return t[:n].copy(), ee[:n], pfe[:n]


# --------------------------------------------------------------------------- G2++ (hull_white_2f.rs)
# model2 = (a, b, sigma, eta, rho, r0); state (x, y) starts at (0, 0)
def hw2_var(z, vol, tau):
    """Own-factor variance term, hull_white_2f.rs:66-75."""
    return lib().eqr_hw2_var(z, vol, tau)


def hw2_cross(a, b, sigma, eta, rho, tau):
    """Covariance term, hull_white_2f.rs:79-91."""
    return lib().eqr_hw2_cross(a, b, sigma, eta, rho, tau)


def zcb2(model2, x, y, t, maturity):
    """P(t,T|x,y), hull_white_2f.rs:96-106."""
    return lib().eqr_zcb2(*model2, x, y, t, maturity)


def philox_normals2(seed, n_paths, n_steps):
    """Z[2][step][path]: x shocks, then independent y shocks (Philox counter stream 1)."""
    z = np.empty((2, n_steps, n_paths), dtype=np.float64)
    lib().eqr_philox_normals2(seed, n_paths, n_steps, _dp(z))
    return z


def simulate2(model2, dt, z, scheme=0):
    """(x, y)[step][path], hull_white_2f.rs:116-140 (scheme 0 Euler, 1 exact)."""
    z = _f64(z)
    _, n_steps, n_paths = z.shape
    x = np.empty((n_steps + 1, n_paths))
    y = np.empty((n_steps + 1, n_paths))
    a, b, sigma, eta, rho, _r0 = model2
    lib().eqr_simulate2(a, b, sigma, eta, rho, dt, n_steps, n_paths, scheme, _dp(z), _dp(x), _dp(y))
    return x, y


def par_rate2(model2, start, pay, acc):
    """irs.rs:126-145 at state (0, 0)."""
    p, q = _f64(pay), _f64(acc)
    return lib().eqr_par_rate2(*model2, start, _dp(p), _dp(q), len(p))


def irs_npv_reset2(model2, x, y, t, notional, fixed_rate, start, pay, acc):
    p, q = _f64(pay), _f64(acc)
    return lib().eqr_irs_npv_reset2(*model2, x, y, t, notional, fixed_rate, start, _dp(p), _dp(q), len(p))


def irs_value2(model2, x, y, t, xf, yf, notional, fixed_rate, start, pay, acc):
    p, q = _f64(pay), _f64(acc)
    return lib().eqr_irs_value2(*model2, x, y, t, xf, yf, notional, fixed_rate, start, _dp(p), _dp(q), len(p))


def unilateral_cva2(model2, times, ee, hazard_rate, recovery_rate):
    """exposure.rs:452-466 (discount at state (0, 0))."""
    t, e = _f64(times), _f64(ee)
    return lib().eqr_unilateral_cva2(*model2, _dp(t), _dp(e), len(t), hazard_rate, recovery_rate)


def exposure_profile2(model2, notional, fixed_rate, use_par_rate, start, pay, acc, times, n_steps, z):
    """exposure.rs:305-370 on explicit draws z[2][step][path]; returns (times, ee, pfe)."""
    p, q, t = _f64(pay), _f64(acc), _f64(times)
    z = _f64(z)
    n_paths = z.shape[2]
    ee = np.zeros(len(t))
    pfe = np.zeros(len(t))
    n = lib().eqr_exposure_profile2(*model2, notional, fixed_rate, int(use_par_rate), start, _dp(p), _dp(q),
                                    len(p), _dp(t), len(t), n_steps, n_paths, _dp(z), _dp(ee), _dp(pfe))
    return t[:n].copy(), ee[:n], pfe[:n]


# --------------------------------------------------------------------------- XVA reducer
def collateral_adjust(value, csa=None):
    """quant-engine/src/xva.rs:48-64; csa = dict(threshold, minimum_transfer_amount, independent_amount,
    variation_margin) or None."""
    c = csa or {}
    return lib().eqr_collateral_adjust(value, int(csa is not None), int(c.get("variation_margin", True)),
                                       c.get("threshold", 0.0), c.get("minimum_transfer_amount", 0.0),
                                       c.get("independent_amount", 0.0))


def xva_calculate(times, q, df=None, p=None, counterparty=(0.0, 0.4, ()), own=(0.0, 0.4, ()), funding_spread=0.0,
                  borrowing_spread=0.0, capital_factor=0.08, initial_margin_total=0.0, csa=None, confidence=0.95):
    """quant-engine/src/xva.rs:97-183 on q (and p) [time][path]; counterparty / own = (hazard, recovery,
    survival probabilities). Returns (ee, pfe, {cva, dva, fva, mva, kva})."""
    t, qq = _f64(times), _f64(q)
    n_times, n_paths = qq.shape
    d = None if df is None else _f64(df)
    pp = None if p is None else _f64(p)
    csp, osp = _f64(counterparty[2]), _f64(own[2])
    c = csa or {}
    ee, pfe, m = np.zeros(n_times), np.zeros(n_times), np.zeros(5)
    lib().eqr_xva_calculate(_dp(t), n_times, n_paths, _dp(qq), _dp(d), _dp(pp), counterparty[0], counterparty[1],
                            _dp(csp) if len(csp) else None, len(csp), own[0], own[1], _dp(osp) if len(osp) else None,
                            len(osp), funding_spread, borrowing_spread, capital_factor, initial_margin_total,
                            int(csa is not None), int(c.get("variation_margin", True)), c.get("threshold", 0.0),
                            c.get("minimum_transfer_amount", 0.0), c.get("independent_amount", 0.0), confidence,
                            _dp(ee), _dp(pfe), _dp(m))
    return ee, pfe, dict(zip(("cva", "dva", "fva", "mva", "kva"), m))


# --------------------------------------------------------------------------- market curve
def curve_zero_rate(pillars, zeros, t):
    p, z = _f64(pillars), _f64(zeros)
    return lib().eqr_curve_zero_rate(_dp(p), _dp(z), len(p), t)


def curve_df(pillars, zeros, t):
    p, z = _f64(pillars), _f64(zeros)
    return lib().eqr_curve_df(_dp(p), _dp(z), len(p), t)


def par_rate_from_market(pillars, zeros, start, pay, acc):
    p, z, a, c = _f64(pillars), _f64(zeros), _f64(pay), _f64(acc)
    return lib().eqr_par_rate_from_market(_dp(p), _dp(z), len(p), start, _dp(a), _dp(c), len(a))


def npv_from_market(pillars, zeros, notional, fixed_rate, start, pay, acc):
    p, z, a, c = _f64(pillars), _f64(zeros), _f64(pay), _f64(acc)
    return lib().eqr_npv_from_market(_dp(p), _dp(z), len(p), notional, fixed_rate, start, _dp(a),
                                     _dp(c), len(a))


def dv01(pillars, zeros, notional, fixed_rate, start, pay, acc, bump=1e-4):
    p, z, a, c = _f64(pillars), _f64(zeros), _f64(pay), _f64(acc)
    return lib().eqr_dv01(_dp(p), _dp(z), len(p), notional, fixed_rate, start, _dp(a), _dp(c),
                          len(a), bump)


def npv_from_market_at(pillars, zeros, notional, fixed_rate, start, pay, acc, valuation_time):
    """npv_from_market at a valuation time (measure.cpp:146-184); NaN when it is after the start."""
    p, z, a, c = _f64(pillars), _f64(zeros), _f64(pay), _f64(acc)
    return lib().eqr_npv_from_market_at(_dp(p), _dp(z), len(p), notional, fixed_rate, start, _dp(a), _dp(c),
                                        len(a), valuation_time)


def dv01_at(pillars, zeros, notional, fixed_rate, start, pay, acc, valuation_time, bump=1e-4):
    p, z, a, c = _f64(pillars), _f64(zeros), _f64(pay), _f64(acc)
    return lib().eqr_dv01_at(_dp(p), _dp(z), len(p), notional, fixed_rate, start, _dp(a), _dp(c), len(a), bump,
                             valuation_time)


def dv01_bucketed(pillars, zeros, notional, fixed_rate, start, pay, acc, bump=1e-4):
    p, z, a, c = _f64(pillars), _f64(zeros), _f64(pay), _f64(acc)
    out = np.zeros(len(p))
    lib().eqr_dv01_bucketed(_dp(p), _dp(z), len(p), notional, fixed_rate, start, _dp(a), _dp(c),
                            len(a), bump, _dp(out))
    return out


# --------------------------------------------------------------------------- portfolio
def portfolio_exposure(model, port, dates, n_steps, z, scheme=0, hazard_rate=0.0,
                       recovery_rate=0.0, want_trade=False, want_path=False, n_threads=0):
    """Per-trade restatement of the netting-set exposure run.

    ``model`` = (a, b, sigma, r0) for HW1F or (a, b, sigma, eta, rho, r0) for G2++; ``port`` = dict
    with set_id, notional, fixed_rate, start, flow_off, pay, acc (numpy arrays); ``z`` = normals
    [n_steps][n_paths] (HW1F) or [2][n_steps][n_paths] (G2++).
    """
    two = len(model) == 6
    set_id = np.ascontiguousarray(port["set_id"], dtype=np.int32)
    flow_off = np.ascontiguousarray(port["flow_off"], dtype=np.int32)
    notional, fixed, start = _f64(port["notional"]), _f64(port["fixed_rate"]), _f64(port["start"])
    pay, acc = _f64(port["pay"]), _f64(port["acc"])
    n_trades = len(notional)
    n_sets = int(set_id.max()) + 1 if n_trades else 0
    dates = _f64(dates)
    z = _f64(z)
    n_paths = z.shape[-1]
    nd = len(dates)
    out = {
        "set_ee": np.zeros((n_sets, nd)), "set_pfe": np.zeros((n_sets, nd)),
        "set_cva": np.zeros(n_sets),
    }
    if want_trade:
        out.update(trade_ee=np.zeros((n_trades, nd)), trade_pfe=np.zeros((n_trades, nd)),
                   trade_cva=np.zeros(n_trades))
    if want_path:
        out.update(path_mtm=np.zeros((nd, n_trades, n_paths)), set_mtm=np.zeros((nd, n_sets, n_paths)))
    fn = lib().eqr_portfolio_exposure2 if two else lib().eqr_portfolio_exposure
    fn(
        *model, n_trades, n_sets, set_id.ctypes.data_as(_PI32), _dp(notional), _dp(fixed),
        _dp(start), flow_off.ctypes.data_as(_PI32), _dp(pay), _dp(acc), _dp(dates), nd, n_steps,
        n_paths, scheme, _dp(z), hazard_rate, recovery_rate, _dp(out["set_ee"]),
        _dp(out["set_pfe"]), _dp(out["set_cva"]), _dp(out.get("trade_ee")),
        _dp(out.get("trade_pfe")), _dp(out.get("trade_cva")), _dp(out.get("path_mtm")),
        _dp(out.get("set_mtm")), n_threads)
    return out


def portfolio_exposure_fused(model, port, dates, n_steps, z=None, n_paths=None, seed=0, scheme=0, hazard_rate=0.0,
                             recovery_rate=0.0, want_trade=False, share_classes=False, want_set_mtm=False,
                             n_threads=0):
    """The netting-set exposure run (HW1F) as a fast CPU engine organises it (eq_ref_fused.c): one
    simulation, each distinct zero-coupon bond once per (path, date), OpenMP over paths, PFE95 by
    selection. ``z`` = normals [n_steps][n_paths], or None to draw the Philox stream of ``seed`` for
    ``n_paths`` paths inside the run. ``share_classes``: trades of one schedule share their legs (values
    then agree with portfolio_exposure to rounding; without it set values are bit-identical)."""
    set_id = np.ascontiguousarray(port["set_id"], dtype=np.int32)
    flow_off = np.ascontiguousarray(port["flow_off"], dtype=np.int32)
    notional, fixed, start = _f64(port["notional"]), _f64(port["fixed_rate"]), _f64(port["start"])
    pay, acc = _f64(port["pay"]), _f64(port["acc"])
    n_trades = len(notional)
    n_sets = int(set_id.max()) + 1
    dates = _f64(dates)
    if z is not None:
        z = _f64(z)
        n_paths = z.shape[1]
    nd = len(dates)
    out = {"set_ee": np.zeros((n_sets, nd)), "set_pfe": np.zeros((n_sets, nd)), "set_cva": np.zeros(n_sets)}
    if want_trade:
        out.update(trade_ee=np.zeros((n_trades, nd)), trade_cva=np.zeros(n_trades))
    if want_set_mtm:
        out["set_mtm"] = np.zeros((nd, n_sets, int(n_paths)))
    a, b, sigma, r0 = model
    rc = lib().eqr_portfolio_exposure_fused(
        a, b, sigma, r0, n_trades, n_sets, set_id.ctypes.data_as(_PI32), _dp(notional), _dp(fixed), _dp(start),
        flow_off.ctypes.data_as(_PI32), _dp(pay), _dp(acc), _dp(dates), nd, n_steps, int(n_paths), scheme, _dp(z),
        int(seed), hazard_rate, recovery_rate, int(bool(share_classes)), _dp(out["set_ee"]), _dp(out["set_pfe"]),
        _dp(out["set_cva"]), _dp(out.get("trade_ee")), _dp(out.get("trade_cva")), _dp(out.get("set_mtm")), n_threads)
    if rc != 0:
        raise ValueError("portfolio_exposure_fused: empty portfolio, dates or paths")
    return out


def max_threads():
    return lib().eqr_max_threads()


# --------------------------------------------------------------------------- Bermudan swaptions
def _i32(x):
    return np.ascontiguousarray(np.asarray(x, dtype=np.int32))


def _book(book):
    return (_f64(book["notional"]), _f64(book["strike"]), _i32(book["first_ex"]), _i32(book["n_ex"]))


def berm_simulate(model, dt, z):
    """Exact-scheme short rate and trapezoid log-numeraire (hull_white.rs:114-120) on one grid:
    z [n_steps][n_paths] -> (r, lb), each [n_steps+1][n_paths]."""
    a, b, sigma, r0 = model
    z = _f64(z)
    n_steps, n_paths = z.shape
    r = np.empty((n_steps + 1, n_paths))
    lb = np.empty((n_steps + 1, n_paths))
    lib().eqr_berm_simulate(a, b, sigma, r0, dt, n_steps, n_paths, _dp(z), _dp(r), _dp(lb))
    return r, lb


def berm_x(model, t, r):
    a, b, sigma, r0 = model
    return lib().eqr_berm_x(a, b, sigma, r0, t, r)


def berm_policy(model, book, steps_per_year, r, lb):
    """Longstaff-Schwartz exercise policy (degree-3 basis in the standardised short rate) fitted on
    the given paths; returns ex_coef [n][max_ex][4], tau (exercise year, -1 never) and y0 (pathwise
    cashflow discounted to 0) [n][n_paths], and the in-sample price [n]."""
    a, b, sigma, r0 = model
    notional, strike, first_ex, n_ex = _book(book)
    n = len(notional)
    max_ex = int(n_ex.max())
    r, lb = _f64(r), _f64(lb)
    n_paths = r.shape[1]
    coef = np.zeros((n, max_ex, 4))
    tau = np.zeros((n, n_paths), dtype=np.int32)
    y0 = np.zeros((n, n_paths))
    price = np.zeros(n)
    lib().eqr_berm_policy(a, b, sigma, r0, n, _dp(notional), _dp(strike), first_ex.ctypes.data_as(_PI32),
                          n_ex.ctypes.data_as(_PI32), max_ex, steps_per_year, n_paths, _dp(r), _dp(lb),
                          _dp(coef), tau.ctypes.data_as(_PI32), _dp(y0), _dp(price))
    return {"ex_coef": coef, "tau": tau, "y0": y0, "price": price}


def berm_apply(model, book, steps_per_year, r, lb, ex_coef):
    """Forward application of a fitted policy to other paths: (tau, y0), each [n][n_paths]."""
    a, b, sigma, r0 = model
    notional, strike, first_ex, n_ex = _book(book)
    n = len(notional)
    ex_coef = _f64(ex_coef)
    r, lb = _f64(r), _f64(lb)
    n_paths = r.shape[1]
    tau = np.zeros((n, n_paths), dtype=np.int32)
    y0 = np.zeros((n, n_paths))
    lib().eqr_berm_apply(a, b, sigma, r0, n, _dp(notional), _dp(strike), first_ex.ctypes.data_as(_PI32),
                         n_ex.ctypes.data_as(_PI32), ex_coef.shape[1], steps_per_year, n_paths, _dp(r), _dp(lb),
                         _dp(ex_coef), tau.ctypes.data_as(_PI32), _dp(y0))
    return tau, y0


def berm_judge(model, steps_per_year, r, lb, tau, y0, step):
    """Continuation regression of one trade at grid step `step` over its alive paths:
    (coef[4], n_alive, EE = mean over all paths of alive * max(C(x), 0))."""
    a, b, sigma, r0 = model
    r, lb = _f64(r), _f64(lb)
    tau, y0 = _i32(tau), _f64(y0)
    n_paths = r.shape[1]
    t = step / float(steps_per_year)
    coef = np.zeros(4)
    n = lib().eqr_berm_judge_fit(a, b, sigma, r0, n_paths, step, t, _dp(r), _dp(lb),
                                 tau.ctypes.data_as(_PI32), _dp(y0), _dp(coef))
    ee = lib().eqr_berm_ee(a, b, sigma, r0, n_paths, step, t, _dp(r), tau.ctypes.data_as(_PI32), _dp(coef))
    return coef, int(n), ee
