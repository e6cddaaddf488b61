"""ctypes binding of the C ABI in ``include/eqhip.h`` (libeqhip.so, HIP kernels for gfx950).

This is the binding a Python maintainer of the reference would add beside its nanobind module;
it loads the in-tree ``libeqhip.so`` and fails loudly when the library is missing. It never falls
back to a CPU implementation.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass

import numpy as np

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_PKG, "libeqhip.so")

OK, E_INVALID, E_NO_DEVICE, E_HIP, E_CAPACITY = 0, 1, 2, 3, 4
RNG_HOST_DRAWS, RNG_PHILOX = 0, 1
DISC_EULER, DISC_EXACT = 0, 1

_D, _I32, _U32, _U64 = ctypes.c_double, ctypes.c_int32, ctypes.c_uint32, ctypes.c_uint64
_PD = ctypes.POINTER(ctypes.c_double)
_PI32 = ctypes.POINTER(ctypes.c_int32)

# Every symbol declared in include/eqhip.h.
EXPORTED = [
    "eqhip_is_gpu_backend_available", "eqhip_last_error", "eqhip_context_create",
    "eqhip_context_destroy", "eqhip_irs_hull_white_exposure_profile",
    "eqhip_irs_hull_white_exposure_profile_batch", "eqhip_unilateral_cva_from_exposure",
    "eqhip_unilateral_cva_from_exposure_batch", "eqhip_pv_dv01", "eqhip_plan_create",
    "eqhip_plan_execute", "eqhip_plan_timing", "eqhip_plan_destroy", "eqhip_exposure_run",
    "eqhip_batch_create", "eqhip_batch_execute", "eqhip_batch_timing", "eqhip_batch_destroy",
    "eqhip_cva_adjoint", "eqhip_pv_sensitivities", "eqhip_berm_create", "eqhip_berm_fit_policy", "eqhip_berm_pathwise",
    "eqhip_berm_exposure_lsm", "eqhip_berm_train", "eqhip_berm_exposure_mlp", "eqhip_berm_exposure_mlp_lattice",
    "eqhip_berm_mse",
    "eqhip_berm_get_params", "eqhip_berm_set_params", "eqhip_berm_mlp_grad", "eqhip_berm_mlp_forward",
    "eqhip_berm_adam_step", "eqhip_berm_get_timing", "eqhip_berm_destroy",
    "eqhip_irs_hull_white_2f_exposure_profile", "eqhip_irs_hull_white_2f_exposure_profile_batch",
    "eqhip_unilateral_cva_from_exposure_2f", "eqhip_unilateral_cva_from_exposure_2f_batch", "eqhip_plan_create_2f",
    "eqhip_xva_calculate", "eqhip_plan_exposure_cube", "eqhip_plan_xva", "eqhip_scenarios_create",
    "eqhip_scenarios_destroy", "eqhip_irs_hull_white_exposure_profile_batch_on", "eqhip_pv_dv01_at",
    "eqhip_unilateral_cva_from_exposure_batch_noalloc", "eqhip_plan_set_variant", "eqhip_bond_exp",
]

PLAN_TRADE_EE, PLAN_KEEP_CUBE, PLAN_AUTOTUNE, PLAN_TRADE_PFE = 1, 2, 4, 8  # flags of eqhip_plan_create's
# want_trade_ee argument


class EqhipError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"eqhip error {code}: {msg}")
        self.code = code


class ModelHW1F(ctypes.Structure):
    _fields_ = [("a", _D), ("b", _D), ("sigma", _D), ("r0", _D)]


class ModelHW2F(ctypes.Structure):
    _fields_ = [("a", _D), ("b", _D), ("sigma", _D), ("eta", _D), ("rho", _D), ("r0", _D)]


class PortfolioC(ctypes.Structure):
    _fields_ = [("n_trades", _I32), ("n_sets", _I32), ("set_id", _PI32), ("notional", _PD),
                ("fixed_rate", _PD), ("start", _PD), ("flow_offset", _PI32),
                ("payment_times", _PD), ("accruals", _PD)]


class McC(ctypes.Structure):
    _fields_ = [("n_paths", _U64), ("n_steps", _U32), ("seed", _U64), ("rng", _I32), ("disc", _I32)]


class CreditC(ctypes.Structure):
    _fields_ = [("hazard_rate", _D), ("recovery_rate", _D)]


class OutputsC(ctypes.Structure):
    _fields_ = [("set_ee", _PD), ("set_pfe", _PD), ("set_cva", _PD), ("trade_ee", _PD),
                ("trade_cva", _PD), ("path_mtm", _PD), ("set_mtm", _PD), ("trade_pfe", _PD)]


class CsaC(ctypes.Structure):
    _fields_ = [("threshold", _D), ("minimum_transfer_amount", _D), ("independent_amount", _D),
                ("variation_margin", _I32)]


class DefaultSpecC(ctypes.Structure):
    _fields_ = [("hazard_rate", _D), ("recovery_rate", _D), ("survival_probabilities", _PD), ("n_survival", _I32)]


class XvaParamsC(ctypes.Structure):
    _fields_ = [("counterparty_default", DefaultSpecC), ("own_default", DefaultSpecC), ("funding_spread", _D),
                ("borrowing_spread", _D), ("capital_factor", _D), ("initial_margin_total", _D),
                ("csa", ctypes.POINTER(CsaC)), ("confidence_level", _D)]


class TimingC(ctypes.Structure):
    _fields_ = [("simulate_ms", _D), ("value_ms", _D), ("reduce_ms", _D), ("cva_ms", _D),
                ("total_ms", _D), ("n_value_blocks", _U64), ("paths_per_tile", _U32),
                ("tiles_per_block", _U32), ("xva_rows_ms", _D), ("xva_select_ms", _D), ("trades_per_item", _U32),
                ("n_parts", _U32), ("trade_pfe_ms", _D), ("trade_pfe_groups", _U32)]


_lib = None


def lib():
    """Load libeqhip.so (raises if it was not built: there is no fallback)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: run __graft_entry__.build() (hipcc, gfx950)")
        L = ctypes.CDLL(LIB_PATH)
        L.eqhip_is_gpu_backend_available.restype = ctypes.c_int
        L.eqhip_last_error.restype = ctypes.c_size_t
        L.eqhip_last_error.argtypes = [ctypes.c_char_p, ctypes.c_size_t]
        L.eqhip_context_create.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_void_p)]
        L.eqhip_context_destroy.argtypes = [ctypes.c_void_p]
        prof = [ctypes.c_void_p, _D, _D, _D, _D]
        tail = [_D, _PD, _PD, _I32, _PD, _I32, _U32, _U64, _U64, _I32, _I32, _PD, _I32, _PD, _PD, _PD, _PI32]
        L.eqhip_irs_hull_white_exposure_profile.argtypes = prof + [_D, _D, ctypes.c_int] + tail
        L.eqhip_irs_hull_white_exposure_profile_batch.argtypes = prof + [_PD, _PD, _I32] + tail
        L.eqhip_unilateral_cva_from_exposure.argtypes = prof + [_PD, _PD, _I32, _D, _D, _PD]
        L.eqhip_unilateral_cva_from_exposure_batch.argtypes = prof + [_PD, _PD, _I32, _I32, _D, _D, _PD]
        L.eqhip_unilateral_cva_from_exposure_batch_noalloc.argtypes = prof + [
            ctypes.c_void_p, ctypes.c_void_p, _I32, _I32, _D, _D, ctypes.c_void_p, ctypes.c_void_p]
        prof2 = [ctypes.c_void_p, _D, _D, _D, _D, _D, _D]
        L.eqhip_irs_hull_white_2f_exposure_profile.argtypes = prof2 + [_D, _D, ctypes.c_int] + tail
        L.eqhip_irs_hull_white_2f_exposure_profile_batch.argtypes = prof2 + [_PD, _PD, _I32] + tail
        L.eqhip_unilateral_cva_from_exposure_2f.argtypes = prof2 + [_PD, _PD, _I32, _D, _D, _PD]
        L.eqhip_unilateral_cva_from_exposure_2f_batch.argtypes = prof2 + [_PD, _PD, _I32, _I32, _D, _D, _PD]
        L.eqhip_plan_create_2f.argtypes = [ctypes.c_void_p, ctypes.POINTER(ModelHW2F), ctypes.POINTER(PortfolioC),
                                           _PD, _I32, ctypes.POINTER(McC), _I32, ctypes.POINTER(ctypes.c_void_p)]
        L.eqhip_xva_calculate.argtypes = [ctypes.c_void_p, ctypes.POINTER(XvaParamsC), _PD, _I32, _U64, _PD, _PD,
                                          _PD, _PD, _PD, _PD]
        L.eqhip_plan_exposure_cube.argtypes = [ctypes.c_void_p, _I32, _PD, _PD]
        L.eqhip_scenarios_create.argtypes = [ctypes.c_void_p, ctypes.POINTER(ModelHW1F), _D, ctypes.POINTER(McC),
                                             ctypes.POINTER(ctypes.c_void_p)]
        L.eqhip_scenarios_destroy.argtypes = [ctypes.c_void_p]
        L.eqhip_irs_hull_white_exposure_profile_batch_on.argtypes = [
            ctypes.c_void_p, ctypes.c_void_p, _PD, _PD, _I32, _D, _PD, _PD, _I32, _PD, _I32, _PD, _PD, _PD, _PI32]
        L.eqhip_plan_xva.argtypes = [ctypes.c_void_p, _I32, ctypes.POINTER(XvaParamsC), _PD, _PD, _PD]
        L.eqhip_pv_dv01.argtypes = [ctypes.c_void_p, _PD, _PD, _I32, ctypes.POINTER(PortfolioC), _D, _PD,
                                    _PD, _PD]
        L.eqhip_pv_dv01_at.argtypes = [ctypes.c_void_p, _PD, _PD, _I32, ctypes.POINTER(PortfolioC), _D, _D, _PD,
                                       _PD, _PD]
        L.eqhip_bond_exp.argtypes = [ctypes.c_void_p, _PD, _PD, _I32]
        L.eqhip_plan_create.argtypes = [ctypes.c_void_p, ctypes.POINTER(ModelHW1F),
                                        ctypes.POINTER(PortfolioC), _PD, _I32, ctypes.POINTER(McC),
                                        _I32, ctypes.POINTER(ctypes.c_void_p)]
        L.eqhip_plan_execute.argtypes = [ctypes.c_void_p, _PD, ctypes.POINTER(CreditC),
                                         ctypes.POINTER(OutputsC)]
        L.eqhip_plan_timing.argtypes = [ctypes.c_void_p, ctypes.POINTER(TimingC)]
        L.eqhip_plan_destroy.argtypes = [ctypes.c_void_p]
        L.eqhip_plan_set_variant.argtypes = [ctypes.c_void_p, _I32, _I32]
        L.eqhip_exposure_run.argtypes = [ctypes.c_void_p, ctypes.POINTER(ModelHW1F),
                                         ctypes.POINTER(PortfolioC), _PD, _I32, ctypes.POINTER(McC), _PD,
                                         ctypes.POINTER(CreditC), ctypes.POINTER(OutputsC)]
        L.eqhip_batch_create.argtypes = [ctypes.c_void_p, ctypes.POINTER(ModelHW1F), _I32,
                                         ctypes.POINTER(PortfolioC), ctypes.POINTER(ctypes.c_uint64), _PD, _I32,
                                         _U32, _U64, _I32, _I32, _I32, ctypes.POINTER(ctypes.c_void_p)]
        L.eqhip_batch_execute.argtypes = [ctypes.c_void_p, ctypes.POINTER(CreditC), ctypes.POINTER(OutputsC)]
        L.eqhip_batch_timing.argtypes = [ctypes.c_void_p, ctypes.POINTER(TimingC), ctypes.c_char_p,
                                         ctypes.c_size_t]
        L.eqhip_batch_destroy.argtypes = [ctypes.c_void_p]
        L.eqhip_cva_adjoint.argtypes = [ctypes.c_void_p, _PD, _I32, ctypes.POINTER(CreditC), _PD, _PD, _PD]
        L.eqhip_pv_sensitivities.argtypes = [ctypes.c_void_p, _PD, _PD, _I32, ctypes.POINTER(PortfolioC),
                                             ctypes.POINTER(ModelHW1F), _PD, _PD, _PD]
        _lib = L
    return _lib


def last_error() -> str:
    buf = ctypes.create_string_buffer(4096)
    lib().eqhip_last_error(buf, len(buf))
    return buf.value.decode(errors="replace")


def _check(code):
    if code != OK:
        raise EqhipError(code, last_error())


def is_gpu_backend_available() -> bool:
    return bool(lib().eqhip_is_gpu_backend_available())


def _f64(x):
    return np.ascontiguousarray(np.asarray(x, dtype=np.float64))


def _dp(x):
    return None if x is None else x.ctypes.data_as(_PD)


@dataclass
class Portfolio:
    """SoA IRS portfolio (receiver swaps carry a negative notional)."""
    set_id: np.ndarray
    notional: np.ndarray
    fixed_rate: np.ndarray
    start: np.ndarray
    flow_off: np.ndarray
    pay: np.ndarray
    acc: np.ndarray

    @property
    def n_trades(self):
        return len(self.notional)

    @property
    def n_sets(self):
        return int(self.set_id.max()) + 1 if len(self.set_id) else 0

    def as_dict(self):
        return dict(set_id=self.set_id, notional=self.notional, fixed_rate=self.fixed_rate,
                    start=self.start, flow_off=self.flow_off, pay=self.pay, acc=self.acc)

    def _c(self):
        self.set_id = np.ascontiguousarray(self.set_id, dtype=np.int32)
        self.flow_off = np.ascontiguousarray(self.flow_off, dtype=np.int32)
        for name in ("notional", "fixed_rate", "start", "pay", "acc"):
            setattr(self, name, _f64(getattr(self, name)))
        return PortfolioC(self.n_trades, self.n_sets, self.set_id.ctypes.data_as(_PI32),
                          _dp(self.notional), _dp(self.fixed_rate), _dp(self.start),
                          self.flow_off.ctypes.data_as(_PI32), _dp(self.pay), _dp(self.acc))


class Scenarios:
    """Device-resident HW1F scenario set (eqhip_scenarios_create / destroy): the short rate at every grid
    step of t_max / n_steps on n_paths Philox paths, valued on by Context.exposure_profile_batch_on."""

    def __init__(self, ctx, model, t_max, n_paths, n_steps, seed=0, disc=DISC_EULER):
        self.ctx = ctx
        self.handle = ctypes.c_void_p()
        m = ModelHW1F(*model)
        mc = McC(n_paths, n_steps, seed, RNG_PHILOX, disc)
        _check(lib().eqhip_scenarios_create(ctx.handle, ctypes.byref(m), float(t_max), ctypes.byref(mc),
                                            ctypes.byref(self.handle)))

    def close(self):
        if self.handle:
            lib().eqhip_scenarios_destroy(self.handle)
            self.handle = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Plan:
    """Device-resident exposure job (eqhip_plan_create / execute / destroy).

    ``model`` is (a, b, sigma, r0) for HullWhite1F or (a, b, sigma, eta, rho, r0) for HullWhite2F
    (eqhip_plan_create_2f); a two-factor plan takes host draws [2][n_steps][n_paths].
    """

    def __init__(self, ctx, model, portfolio: Portfolio, dates, n_paths, n_steps, seed=0,
                 rng=RNG_PHILOX, disc=DISC_EULER, want_trade_ee=False, keep_cube=False, autotune=False,
                 want_trade_pfe=False):
        self.ctx = ctx
        self.portfolio = portfolio
        self.dates = _f64(dates)
        self.n_paths, self.n_steps = int(n_paths), int(n_steps)
        self.rng = rng
        self.want_trade_ee = bool(want_trade_ee)
        self.want_trade_pfe = bool(want_trade_pfe)
        self.keep_cube = bool(keep_cube)
        self._pc = portfolio._c()
        self.factors = 2 if len(model) == 6 else 1
        m = ModelHW2F(*model) if self.factors == 2 else ModelHW1F(*model)
        create = lib().eqhip_plan_create_2f if self.factors == 2 else lib().eqhip_plan_create
        mc = McC(self.n_paths, self.n_steps, int(seed), rng, disc)
        h = ctypes.c_void_p()
        flags = ((PLAN_TRADE_EE if want_trade_ee else 0) | (PLAN_KEEP_CUBE if keep_cube else 0) |
                 (PLAN_AUTOTUNE if autotune else 0) | (PLAN_TRADE_PFE if want_trade_pfe else 0))
        _check(create(ctx.handle, ctypes.byref(m), ctypes.byref(self._pc), _dp(self.dates), len(self.dates),
                      ctypes.byref(mc), flags, ctypes.byref(h)))
        self.handle = h

    def execute(self, host_draws=None, hazard_rate=0.0, recovery_rate=0.0, want_cva=True,
                want_path=False, want_set_mtm=False):
        ns, nd, nt = self.portfolio.n_sets, len(self.dates), self.portfolio.n_trades
        res = {"set_ee": np.zeros((ns, nd)), "set_pfe": np.zeros((ns, nd))}
        if want_cva:
            res["set_cva"] = np.zeros(ns)
        if self.want_trade_ee:
            res["trade_ee"] = np.zeros((nt, nd))
            if want_cva:
                res["trade_cva"] = np.zeros(nt)
        if want_path:
            res["path_mtm"] = np.zeros((nd, nt, self.n_paths))
        if want_set_mtm:
            res["set_mtm"] = np.zeros((nd, ns, self.n_paths))
        if self.want_trade_pfe:
            res["trade_pfe"] = np.zeros((nt, nd))
        out = OutputsC(*[_dp(res.get(k)) for k in ("set_ee", "set_pfe", "set_cva", "trade_ee", "trade_cva",
                                                    "path_mtm", "set_mtm", "trade_pfe")])
        z = None
        if self.rng == RNG_HOST_DRAWS:
            z = _f64(host_draws)
            shape = (self.n_steps, self.n_paths) if self.factors == 1 else (2, self.n_steps, self.n_paths)
            if z.shape != shape:
                raise ValueError(f"host_draws must have shape {shape}")
        cr = CreditC(hazard_rate, recovery_rate)
        _check(lib().eqhip_plan_execute(self.handle, _dp(z), ctypes.byref(cr), ctypes.byref(out)))
        return res

    def cva_adjoint(self, set_index=0, hazard_rate=0.0, recovery_rate=0.0, host_draws=None):
        """CVA of one netting set and dCVA/d(a, b, sigma, r0, hazard, recovery) by pathwise adjoint."""
        z = None
        if self.rng == RNG_HOST_DRAWS:
            z = _f64(host_draws)
        cva, check = ctypes.c_double(), ctypes.c_double()
        grad = np.zeros(6)
        cr = CreditC(hazard_rate, recovery_rate)
        _check(lib().eqhip_cva_adjoint(self.handle, _dp(z), set_index, ctypes.byref(cr), ctypes.byref(cva),
                                       _dp(grad), ctypes.byref(check)))
        return cva.value, dict(zip(("a", "b", "sigma", "r0", "hazard_rate", "recovery_rate"), grad)), check.value

    def exposure_cube(self, set_index=0):
        """Signed netting-set values [n_dates][n_paths] of the last execute and the model discount
        factors P(0, t_k) (eqhip_plan_exposure_cube; the plan needs keep_cube=True)."""
        values = np.zeros((len(self.dates), self.n_paths))
        df = np.zeros(len(self.dates))
        _check(lib().eqhip_plan_exposure_cube(self.handle, set_index, _dp(values), _dp(df)))
        return values, df

    def xva(self, params, set_index=0):
        """XVA reductions of the resident cube of one netting set (eqhip_plan_xva): (ee, pfe, measures[5])."""
        nd = len(self.dates)
        ee, pfe, m = np.zeros(nd), np.zeros(nd), np.zeros(5)
        _check(lib().eqhip_plan_xva(self.handle, set_index, ctypes.byref(params), _dp(ee), _dp(pfe), _dp(m)))
        return ee, pfe, m

    def timing(self):
        t = TimingC()
        _check(lib().eqhip_plan_timing(self.handle, ctypes.byref(t)))
        return {f: getattr(t, f) for f, _ in TimingC._fields_}

    def set_variant(self, paths_per_lane, trades_per_item):
        """Valuation variant (eqhip_plan_set_variant): paths per lane 1 / 2 / 4, trades per item 2 / 4 / 8;
        paths_per_lane 0 selects path mode (one thread per path, for plans without per-trade outputs)."""
        _check(lib().eqhip_plan_set_variant(self.handle, int(paths_per_lane), int(trades_per_item)))

    def close(self):
        if getattr(self, "handle", None):
            lib().eqhip_plan_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Context:
    """One HIP stream on one gfx950 device (eqhip_context)."""

    def __init__(self, device=0):
        h = ctypes.c_void_p()
        _check(lib().eqhip_context_create(device, ctypes.byref(h)))
        self.handle = h

    def close(self):
        if getattr(self, "handle", None):
            lib().eqhip_context_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # reference-shaped entry points -------------------------------------------------------
    def exposure_profile(self, a, b, sigma, r0, notional, fixed_rate, use_par_rate, start, pay, acc,
                         times, n_steps, n_paths, seed, rng=RNG_PHILOX, disc=DISC_EULER, host_draws=None,
                         allow_any_date=False):
        pay, acc, times = _f64(pay), _f64(acc), _f64(times)
        nd = len(times)
        ot, oe, op = np.zeros(nd), np.zeros(nd), np.zeros(nd)
        n_out = ctypes.c_int32()
        z = _f64(host_draws) if host_draws is not None else None
        _check(lib().eqhip_irs_hull_white_exposure_profile(
            self.handle, a, b, sigma, r0, notional, fixed_rate, int(bool(use_par_rate)), start, _dp(pay),
            _dp(acc), len(pay), _dp(times), nd, n_steps, n_paths, seed, rng, disc, _dp(z),
            int(allow_any_date), _dp(ot), _dp(oe), _dp(op), ctypes.byref(n_out)))
        n = n_out.value
        return ot[:n], oe[:n], op[:n]

    def exposure_profile_batch(self, a, b, sigma, r0, notionals, fixed_rates, start, pay, acc, times,
                               n_steps, n_paths, seed, rng=RNG_PHILOX, disc=DISC_EULER, host_draws=None,
                               allow_any_date=False):
        notionals, fixed_rates = _f64(notionals), _f64(fixed_rates)
        pay, acc, times = _f64(pay), _f64(acc), _f64(times)
        nt, nd = len(notionals), len(times)
        ot, oe, op = np.zeros(nd), np.zeros((nt, nd)), np.zeros((nt, nd))
        n_out = ctypes.c_int32()
        z = _f64(host_draws) if host_draws is not None else None
        _check(lib().eqhip_irs_hull_white_exposure_profile_batch(
            self.handle, a, b, sigma, r0, _dp(notionals), _dp(fixed_rates), nt, start, _dp(pay), _dp(acc),
            len(pay), _dp(times), nd, n_steps, n_paths, seed, rng, disc, _dp(z), int(allow_any_date), _dp(ot),
            _dp(oe), _dp(op), ctypes.byref(n_out)))
        n = n_out.value
        return ot[:n], oe[:, :n], op[:, :n]

    def scenarios(self, model, t_max, n_paths, n_steps, seed=0, disc=DISC_EULER):
        """A scenario set (HW1F paths on the grid of t_max / n_steps, every step kept on the device)."""
        return Scenarios(self, model, t_max, n_paths, n_steps, seed, disc)

    def exposure_profile_batch_on(self, scenarios, notionals, fixed_rates, start, pay, acc, times):
        notionals, fixed_rates = _f64(notionals), _f64(fixed_rates)
        pay, acc, times = _f64(pay), _f64(acc), _f64(times)
        nt, nd = len(notionals), len(times)
        ot, oe, op = np.zeros(nd), np.zeros((nt, nd)), np.zeros((nt, nd))
        n_out = ctypes.c_int32()
        _check(lib().eqhip_irs_hull_white_exposure_profile_batch_on(
            self.handle, scenarios.handle, _dp(notionals), _dp(fixed_rates), nt, start, _dp(pay), _dp(acc), len(pay),
            _dp(times), nd, _dp(ot), _dp(oe), _dp(op), ctypes.byref(n_out)))
        n = n_out.value
        return ot[:n], oe[:, :n], op[:, :n]

    def unilateral_cva(self, a, b, sigma, r0, times, ee, hazard_rate, recovery_rate):
        times, ee = _f64(times), _f64(ee)
        out = ctypes.c_double()
        _check(lib().eqhip_unilateral_cva_from_exposure(self.handle, a, b, sigma, r0, _dp(times), _dp(ee),
                                                        len(times), hazard_rate, recovery_rate,
                                                        ctypes.byref(out)))
        return out.value

    def unilateral_cva_batch(self, a, b, sigma, r0, times, ee, hazard_rate, recovery_rate):
        times, ee = _f64(times), np.atleast_2d(_f64(ee))
        out = np.zeros(ee.shape[0])
        _check(lib().eqhip_unilateral_cva_from_exposure_batch(self.handle, a, b, sigma, r0, _dp(times),
                                                              _dp(ee), ee.shape[0], len(times), hazard_rate,
                                                              recovery_rate, _dp(out)))
        return out

    def unilateral_cva_batch_device(self, a, b, sigma, r0, d_times, d_ee, n_profiles, n_dates, hazard_rate,
                                    recovery_rate, d_out_cva, stream=None):
        """eqhip_unilateral_cva_from_exposure_batch_noalloc on device memory (integer addresses, e.g. a
        torch tensor's data_ptr()): no allocation; with a stream (an integer hipStream_t) the call only
        enqueues, without one it waits for the result."""
        _check(lib().eqhip_unilateral_cva_from_exposure_batch_noalloc(
            self.handle, a, b, sigma, r0, ctypes.c_void_p(d_times), ctypes.c_void_p(d_ee), int(n_profiles),
            int(n_dates), hazard_rate, recovery_rate, ctypes.c_void_p(d_out_cva),
            ctypes.c_void_p(stream) if stream else None))

    def exposure_profile_2f(self, a, b, sigma, eta, rho, r0, notional, fixed_rate, use_par_rate, start, pay, acc,
                            times, n_steps, n_paths, seed, rng=RNG_PHILOX, disc=DISC_EULER, host_draws=None,
                            allow_any_date=False):
        pay, acc, times = _f64(pay), _f64(acc), _f64(times)
        nd = len(times)
        ot, oe, op = np.zeros(nd), np.zeros(nd), np.zeros(nd)
        n_out = ctypes.c_int32()
        z = _f64(host_draws) if host_draws is not None else None
        _check(lib().eqhip_irs_hull_white_2f_exposure_profile(
            self.handle, a, b, sigma, eta, rho, r0, notional, fixed_rate, int(bool(use_par_rate)), start, _dp(pay),
            _dp(acc), len(pay), _dp(times), nd, n_steps, n_paths, seed, rng, disc, _dp(z), int(allow_any_date),
            _dp(ot), _dp(oe), _dp(op), ctypes.byref(n_out)))
        n = n_out.value
        return ot[:n], oe[:n], op[:n]

    def exposure_profile_2f_batch(self, a, b, sigma, eta, rho, r0, notionals, fixed_rates, start, pay, acc, times,
                                  n_steps, n_paths, seed, rng=RNG_PHILOX, disc=DISC_EULER, host_draws=None,
                                  allow_any_date=False):
        notionals, fixed_rates = _f64(notionals), _f64(fixed_rates)
        pay, acc, times = _f64(pay), _f64(acc), _f64(times)
        nt, nd = len(notionals), len(times)
        ot, oe, op = np.zeros(nd), np.zeros((nt, nd)), np.zeros((nt, nd))
        n_out = ctypes.c_int32()
        z = _f64(host_draws) if host_draws is not None else None
        _check(lib().eqhip_irs_hull_white_2f_exposure_profile_batch(
            self.handle, a, b, sigma, eta, rho, r0, _dp(notionals), _dp(fixed_rates), nt, start, _dp(pay),
            _dp(acc), len(pay), _dp(times), nd, n_steps, n_paths, seed, rng, disc, _dp(z), int(allow_any_date),
            _dp(ot), _dp(oe), _dp(op), ctypes.byref(n_out)))
        n = n_out.value
        return ot[:n], oe[:, :n], op[:, :n]

    def unilateral_cva_2f(self, a, b, sigma, eta, rho, r0, times, ee, hazard_rate, recovery_rate):
        times, ee = _f64(times), _f64(ee)
        out = ctypes.c_double()
        _check(lib().eqhip_unilateral_cva_from_exposure_2f(self.handle, a, b, sigma, eta, rho, r0, _dp(times),
                                                           _dp(ee), len(times), hazard_rate, recovery_rate,
                                                           ctypes.byref(out)))
        return out.value

    def unilateral_cva_2f_batch(self, a, b, sigma, eta, rho, r0, times, ee, hazard_rate, recovery_rate):
        times, ee = _f64(times), np.atleast_2d(_f64(ee))
        out = np.zeros(ee.shape[0])
        _check(lib().eqhip_unilateral_cva_from_exposure_2f_batch(self.handle, a, b, sigma, eta, rho, r0,
                                                                 _dp(times), _dp(ee), ee.shape[0], len(times),
                                                                 hazard_rate, recovery_rate, _dp(out)))
        return out

    def xva_calculate(self, params, times, q_paths, discount_factors=None, p_paths=None):
        """eqhip_xva_calculate on host cubes [n_times][n_paths]: (ee, pfe, measures[5])."""
        times, q = _f64(times), np.ascontiguousarray(_f64(q_paths))
        if q.ndim != 2 or q.shape[0] != len(times):
            raise ValueError("exposure times and rows must be non-empty and have equal length")
        p = None if p_paths is None else np.ascontiguousarray(_f64(p_paths))
        if p is not None and p.shape != q.shape:
            raise ValueError("Q and P exposure grids must match")
        df = None if discount_factors is None or len(discount_factors) == 0 else _f64(discount_factors)
        if df is not None and len(df) != len(times):
            raise ValueError("discount_factors must match times and be non-negative")
        nt = len(times)
        ee, pfe, m = np.zeros(nt), np.zeros(nt), np.zeros(5)
        _check(lib().eqhip_xva_calculate(self.handle, ctypes.byref(params), _dp(times), nt, q.shape[1], _dp(q),
                                         _dp(df), _dp(p), _dp(ee), _dp(pfe), _dp(m)))
        return ee, pfe, m

    def pv_dv01(self, pillars, zero_rates, portfolio: Portfolio, bump=1e-4, bucketed=True, valuation_time=None):
        """PV, parallel DV01 and (optionally) the bucketed ladder of every trade from the market curve;
        with a valuation time, PV and DV01 are valued then (eqhip_pv_dv01_at), the ladder at t = 0."""
        pillars, zero_rates = _f64(pillars), _f64(zero_rates)
        pc = portfolio._c()
        nt = portfolio.n_trades
        pv, dv = np.zeros(nt), np.zeros(nt)
        bk = np.zeros((nt, len(pillars))) if bucketed else None
        if valuation_time is None:
            _check(lib().eqhip_pv_dv01(self.handle, _dp(pillars), _dp(zero_rates), len(pillars), ctypes.byref(pc),
                                       bump, _dp(pv), _dp(dv), _dp(bk)))
        else:
            _check(lib().eqhip_pv_dv01_at(self.handle, _dp(pillars), _dp(zero_rates), len(pillars), ctypes.byref(pc),
                                          bump, float(valuation_time), _dp(pv), _dp(dv), _dp(bk)))
        return pv, dv, bk

    def bond_exp(self, x):
        """The valuation kernels' FP64 exp of every value of x, computed on the device."""
        x = _f64(x)
        y = np.empty_like(x)
        _check(lib().eqhip_bond_exp(self.handle, _dp(x), _dp(y), len(x)))
        return y

    def pv_sensitivities(self, pillars, zero_rates, portfolio: Portfolio, model=None):
        """Analytic dPV/dz per pillar [n_trades][n_pillars]; with a model also the model NPV at t = 0
        and its (a, b, sigma, r0) derivatives [n_trades][4]."""
        pillars, zero_rates = _f64(pillars), _f64(zero_rates)
        pc = portfolio._c()
        nt = portfolio.n_trades
        dz = np.zeros((nt, len(pillars)))
        npv = np.zeros(nt) if model is not None else None
        gr = np.zeros((nt, 4)) if model is not None else None
        m = ctypes.byref(ModelHW1F(*model)) if model is not None else None
        _check(lib().eqhip_pv_sensitivities(self.handle, _dp(pillars), _dp(zero_rates), len(pillars),
                                            ctypes.byref(pc), m, _dp(dz), _dp(npv), _dp(gr)))
        return dz, npv, gr

    # portfolio --------------------------------------------------------------------------------
    def plan(self, model, portfolio, dates, n_paths, n_steps, **kw):
        return Plan(self, model, portfolio, dates, n_paths, n_steps, **kw)

    def exposure_run(self, model, portfolio, dates, n_paths, n_steps, seed=0, rng=RNG_PHILOX,
                     disc=DISC_EULER, host_draws=None, hazard_rate=0.0, recovery_rate=0.0,
                     want_trade_ee=False, want_path=False, want_set_mtm=False, want_trade_pfe=False):
        plan = Plan(self, model, portfolio, dates, n_paths, n_steps, seed=seed, rng=rng, disc=disc,
                    want_trade_ee=want_trade_ee or want_path, want_trade_pfe=want_trade_pfe)
        try:
            return plan.execute(host_draws, hazard_rate, recovery_rate, True, want_path, want_set_mtm)
        finally:
            plan.close()


class Batch:
    """Many netting-set jobs in one simulation and one valuation launch (eqhip_batch_*)."""

    def __init__(self, ctx, model, jobs, dates, n_steps, seed=0, disc=DISC_EULER, want_trade_ee=False):
        """jobs: list of (Portfolio, n_paths)."""
        self.ctx = ctx
        self.jobs = jobs
        self.dates = _f64(dates)
        self.want_trade_ee = bool(want_trade_ee)
        self._pcs = (PortfolioC * len(jobs))(*[p._c() for p, _ in jobs])
        self._np = (ctypes.c_uint64 * len(jobs))(*[int(n) for _, n in jobs])
        m = ModelHW1F(*model)
        h = ctypes.c_void_p()
        _check(lib().eqhip_batch_create(ctx.handle, ctypes.byref(m), len(jobs), self._pcs, self._np,
                                        _dp(self.dates), len(self.dates), int(n_steps), int(seed), RNG_PHILOX,
                                        disc, int(want_trade_ee), ctypes.byref(h)))
        self.handle = h

    def execute(self, hazard_rate=0.0, recovery_rate=0.0):
        nd = len(self.dates)
        results, outs = [], (OutputsC * len(self.jobs))()
        for j, (p, _) in enumerate(self.jobs):
            r = {"set_ee": np.zeros((p.n_sets, nd)), "set_pfe": np.zeros((p.n_sets, nd)),
                 "set_cva": np.zeros(p.n_sets)}
            if self.want_trade_ee:
                r["trade_ee"] = np.zeros((p.n_trades, nd))
                r["trade_cva"] = np.zeros(p.n_trades)
            outs[j] = OutputsC(*[_dp(r.get(k)) for k in ("set_ee", "set_pfe", "set_cva", "trade_ee", "trade_cva",
                                                         "path_mtm", "set_mtm")])
            results.append(r)
        cr = CreditC(hazard_rate, recovery_rate)
        _check(lib().eqhip_batch_execute(self.handle, ctypes.byref(cr), outs))
        return results

    def timing(self):
        t = TimingC()
        buf = ctypes.create_string_buffer(8192)
        _check(lib().eqhip_batch_timing(self.handle, ctypes.byref(t), buf, len(buf)))
        d = {f: getattr(t, f) for f, _ in TimingC._fields_}
        d["tune_log"] = buf.value.decode()
        return d

    def close(self):
        if getattr(self, "handle", None):
            lib().eqhip_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def singleton_portfolio(notionals, fixed_rates, start, pay, acc):
    """Homogeneous batch with one netting set per trade (the reference's per-trade profiles)."""
    n = len(notionals)
    m = len(pay)
    return Portfolio(set_id=np.arange(n, dtype=np.int32), notional=_f64(notionals), fixed_rate=_f64(fixed_rates),
                     start=np.full(n, float(start)), flow_off=np.arange(0, (n + 1) * m, m, dtype=np.int32),
                     pay=np.tile(_f64(pay), n), acc=np.tile(_f64(acc), n))
