"""ctypes binding of the ``eqhip_berm_*`` entry points (include/eqhip.h): a book of Bermudan payer
swaptions under HW1F, its Longstaff-Schwartz exercise policy, the polynomial exposure judge and the
BF16 continuation networks trained and evaluated on the GPU. Like ``eqhip``, it loads the in-tree
``libeqhip.so`` and has no CPU fallback.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

from . import eqhip
from .eqhip import ModelHW1F, _check, _dp, _f64

_D, _F, _I32, _U32, _U64 = ctypes.c_double, ctypes.c_float, ctypes.c_int32, ctypes.c_uint32, ctypes.c_uint64
_PD, _PF, _PI32 = ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_int32)
_VP = ctypes.c_void_p

N_PARAMS = 8641


class BookC(ctypes.Structure):
    _fields_ = [("n_trades", _I32), ("notional", _PD), ("strike", _PD), ("first_exercise", _PI32),
                ("n_exercise", _PI32)]


class BermMcC(ctypes.Structure):
    _fields_ = [("n_train", _U64), ("n_infer", _U64), ("steps_per_year", _U32), ("horizon_years", _U32),
                ("seed_train", _U64), ("seed_infer", _U64), ("rng", _I32)]


class MlpCfgC(ctypes.Structure):
    _fields_ = [("steps", _U32), ("batch", _U32), ("lr", _F), ("lr_final", _F), ("beta1", _F), ("beta2", _F),
                ("eps", _F), ("seed", _U64), ("r_shift", _D), ("r_scale", _D), ("lb_scale", _D),
                ("lb_shift", _D), ("t_scale", _D), ("t_shift", _D), ("y_scale", _D)]


class BermTimingC(ctypes.Structure):
    _fields_ = [("simulate_ms", _D), ("policy_ms", _D), ("apply_ms", _D), ("judge_ms", _D), ("train_ms", _D),
                ("infer_ms", _D), ("lattice_ms", _D), ("lattice_net_ms", _D), ("lattice_interp_ms", _D)]


_ready = False


def _lib():
    global _ready
    L = eqhip.lib()
    if not _ready:
        L.eqhip_berm_create.argtypes = [_VP, ctypes.POINTER(ModelHW1F), ctypes.POINTER(BookC),
                                        ctypes.POINTER(BermMcC), _PD, _PD, ctypes.POINTER(_VP)]
        L.eqhip_berm_fit_policy.argtypes = [_VP, _PD, _PD, _PD]
        L.eqhip_berm_pathwise.argtypes = [_VP, _I32, _PI32, _PD]
        L.eqhip_berm_exposure_lsm.argtypes = [_VP, _I32, _PD, _PD]
        L.eqhip_berm_train.argtypes = [_VP, ctypes.POINTER(MlpCfgC), _PF]
        L.eqhip_berm_exposure_mlp.argtypes = [_VP, _I32, _PD]
        L.eqhip_berm_exposure_mlp_lattice.argtypes = [_VP, _I32, _I32, _I32, _PD]
        L.eqhip_berm_mse.argtypes = [_VP, _PD, _PD]
        L.eqhip_berm_get_params.argtypes = [_VP, _PF]
        L.eqhip_berm_set_params.argtypes = [_VP, ctypes.POINTER(MlpCfgC), _PF]
        L.eqhip_berm_mlp_grad.argtypes = [_VP, _PF, _PF, _PF, _U32, _PF, _PF]
        L.eqhip_berm_mlp_forward.argtypes = [_VP, _PF, _U32, _PF]
        L.eqhip_berm_adam_step.argtypes = [_VP, _PF, _PF, _PF, _U32, _U32, _F]
        L.eqhip_berm_get_timing.argtypes = [_VP, ctypes.POINTER(BermTimingC)]
        L.eqhip_berm_destroy.argtypes = [_VP]
        _ready = True
    return L


def _fp(x):
    return None if x is None else x.ctypes.data_as(_PF)


def _f32(x):
    return np.ascontiguousarray(np.asarray(x, dtype=np.float32))


@dataclass
class MlpConfig:
    """Training settings and the fixed input / output scaling of the continuation networks."""
    steps: int = 1500
    batch: int = 1024
    lr: float = 2e-3
    lr_final: float = 5e-5
    beta1: float = 0.9
    beta2: float = 0.999
    eps: float = 1e-8
    seed: int = 17
    r_shift: float = 0.03
    r_scale: float = 22.36
    lb_scale: float = 2.0 / 0.9
    lb_shift: float = -1.0
    t_scale: float = 2.0 / 30.0
    t_shift: float = -1.0
    y_scale: float = 10.0

    @classmethod
    def for_model(cls, model, horizon_years, **kw):
        """Scales from the model: r about its mean level in units of twice its stationary deviation,
        the log-numeraire and time mapped to about [-1, 1] over the horizon."""
        a, b, sigma, _ = model
        sd = sigma / np.sqrt(2.0 * a)
        base = dict(r_shift=b, r_scale=1.0 / (2.0 * sd), lb_scale=2.0 / max(b * horizon_years, 1e-3), lb_shift=-1.0,
                    t_scale=2.0 / horizon_years, t_shift=-1.0)
        base.update(kw)
        return cls(**base)

    def c(self):
        return MlpCfgC(self.steps, self.batch, self.lr, self.lr_final, self.beta1, self.beta2, self.eps, self.seed,
                       self.r_shift, self.r_scale, self.lb_scale, self.lb_shift, self.t_scale, self.t_shift,
                       self.y_scale)

    def inputs(self, r, lb, t):
        """Network inputs as the kernels form them: FP64 affine maps rounded to float32."""
        r, lb, t = np.broadcast_arrays(np.asarray(r, dtype=np.float64), np.asarray(lb, dtype=np.float64),
                                       np.asarray(t, dtype=np.float64))
        u0 = ((r - self.r_shift) * self.r_scale).astype(np.float32)
        u1 = (lb * self.lb_scale + self.lb_shift).astype(np.float32)
        u2 = (t * self.t_scale + self.t_shift).astype(np.float32)
        return u0, u1, u2


class BermudanBook:
    """Resident path sets (training and inference) of one Bermudan book on one context."""

    def __init__(self, ctx, model, notional, strike, first_ex, n_ex, n_train, n_infer, steps_per_year=4,
                 horizon_years=30, seed_train=101, seed_infer=202, z_train=None, z_infer=None):
        self.notional, self.strike = _f64(notional), _f64(strike)
        self.first_ex = np.ascontiguousarray(first_ex, dtype=np.int32)
        self.n_ex = np.ascontiguousarray(n_ex, dtype=np.int32)
        self.n_trades = len(self.notional)
        self.spy, self.horizon = steps_per_year, horizon_years
        self.n_steps = steps_per_year * horizon_years
        self.n_train, self.n_infer = n_train, n_infer
        self.max_ex = int(self.n_ex.max())
        self.model = tuple(model)
        book = BookC(self.n_trades, _dp(self.notional), _dp(self.strike), self.first_ex.ctypes.data_as(_PI32),
                     self.n_ex.ctypes.data_as(_PI32))
        host = z_train is not None
        mc = BermMcC(n_train, n_infer, steps_per_year, horizon_years, seed_train, seed_infer,
                     eqhip.RNG_HOST_DRAWS if host else eqhip.RNG_PHILOX)
        zt = _f64(z_train) if host else None
        zi = _f64(z_infer) if host else None
        h = ctypes.c_void_p()
        _check(_lib().eqhip_berm_create(ctx.handle, ctypes.byref(ModelHW1F(*model)), ctypes.byref(book),
                                        ctypes.byref(mc), _dp(zt), _dp(zi), ctypes.byref(h)))
        self.handle = h
        self.ctx = ctx  # the context must outlive the book
        self.cfg = None

    def close(self):
        if getattr(self, "handle", None):
            _lib().eqhip_berm_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def fit_policy(self):
        coef = np.zeros((self.n_trades, self.max_ex, 4))
        pt, pi = np.zeros(self.n_trades), np.zeros(self.n_trades)
        _check(_lib().eqhip_berm_fit_policy(self.handle, _dp(coef), _dp(pt), _dp(pi)))
        return {"ex_coef": coef, "price_train": pt, "price_infer": pi}

    def pathwise(self, set_index, want_y0=True):
        n = self.n_train if set_index == 0 else self.n_infer
        tau = np.zeros((self.n_trades, n), dtype=np.int32)
        y0 = np.zeros((self.n_trades, n)) if want_y0 else None
        _check(_lib().eqhip_berm_pathwise(self.handle, set_index, tau.ctypes.data_as(_PI32), _dp(y0)))
        return tau, y0

    def exposure_lsm(self, set_index=1):
        ee = np.zeros((self.n_trades, self.n_steps))
        coef = np.zeros((self.n_trades, self.n_steps, 4))
        _check(_lib().eqhip_berm_exposure_lsm(self.handle, set_index, _dp(ee), _dp(coef)))
        return ee, coef

    def train(self, cfg: MlpConfig):
        self.cfg = cfg
        curve = np.zeros(cfg.steps, dtype=np.float32)
        _check(_lib().eqhip_berm_train(self.handle, ctypes.byref(cfg.c()), _fp(curve)))
        return curve

    def exposure_mlp(self, set_index=1, lattice=None):
        """Network EE [n_trades][n_steps]: every alive path through the network, or with
        lattice=(nr, nl) the network on a per-date node lattice and bilinear interpolation per path."""
        ee = np.zeros((self.n_trades, self.n_steps))
        if lattice is None:
            _check(_lib().eqhip_berm_exposure_mlp(self.handle, set_index, _dp(ee)))
        else:
            nr, nl = lattice
            _check(_lib().eqhip_berm_exposure_mlp_lattice(self.handle, set_index, int(nr), int(nl), _dp(ee)))
        return ee

    def mse(self):
        a, b = np.zeros(self.n_trades), np.zeros(self.n_trades)
        _check(_lib().eqhip_berm_mse(self.handle, _dp(a), _dp(b)))
        return a, b

    def get_params(self):
        p = np.zeros((self.n_trades, N_PARAMS), dtype=np.float32)
        _check(_lib().eqhip_berm_get_params(self.handle, _fp(p)))
        return p

    def set_params(self, cfg: MlpConfig, params):
        self.cfg = cfg
        p = _f32(params).reshape(self.n_trades, N_PARAMS)
        _check(_lib().eqhip_berm_set_params(self.handle, ctypes.byref(cfg.c()), _fp(p)))

    def mlp_grad(self, u, target, mask):
        u, target, mask = _f32(u), _f32(target), _f32(mask)
        n = target.shape[1]
        loss = np.zeros(self.n_trades, dtype=np.float32)
        grad = np.zeros((self.n_trades, N_PARAMS), dtype=np.float32)
        _check(_lib().eqhip_berm_mlp_grad(self.handle, _fp(u), _fp(target), _fp(mask), n, _fp(loss), _fp(grad)))
        return loss, grad

    def adam_step(self, u, target, mask, step, lr):
        u, target, mask = _f32(u), _f32(target), _f32(mask)
        _check(_lib().eqhip_berm_adam_step(self.handle, _fp(u), _fp(target), _fp(mask), target.shape[1], step, lr))

    def mlp_forward(self, u):
        u = _f32(u)
        n = u.shape[1]
        y = np.zeros((self.n_trades, n), dtype=np.float32)
        _check(_lib().eqhip_berm_mlp_forward(self.handle, _fp(u), n, _fp(y)))
        return y

    def timing(self):
        t = BermTimingC()
        _check(_lib().eqhip_berm_get_timing(self.handle, ctypes.byref(t)))
        return {k: getattr(t, k) for k, _ in BermTimingC._fields_}
