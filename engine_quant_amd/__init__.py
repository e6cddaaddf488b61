"""engine_quant_amd — MI355X-native Monte Carlo exposure path of engine_quant.

Hull-White 1F short-rate simulation, per-path IRS mark-to-market, EE / PFE95 / unilateral CVA
and PV / DV01, computed by hand-written HIP kernels for gfx950 behind the C ABI of
``include/eqhip.h`` (``libeqhip.so``) and the reference's registry-driven pricing API
(``Engine.create_model / create_product / price`` with ``backend="gpu"``).
"""
from . import eqhip  # noqa: F401

__all__ = ["eqhip"]

try:  # C++ host mirror of the reference registry (pybind11); absent until built
    from ._engine import *  # noqa: F401,F403
    from ._engine import Engine, MarketSnapshot, PricingContext, ExecutionContext, MeasureResult  # noqa: F401
    __all__ += ["Engine", "MarketSnapshot", "PricingContext", "ExecutionContext", "MeasureResult"]
except ImportError as _e:  # pragma: no cover - reported by engine()
    _ENGINE_IMPORT_ERROR = _e
