"""The FP64 exp of the per-path zero-coupon bonds (bond_exp and bond_exp_v in eqhip_device.hpp), run on
the device through eqhip_bond_exp and checked against e^x in 256-bit arithmetic: below one ulp over the
range the valuation kernels use and far beyond it, exact overflow to inf and underflow to 0, NaN
propagated. The entry point returns NaN wherever the two forms differ in any bit, so every finite check
below also checks that they agree."""
import math

import numpy as np
import pytest

from engine_quant_amd import eqhip

pytestmark = pytest.mark.gpu

mp = pytest.importorskip("mpmath")
mp.mp.prec = 256


def _ulps(y, x):
    e = mp.exp(mp.mpf(float(x)))
    u = math.ulp(float(e))
    return float(abs(mp.mpf(float(y)) - e) / u)


@pytest.fixture(scope="module")
def ctx():
    if not eqhip.is_gpu_backend_available():
        pytest.skip("needs a gfx950 device")
    c = eqhip.Context(0)
    yield c
    c.close()


def test_bond_exp_within_one_ulp(ctx):
    rng = np.random.default_rng(11)
    ln2 = math.log(2.0)
    x = np.concatenate([
        rng.uniform(-40.0, 2.0, 20000),        # -B r of the bonds: B up to 30 (a small), |r| up to ~1
        rng.uniform(-708.0, 708.0, 5000),
        rng.uniform(-1e-3, 1e-3, 2000),
        (np.arange(-1000, 1000) + 0.5) * ln2,   # reduction ties: |r| at ln2 / 2
        np.array([0.0, -0.0, 1e-300, -1e-300, 708.0, -708.0, 1.0, -1.0, ln2, -ln2]),
    ])
    y = ctx.bond_exp(x)
    worst = max(_ulps(a, b) for a, b in zip(y, x))
    print(f"bond_exp: worst error {worst:.3f} ulp over {len(x)} values", flush=True)
    assert worst < 1.0


def test_bond_exp_edges(ctx):
    x = np.array([709.78, 709.8, 710.0, 1e5, 1e9, 1e10, 1e300, np.inf,
                  -745.2, -746.0, -1e5, -1e9, -1e10, -1e300, -np.inf, np.nan, -740.0, -730.0])
    y = ctx.bond_exp(x)
    assert math.isfinite(y[0]) and _ulps(y[0], x[0]) < 1.0  # the largest finite results
    assert np.all(np.isinf(y[1:8])) and np.all(y[1:8] > 0)  # overflow
    assert np.all(y[8:15] == 0.0)                            # underflow
    assert np.isnan(y[15])
    # subnormal results: ldexp rounds the polynomial's value once more, within one subnormal spacing
    for a, b in zip(y[16:], x[16:]):
        assert a > 0.0 and abs(a - math.exp(b)) <= 2.0 ** -1074


def test_bond_exp_refuses_bad_arguments(ctx):
    with pytest.raises(eqhip.EqhipError, match="n values"):
        eqhip._check(eqhip.lib().eqhip_bond_exp(ctx.handle, None, None, 3))
