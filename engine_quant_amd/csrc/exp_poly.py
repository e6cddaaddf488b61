"""Coefficients of bond_exp (eqhip_device.hpp), the FP64 e^x of the per-path zero-coupon bonds.

Method: Cody-Waite reduction x = n ln2 + r, |r| <= ln2/2, and e^r = 1 + r(1 + r(1/2 + r q(r))) with q a
degree-9 Chebyshev fit (mpmath.chebyfit, near-minimax) of (e^r - 1 - r - r^2/2) / r^3 on
|r| <= 1.0005 ln2/2 (the margin covers the rounding of x log2(e) before rint). The script prints the
coefficients as C hex literals, highest degree first, and the largest error in ulps of the FP64
evaluation (one rounding per fused multiply-add, emulated in 256-bit arithmetic) over a sample of x.

    python engine_quant_amd/csrc/exp_poly.py
"""
import math
import random
import struct

import mpmath as mp

mp.mp.prec = 256
DEG = 9
HALF = mp.log(2) / 2 * mp.mpf("1.0005")
LN2_HI = struct.unpack("<d", struct.pack("<Q", 0x3FE62E42FEE00000))[0]  # 32 significant bits
LN2_LO = float(mp.log(2) - mp.mpf(LN2_HI))
LOG2E = float(1 / mp.log(2))


def q_exact(r):
    r = mp.mpf(r)
    if abs(r) < mp.mpf("1e-30"):
        return mp.mpf(1) / 6
    return (mp.exp(r) - 1 - r - r * r / 2) / r ** 3


def fma(x, y, z):
    return float(mp.mpf(x) * mp.mpf(y) + mp.mpf(z))  # one rounding


def evaluate(x, coeffs):
    n = round(x * LOG2E)
    r = fma(-n, LN2_HI, x)
    r = fma(-n, LN2_LO, r)
    q = coeffs[0]
    for c in coeffs[1:]:
        q = fma(q, r, c)
    p = fma(q, r, 0.5)
    p = fma(p, r, 1.0)
    p = fma(p, r, 1.0)
    return math.ldexp(p, n)


def main():
    fit, err = mp.chebyfit(q_exact, [-HALF, HALF], DEG + 1, error=True)
    coeffs = [float(c) for c in fit]
    print(f"fit error of q: {mp.nstr(err, 4)} (of e^r: {mp.nstr(err * HALF ** 3 / mp.exp(-HALF) * 2 ** 53, 3)} ulp)")
    print(f"log2(e) {LOG2E.hex()}  ln2 head {LN2_HI.hex()}  ln2 tail {LN2_LO.hex()}")
    for i, c in enumerate(coeffs):
        print(f"q{DEG - i}: {c.hex()}")
    rng = random.Random(1)
    xs = [rng.uniform(-40.0, 2.0) for _ in range(20000)] + [rng.uniform(-708.0, 708.0) for _ in range(5000)]
    xs += [rng.uniform(-1e-3, 1e-3) for _ in range(2000)]
    worst = 0.0
    for x in xs:
        y, t = evaluate(x, coeffs), mp.exp(mp.mpf(x))
        worst = max(worst, float(abs(mp.mpf(y) - t)) / math.ulp(float(t)))
    print(f"max error over {len(xs)} points: {worst:.3f} ulp")


if __name__ == "__main__":
    main()
