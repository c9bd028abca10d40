"""Coefficients of exp2_f64's polynomial (bk_kernels.hip).

2^f on [-1/2, 1/2] by its Chebyshev interpolant of degree 10 (mpmath's
chebyfit at 200 bits), each coefficient rounded once to FP64; prints them
highest degree first and the largest relative error of their FMA Horner
evaluation (each step rounded once, as the kernel's fma) on 4,001 points.
Needs mpmath; a development script, not imported by the library.
"""
import mpmath as mp
import numpy as np

mp.mp.prec = 200
DEG = 10
poly, err = mp.chebyfit(lambda x: mp.power(2, x), [-0.5, 0.5], DEG + 1, error=True)
coef = [float(p) for p in poly]
worst = 0.0
for f in np.linspace(-0.5, 0.5, 4001):
    p = coef[0]
    for c in coef[1:]:
        p = float(mp.mpf(p) * mp.mpf(float(f)) + mp.mpf(c))
    exact = mp.power(2, mp.mpf(float(f)))
    worst = max(worst, float(abs((mp.mpf(p) - exact) / exact)))
print("coefficients (degree", DEG, "down to 0):")
for c in coef:
    print(f"  {c!r}")
print(f"interpolation error {float(err):.3e}, with FMA rounding {worst:.3e}")
