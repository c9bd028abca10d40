"""The FP64 exponential of the FP64-context Gaussian kernel (exp2_f64 in
bbrcit_kde_amd/csrc/bk_kernels.hip): its degree-10 polynomial for 2^f on
[-1/2, 1/2], read from the kernel source and evaluated as the kernel does
(Horner form, each FMA rounded once to FP64, emulated in 200-bit arithmetic),
against 2^f itself.  A wrong or missing coefficient moves the error far above
the bound.  CPU only; skipped where mpmath is absent.
"""
from __future__ import annotations

import re
from pathlib import Path

import numpy as np
import pytest

mp = pytest.importorskip("mpmath")

SRC = Path(__file__).resolve().parents[1] / "bbrcit_kde_amd" / "csrc" / "bk_kernels.hip"


def _coefficients():
    text = SRC.read_text()
    body = text[text.index("__device__ __forceinline__ double exp2_f64(double e)"):]
    body = body[:body.index("return __builtin_ldexp")]
    first = re.search(r"double p = ([0-9.e+-]+);", body).group(1)
    rest = re.findall(r"__builtin_fma\(p, f, ([0-9.e+-]+)\)", body)
    return [float(first)] + [float(c) for c in rest]


def _horner_fma(coef, f):
    p = coef[0]
    for c in coef[1:]:
        p = float(mp.mpf(p) * mp.mpf(f) + mp.mpf(c))  # one rounding, as an FMA
    return p


def test_exp2_polynomial_is_accurate_to_a_few_ulp():
    mp.mp.prec = 200
    coef = _coefficients()
    assert len(coef) == 11 and coef[-1] == 1.0
    fs = np.concatenate([np.linspace(-0.5, 0.5, 2001), np.random.default_rng(5).uniform(-0.5, 0.5, 2000),
                         [-0.5, 0.5, 0.0, 2.0 ** -30, -(2.0 ** -30)]])
    worst = 0.0
    for f in fs:
        exact = mp.power(2, mp.mpf(float(f)))
        worst = max(worst, float(abs((mp.mpf(_horner_fma(coef, float(f))) - exact) / exact)))
    assert worst < 5e-16, worst


def test_exp2_range_reduction_and_scale():
    # e = n + f with n = rint(e): 2^e = ldexp(p(f), n), the polynomial's error
    # carried unchanged to every normal result
    mp.mp.prec = 200
    coef = _coefficients()
    for e in (-0.3, -1.5, -7.25, -100.49, -1021.7):
        n = float(np.rint(e))
        got = np.ldexp(_horner_fma(coef, e - n), int(n))
        exact = mp.power(2, mp.mpf(e))
        assert abs((mp.mpf(got) - exact) / exact) < 5e-16, e
