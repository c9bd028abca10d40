"""Pins the FP64 oracle (oracle/kde_oracle.c) before it is trusted.

Checks, in order of strength:
  * known answers printed by the reference's own tests
    (test/test_point2d.cc:43-58, test/test_kernels.cc:49-172);
  * golden fixtures written by the reference itself (tests/golden/);
  * the reference build (oracle/_ref) on fresh inputs, where it exists;
  * closed forms and invariants the mathematics fixes, chosen so that a
    dropped term, a wrong sign or a swapped axis fails one of them.
"""
from __future__ import annotations

from math import gamma, pi, sqrt
from pathlib import Path

import numpy as np
import pytest

GOLDEN = Path(__file__).resolve().parent / "golden"

EPAN, GAUSS, EPANPROD, GAUSSPROD, EPANCONV1, EPANPRODCONV, GAUSSCONV1, GAUSSPRODCONV = range(1, 9)


# ---- known answers of the reference's tests ---------------------------------
def test_point2d_known_answers(oracle):
    # test/test_point2d.cc:43-58: C * U between (1,1) and the origin / (3,2)
    k = lambda kind, h, dx: oracle.normalization(kind, 2, h) * oracle.kernel_u(kind, 2, h, h, dx)
    assert k(GAUSS, 1.0, [1, 1]) == pytest.approx(0.05854983152431917, rel=1e-14)
    assert k(GAUSS, 2.0, [1, 1]) == pytest.approx(0.03098749857741324, rel=1e-14)
    assert k(EPAN, 1.0, [-2, -1]) == 0.0
    assert k(EPAN, 5.0, [-2, -1]) == pytest.approx(0.020371832715762605, rel=1e-14)


def test_kernels_normalisations(oracle):
    # test/test_kernels.cc:49-172 prints these "(c.f. ...)" values
    assert oracle.normalization(EPAN, 1, 1.0) == pytest.approx(0.75)
    assert oracle.normalization(EPAN, 1, 0.5) == pytest.approx(1.5)
    assert oracle.normalization(EPAN, 1, 2.0) == pytest.approx(0.375)
    assert oracle.normalization(EPAN, 2, 1.0) == pytest.approx(0.5 * (1 / pi) * 4)
    for h in (1.0, 0.5, 2.0):
        assert oracle.normalization(GAUSS, 1, h) == pytest.approx((2 * pi) ** -0.5 / h)
    assert oracle.normalization(GAUSS, 2, 1.0) == pytest.approx(1 / (2 * pi))
    for hx, hy in ((1.2, 1.0), (1.5, 1.0), (1.2, 0.7)):
        assert oracle.normalization(GAUSSPROD, 2, hx, hy) == pytest.approx(1 / (2 * pi) / (hx * hy))
        assert oracle.normalization(EPANPROD, 2, hx, hy) == pytest.approx(0.75 ** 2 / (hx * hy))


def test_normalisations_integrate_to_one(oracle):
    # C * integral of U = 1 for every density kernel; the convolution kernels
    # integrate to 1 as well (they are densities of x - y).
    g = np.linspace(-6, 6, 2401)
    dx = g[1] - g[0]
    X, Y = np.meshgrid(g, g, indexing="ij")
    for kind, dim, hx, hy in ((EPAN, 1, 0.7, 0.7), (GAUSS, 1, 0.7, 0.7), (EPANCONV1, 1, 0.7, 0.7),
                              (GAUSSCONV1, 1, 0.7, 0.7)):
        u = np.array([oracle.kernel_u(kind, 1, hx, hy, [x]) for x in g])
        assert oracle.normalization(kind, dim, hx, hy) * u.sum() * dx == pytest.approx(1.0, abs=2e-4)
    sub = slice(None, None, 4)
    xs, ys = X[sub, sub].ravel(), Y[sub, sub].ravel()
    d2 = dx * 4
    for kind, hx, hy in ((EPAN, 0.9, 0.9), (GAUSS, 0.8, 0.8), (EPANPROD, 0.9, 1.3), (GAUSSPROD, 0.7, 1.1),
                         (EPANPRODCONV, 0.9, 1.3)):
        u = np.array([oracle.kernel_u(kind, 2, hx, hy, [a, b]) for a, b in zip(xs, ys)])
        assert oracle.normalization(kind, 2, hx, hy) * u.sum() * d2 * d2 == pytest.approx(1.0, abs=5e-3)


def test_convolution_kernels_are_self_convolutions(oracle):
    # K*(d) = integral K(t) K(d - t) dt for the Epanechnikov 1-D pair and the
    # Gaussian 1-D pair (ConvKernelAssociator.h:23-71).
    t = np.linspace(-4, 4, 8001)
    dt = t[1] - t[0]
    for k, kc, h in ((EPAN, EPANCONV1, 0.6), (GAUSS, GAUSSCONV1, 0.5)):
        c, cc = oracle.normalization(k, 1, h), oracle.normalization(kc, 1, h)
        K = lambda x: c * np.array([oracle.kernel_u(k, 1, h, h, [v]) for v in np.atleast_1d(x)])
        for d in (0.0, 0.3, 0.9):
            conv = (K(t) * K(d - t)).sum() * dt
            assert cc * oracle.kernel_u(kc, 1, h, h, [d]) == pytest.approx(conv, rel=2e-3, abs=1e-6)


def test_gaussian_product_conv_is_reference_bug_compatible(oracle):
    # The float/double GaussianProductConv2dTraits evaluate exp(-0.5 arg)
    # (KernelTraits.h:432, 445): U* equals U of GaussianProductKernel2d and
    # C* = C/2.  The true convolution would be exp(-0.25 arg).
    hx, hy = 0.4, 0.7
    for dx in ([0.0, 0.0], [0.3, -0.2], [1.1, 0.5]):
        assert oracle.kernel_u(GAUSSPRODCONV, 2, hx, hy, dx) == pytest.approx(
            oracle.kernel_u(GAUSSPROD, 2, hx, hy, dx), rel=1e-15)
    assert oracle.normalization(GAUSSPRODCONV, 2, hx, hy) == pytest.approx(
        0.5 * oracle.normalization(GAUSSPROD, 2, hx, hy), rel=1e-15)


# ---- golden fixtures from the reference ----------------------------------------
def _direct_fixture():
    return np.load(GOLDEN / "direct_ref.npz")


@pytest.mark.parametrize("i", range(10))
def test_direct_matches_reference_fixture(oracle, i):
    f = _direct_fixture()
    kind, dim, hx, hy = f["cases"][i]
    kind, dim = int(kind), int(dim)
    pts, w, q = f["pts"][:, :dim], f["w"], f["q"][:, :dim]
    want = f[f"direct_{i}_kf0"]  # KernelDensity<D, Kernel<double>, double>::direct_eval
    got = oracle.direct(kind, hx, hy, pts, q, mass=w / w.sum(), normalize=True)
    if (kind, dim) == (GAUSS, 1):
        # GaussianTraits<1,double>::normalization(float h) narrows h to float
        # (KernelTraits.h:143); the rest of the value is FP64.
        got = got * (hx / float(np.float32(hx)))
    np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-300)


@pytest.mark.parametrize("i", range(10))
def test_direct_float_kernels_within_fp32(oracle, i):
    # Kernel<float> evaluates each term in FP32 (the reference GPU callers'
    # type): it agrees with the FP64 oracle to FP32 accuracy.
    f = _direct_fixture()
    kind, dim, hx, hy = f["cases"][i]
    kind, dim = int(kind), int(dim)
    pts, w, q = f["pts"][:, :dim], f["w"], f["q"][:, :dim]
    want = f[f"direct_{i}_kf1"]
    got = oracle.direct(kind, float(np.float32(hx)), float(np.float32(hy)), pts, q, mass=w / w.sum(),
                        normalize=True)
    big = np.abs(got) > 1e-6 * np.abs(got).max()
    np.testing.assert_allclose(got[big], want[big], rtol=2e-5)


@pytest.mark.parametrize("i", range(4))
def test_cv_scores_match_reference_fixture(oracle, i):
    f = np.load(GOLDEN / "cv_ref.npz")
    kind, dim, hx, hy = f["cases"][i]
    kind, dim = int(kind), int(dim)
    pts, w = f["pts"][:, :dim], f["w"]
    wn = w / w.sum()
    lcv = oracle.likelihood_cv(kind, hx, hy, pts, wn)
    lscv = oracle.lsq_conv_cv(kind, hx, hy, pts, wn)
    if (kind, dim) == (GAUSS, 1):
        tol = 1e-6  # float-narrowed normalisation of the 1-D double Gaussian
    else:
        tol = 1e-9
    assert lcv == pytest.approx(float(f[f"lcv_{i}"]), rel=tol, abs=tol)
    assert lscv == pytest.approx(float(f[f"lscv_{i}"]), rel=tol, abs=tol)


def test_adapt_matches_reference_fixture(oracle):
    f = np.load(GOLDEN / "adapt_ref.npz")
    for alpha in (0.0, 0.5, 1.0):
        tag = str(alpha).replace(".", "p")
        c, w = f[f"coords_{tag}"], f[f"w_{tag}"]
        g, lam = oracle.adapt(GAUSSPROD, 0.3, 0.45, c, w, alpha)
        np.testing.assert_allclose(lam, f[f"abw_{tag}"], rtol=1e-9)
        np.testing.assert_allclose(w * lam ** -2.0, f[f"mass_{tag}"], rtol=1e-9)


def test_tree_bounds_bracket_oracle(oracle):
    # The reference's dual tree returns bounds that bracket the exact
    # density and meet its error guarantee; the oracle's exact value must
    # lie inside them.
    f = np.load(GOLDEN / "tree_ref.npz")
    pts, w = f["pts"], f["w"]
    for kind, hx, hy, tag in ((GAUSS, 0.3, 0.3, "gauss"), (EPANPROD, 0.4, 0.55, "epanprod")):
        q = f[f"q_{tag}"]
        exact = oracle.direct(kind, hx, hy, pts, q, mass=w / w.sum(), normalize=True)
        lo, up = f[f"lower_{tag}"], f[f"upper_{tag}"]
        slack = 1e-12 * np.maximum(np.abs(exact), 1.0)
        assert np.all(lo <= exact + slack) and np.all(exact <= up + slack)
        mid = 0.5 * (lo + up)
        assert np.all(np.abs(mid - exact) <= np.maximum(1e-6, 1e-3 * np.abs(exact)) + slack)


def test_numint_lscv_matches_riemann_sum(oracle):
    # lsq_numint_cross_validate = sum_grid f^2 dx dy - 2 * LOO term
    # (KernelDensityImpl.h:14-97), grid start + i*delta, end exclusive.
    f = np.load(GOLDEN / "numint_ref.npz")
    pts, w = f["pts"], f["w"]
    wn = w / w.sum()
    hx, hy = 0.3, 0.45
    g = -4.0 + np.arange(100) * 0.08
    X, Y = np.meshgrid(g, g, indexing="xy")
    grid = np.column_stack([X.ravel(), Y.ravel()])
    fg = oracle.direct(GAUSSPROD, hx, hy, pts, grid, mass=wn, normalize=True)
    loo = oracle.direct(GAUSSPROD, hx, hy, pts, pts, mass=wn, exclude_self=True, normalize=True)
    val = (fg ** 2).sum() * 0.08 * 0.08 - 2.0 * (wn * loo).sum()
    assert val == pytest.approx(float(f["numint_gaussprod"]), rel=1e-9)


# ---- live comparison against the reference build --------------------------------
def test_live_reference_direct_random(oracle, reference):
    rng = np.random.default_rng(7)
    pts = rng.normal(size=(300, 2))
    w = rng.uniform(0.2, 2.0, 300)
    q = rng.uniform(-3, 3, (150, 2))
    for kind, hx, hy in ((EPAN, 0.5, 0.5), (GAUSS, 0.4, 0.4), (EPANPROD, 0.6, 0.3), (GAUSSPROD, 0.2, 0.7),
                         (EPANPRODCONV, 0.6, 0.3), (GAUSSPRODCONV, 0.2, 0.7)):
        want = reference.direct_eval(kind, 0, hx, hy, pts, w, q)
        got = oracle.direct(kind, hx, hy, pts, q, mass=w / w.sum(), normalize=True)
        np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-300)


# ---- closed forms and invariants -------------------------------------------------
def test_single_and_two_point_closed_forms(oracle):
    h = 0.3
    x = np.array([[0.2, -0.1]])
    q = np.array([[0.5, 0.3], [0.2, -0.1]])
    got = oracle.direct(GAUSS, h, h, x, q, mass=[1.0], normalize=True)
    d2 = ((q - x) ** 2).sum(1)
    np.testing.assert_allclose(got, np.exp(-0.5 * d2 / h ** 2) / (2 * pi * h ** 2), rtol=1e-14)
    # two points with masses 0.25/0.75 and abw 2 on the second point
    x2 = np.array([[0.0, 0.0], [1.0, 0.5]])
    got = oracle.direct(GAUSSPROD, 0.4, 0.6, x2, q, mass=[0.25, 0.75], abw=[1.0, 2.0], normalize=True)
    u = lambda dx, dy, a: np.exp(-0.5 * ((dx / (0.4 * a)) ** 2 + (dy / (0.6 * a)) ** 2))
    want = (0.25 * u(q[:, 0], q[:, 1], 1) + 0.75 * u(q[:, 0] - 1, q[:, 1] - 0.5, 2)) / (2 * pi * 0.4 * 0.6)
    np.testing.assert_allclose(got, want, rtol=1e-14)


def test_axes_are_not_swapped(oracle):
    # product kernels: hx acts on x only
    u = oracle.kernel_u(EPANPROD, 2, 1.0, 0.1, [0.5, 0.0])
    assert u == pytest.approx(0.75)
    assert oracle.kernel_u(EPANPROD, 2, 1.0, 0.1, [0.0, 0.5]) == 0.0
    assert oracle.kernel_u(GAUSSPROD, 2, 1.0, 0.5, [0.0, 1.0]) == pytest.approx(np.exp(-2.0))


def test_epanechnikov_conv_kernel_values(oracle):
    # k(u) = (2-|u|)^3 (u^2 + 6|u| + 4)/32: k(0) = 1, k(1) = 11/32, k(2) = 0
    for u, want in ((0.0, 1.0), (1.0, 11 / 32), (-1.0, 11 / 32), (2.0, 0.0), (3.0, 0.0)):
        assert oracle.kernel_u(EPANCONV1, 1, 1.0, 1.0, [u]) == pytest.approx(want, abs=1e-15)


def test_scaling_identity(oracle):
    # f(q; h, X) = s^-D f(q/s; h/s, X/s)
    rng = np.random.default_rng(3)
    x = rng.normal(size=(200, 2))
    q = rng.normal(size=(50, 2))
    s = 2.5
    for kind in (EPAN, GAUSS):
        a = oracle.direct(kind, 0.4, 0.4, x, q, mass=np.full(200, 1 / 200), normalize=True)
        b = oracle.direct(kind, 0.4 / s, 0.4 / s, x / s, q / s, mass=np.full(200, 1 / 200), normalize=True)
        np.testing.assert_allclose(a, b / s ** 2, rtol=1e-12)


def test_linearity_and_permutation(oracle):
    rng = np.random.default_rng(4)
    x = rng.normal(size=(300, 2))
    q = rng.normal(size=(40, 2))
    m1, m2 = rng.random(300), rng.random(300)
    for kind in (GAUSSPROD, EPANPRODCONV):
        a = oracle.direct(kind, 0.3, 0.5, x, q, mass=m1 + m2)
        b = oracle.direct(kind, 0.3, 0.5, x, q, mass=m1) + oracle.direct(kind, 0.3, 0.5, x, q, mass=m2)
        np.testing.assert_allclose(a, b, rtol=1e-13)
        p = rng.permutation(300)
        c = oracle.direct(kind, 0.3, 0.5, x[p], q, mass=m1[p] + m2[p])
        np.testing.assert_allclose(a, c, rtol=1e-13)


def test_loo_exclusion_equals_subtraction_when_well_conditioned(oracle):
    # sum over j != i equals the all-pairs sum minus m_i U(0) = m_i
    rng = np.random.default_rng(5)
    x = rng.normal(size=(250, 2)) * 0.3
    m = rng.uniform(0.5, 1.5, 250)
    full = oracle.direct(GAUSS, 0.5, 0.5, x, x, mass=m)
    loo = oracle.direct(GAUSS, 0.5, 0.5, x, x, mass=m, exclude_self=True)
    np.testing.assert_allclose(loo, full - m, rtol=1e-11)


def test_gaussian_lscv_closed_form_and_reference_relation(oracle):
    # Mathematically, int fhat^2 = sum_ij w_i w_j N(x_i - x_j; 0, 2 H) for a
    # Gaussian product kernel.  The reference's convolution kernel is
    # exp(-0.5 arg) with C/2, so its "square term" is instead
    # 0.5 * sum_i w_i fhat(x_i).  Both relations are checked.
    rng = np.random.default_rng(6)
    x = rng.normal(size=(200, 2))
    w = rng.uniform(0.5, 1.5, 200)
    w /= w.sum()
    hx, hy = 0.3, 0.5
    d = x[:, None, :] - x[None, :, :]
    exact_sq = (w[:, None] * w[None, :] * np.exp(-0.25 * ((d[..., 0] / hx) ** 2 + (d[..., 1] / hy) ** 2))
                / (4 * pi * hx * hy)).sum()
    f_all = oracle.direct(GAUSSPROD, hx, hy, x, x, mass=w, normalize=True)
    f_loo = oracle.direct(GAUSSPROD, hx, hy, x, x, mass=w, exclude_self=True, normalize=True)
    lscv = oracle.lsq_conv_cv(GAUSSPROD, hx, hy, x, w)
    assert lscv == pytest.approx(0.5 * (w * f_all).sum() - 2 * (w * f_loo).sum(), rel=1e-12)
    # the mathematically correct square term differs from the reference's
    assert abs(exact_sq - 0.5 * (w * f_all).sum()) > 1e-3 * exact_sq
    # with a 1-D Gaussian the convolution kernel is exp(-0.25 arg): correct
    x1 = x[:, :1]
    d1 = x1[:, None, 0] - x1[None, :, 0]
    exact1 = (w[:, None] * w[None, :] * np.exp(-0.25 * (d1 / hx) ** 2) / (sqrt(4 * pi) * hx)).sum()
    f1_loo = oracle.direct(GAUSS, hx, hx, x1, x1, mass=w, exclude_self=True, normalize=True)
    assert oracle.lsq_conv_cv(GAUSS, hx, hx, x1, w) == pytest.approx(exact1 - 2 * (w * f1_loo).sum(), rel=1e-12)


def test_adapt_invariants(oracle):
    rng = np.random.default_rng(8)
    x = rng.normal(size=(300, 2))
    w = rng.uniform(0.5, 1.5, 300)
    w /= w.sum()
    g0, lam0 = oracle.adapt(GAUSSPROD, 0.3, 0.3, x, w, 0.0)
    np.testing.assert_allclose(lam0, 1.0)
    g, lam = oracle.adapt(GAUSSPROD, 0.3, 0.3, x, w, 0.5)
    # geometric-mean normaliser: sum_i w_i log(lambda_i) = 0
    assert abs((w * np.log(lam)).sum()) < 1e-12
    # lambda = (f/g)^-alpha with f the full pilot density (self included)
    f = oracle.direct(GAUSSPROD, 0.3, 0.3, x, x, mass=w, normalize=True)
    np.testing.assert_allclose(lam, (f / g) ** -0.5, rtol=1e-12)
    # two symmetric points: equal densities, so lambda = 1 for both
    _, lam2 = oracle.adapt(GAUSS, 0.5, 0.5, np.array([[0.0, 0.0], [1.0, 0.0]]), np.array([0.5, 0.5]), 0.7)
    np.testing.assert_allclose(lam2, 1.0, rtol=1e-14)


def test_epanechnikov_general_d_normalisation(oracle):
    for D in (3, 5, 16):
        want = 0.5 * (D + 2) / (pi ** (D / 2) / gamma(1 + D / 2))
        assert oracle.normalization(EPAN, D, 1.0) == pytest.approx(want, rel=1e-14)
        assert oracle.normalization(GAUSS, D, 0.5) == pytest.approx((2 * pi) ** (-D / 2) / 0.5 ** D, rel=1e-14)


def test_loo_at_subset_equals_full_leave_one_out(oracle):
    # bko_direct_loo_at(idx) must return exactly the idx entries of the
    # whole-set leave-one-out sums (same loop and order), for every kind used
    # by the full-size GPU checks, with adaptive bandwidths
    rng = np.random.default_rng(31)
    pts = rng.normal(size=(700, 2))
    mass = rng.uniform(0.5, 1.5, 700)
    abw = rng.uniform(0.5, 2.0, 700)
    idx = np.array([0, 5, 699, 350, 5, 123])
    for kind, (hx, hy) in ((2, (0.2, 0.2)), (4, (0.15, 0.25)), (3, (0.3, 0.4))):
        full = oracle.direct(kind, hx, hy, pts, pts, mass=mass, abw=abw, exclude_self=True)
        sub = oracle.direct_loo_at(kind, hx, hy, pts, idx, mass=mass, abw=abw)
        np.testing.assert_array_equal(sub, full[idx])
    with pytest.raises(RuntimeError):
        oracle.direct_loo_at(2, 0.2, 0.2, pts, [700])


def test_direct_high_dim_matches_reference_fixture(oracle):
    # generic GaussianKernel<D> / EpanechnikovKernel<D>, D = 3, 8, 16, written by
    # the reference's own KernelDensity::direct_eval (weights normalised to sum 1)
    f = np.load(GOLDEN / "direct_hd_ref.npz")
    for i, (kind, dim, h) in enumerate(f["cases"]):
        p, w, q = f[f"pts_{i}"], f[f"w_{i}"], f[f"q_{i}"]
        assert p.shape[1] == int(dim)
        want = f[f"direct_{i}"]
        got = oracle.direct(int(kind), h, h, p, q, mass=w / w.sum(), normalize=True)
        assert (want > 0).mean() > 0.9
        np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-14 * np.abs(want).max())
