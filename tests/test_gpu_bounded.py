"""GPU parity of the bounded (dual-tree) evaluation, bk_eval_bounded.

The reference's KernelDensity::eval(query_tree, kernel, rel_err, abs_err)
(include/KernelDensityImpl.h:681-737) returns, per query, bounds with
    lower <= f <= upper,  upper - lower <= max(abs_err, rel_err * lower)
(its report_error check, 1114-1143 of the same file).  These tests check that
contract against the FP64 oracle's exact sums: the bounds must bracket the
exact value up to the arithmetic's own error (FP32 contexts: 1e-5 relative
above a floor of 1e-10 times an upper bound of the sum, the floor itself
below it; FP64 contexts: 1e-12), and the gap must meet the tolerance.  The
cases use tolerances loose enough that tiles are pruned, so both the pruned
and the summed parts are exercised; the pruning counts are checked too.
"""
from __future__ import annotations

import numpy as np
import pytest

from bbrcit_kde_amd import datasets as ds
from tests.test_gpu_direct import f32

pytestmark = pytest.mark.gpu

EPAN, GAUSS, EPANPROD, GAUSSPROD, EPANCONV1, EPANPRODCONV, GAUSSCONV1, GAUSSPRODCONV = range(1, 9)
CASES_2D = [(GAUSS, (0.05, 0.05)), (GAUSSPROD, (0.04, 0.07)), (GAUSSPRODCONV, (0.04, 0.07)),
            (EPAN, (0.15, 0.15)), (EPANPROD, (0.12, 0.2)), (EPANPRODCONV, (0.12, 0.2))]
CASES_1D = [(GAUSS, (0.05, 0.05)), (EPAN, (0.1, 0.1)), (EPANCONV1, (0.1, 0.1)), (GAUSSCONV1, (0.05, 0.05))]


def kd_points(n, seed, dim):
    pts, w = ds.gaussian_mixture(n, seed=seed)
    pts = pts[:, :dim]
    perm = ds.kd_order(pts, 64)
    return pts[perm], (w / w.sum())[perm]


def check_bounds(lo, up, exact, bound, rel_err, abs_err, f64, what):
    arith = 1e-12 if f64 else 1e-5
    floor = 1e-10 * bound
    slack = arith * np.maximum(np.abs(exact), floor)
    assert np.all(lo <= exact + slack), f"{what}: lower above the exact value"
    assert np.all(exact <= up + slack), f"{what}: upper below the exact value"
    gap = up - lo
    allowed = np.maximum(abs_err, rel_err * lo)
    # the exact part of lower carries the arithmetic's rounding
    assert np.all(gap <= allowed * (1 + arith) + 1e-300), f"{what}: gap {np.max(gap - allowed):.3e} over"
    assert np.all(gap >= 0)


@pytest.mark.parametrize("f64", [False, True])
@pytest.mark.parametrize("adaptive", [False, True])
@pytest.mark.parametrize("kind,h", CASES_2D)
def test_bounded_2d_meets_contract(bk, oracle, kind, h, adaptive, f64):
    n_ref, n_q = 20_000, 6_000
    pts, mass = kd_points(n_ref, 700 + kind, 2)
    q = ds.uniform_box(n_q, -4.0, 4.0, 800 + kind)
    q = q[ds.kd_order(q, 64)]
    abw = np.random.default_rng(3).uniform(0.7, 1.5, n_ref) if adaptive else None
    rel_err, abs_err = 1e-2, 1e-6
    desc = bk.KernelDesc(kind, 2, *h)
    prec = bk.BK_PREC_F64 if f64 else bk.BK_PREC_F32
    with bk.DirectKde(pts, q, mass, abw, precision=prec) as kde:
        lo, up, st = kde.eval_bounded(desc, 0, n_q - 1, rel_err, abs_err)
    P, Q, M, A = (pts, q, mass, abw) if f64 else (f32(pts), f32(q), f32(mass), f32(abw))
    exact = oracle.direct(kind, h[0], h[1], P, Q, mass=M, abw=A, normalize=True)
    bound = oracle.normalization(kind, 2, *h) * mass.sum()
    check_bounds(lo, up, exact, bound, rel_err, abs_err, f64, f"kind {kind} adaptive {adaptive} f64 {f64}")
    # (reference tile, query block) pairs: 256 queries per block on the FP32
    # Gaussian path, 1024 on the Epanechnikov and FP64 paths
    gauss32 = not f64 and kind in (GAUSS, GAUSSPROD, GAUSSPRODCONV)
    total = -(-n_ref // 256) * -(-n_q // (256 if gauss32 else 1024))
    assert st["excluded"] + st["approximated"] + st["evaluated"] == total
    assert st["excluded"] + st["approximated"] > 0 and st["evaluated"] > 0


@pytest.mark.parametrize("f64", [False, True])
@pytest.mark.parametrize("kind,h", CASES_1D)
def test_bounded_1d_meets_contract(bk, oracle, kind, h, f64):
    pts, mass = kd_points(12_000, 900 + kind, 1)
    q = np.sort(ds.uniform_box(5_000, -4.0, 4.0, 901 + kind, dim=1), axis=0)
    desc = bk.KernelDesc(kind, 1, *h)
    prec = bk.BK_PREC_F64 if f64 else bk.BK_PREC_F32
    with bk.DirectKde(pts, q, mass, precision=prec) as kde:
        lo, up, st = kde.eval_bounded(desc, 0, len(q) - 1, 1e-3, 1e-7)
    P, Q, M = (pts, q, mass) if f64 else (f32(pts), f32(q), f32(mass))
    exact = oracle.direct(kind, h[0], h[1], P, Q, mass=M, normalize=True)
    check_bounds(lo, up, exact, oracle.normalization(kind, 1, *h), 1e-3, 1e-7, f64, f"1-D kind {kind} f64 {f64}")
    assert st["evaluated"] > 0


@pytest.mark.parametrize("kind,h", [(EPANPROD, (0.1, 0.15)), (GAUSS, (0.03, 0.03))])
def test_zero_tolerance_is_exact(bk, oracle, kind, h):
    # rel = abs = 0 prunes only tiles whose bounds coincide (compact support:
    # both zero); every other tile is summed, so the bounds are the direct sum
    pts, mass = kd_points(15_000, 41, 2)
    q = ds.uniform_box(3_000, -3.0, 3.0, 42)
    desc = bk.KernelDesc(kind, 2, *h)
    with bk.DirectKde(pts, q, mass) as kde:
        lo, up, st = kde.eval_bounded(desc, 0, len(q) - 1, 0.0, 0.0)
        direct = kde.eval(desc, 0, len(pts) - 1, 0, len(q) - 1)
    np.testing.assert_array_equal(lo, up)
    exact = oracle.direct(kind, h[0], h[1], f32(pts), f32(q), mass=f32(mass), normalize=True)
    check_bounds(lo, up, exact, oracle.normalization(kind, 2, *h), 0.0, 0.0, False, "zero tolerance")
    if kind == EPANPROD:
        np.testing.assert_array_equal(lo, direct)
        assert st["approximated"] == 0


@pytest.mark.parametrize("f64", [False, True])
def test_spatial_order_and_sub_range(bk, oracle, f64):
    # random-order queries, sorted on the device; a sub-range of them
    pts, mass = kd_points(20_000, 51, 2)
    q = ds.uniform_box(7_000, -4.0, 4.0, 52)
    h = (0.04, 0.06)
    desc = bk.KernelDesc(GAUSSPROD, 2, *h)
    prec = bk.BK_PREC_F64 if f64 else bk.BK_PREC_F32
    with bk.DirectKde(pts, q, mass, precision=prec) as kde:
        kde.set_option(bk.BK_OPT_QUERY_ORDER, bk.BK_QUERY_ORDER_SPATIAL)
        lo, up, st = kde.eval_bounded(desc, 1000, 6500, 1e-3, 1e-6)
        lo2, up2, _ = kde.eval_bounded(desc, 1000, 6500, 1e-3, 1e-6)
    P, Q, M = (pts, q, mass) if f64 else (f32(pts), f32(q), f32(mass))
    exact = oracle.direct(GAUSSPROD, h[0], h[1], P, Q[1000:6501], mass=M, normalize=True)
    check_bounds(lo, up, exact, oracle.normalization(GAUSSPROD, 2, *h), 1e-3, 1e-6, f64, f"spatial f64 {f64}")
    np.testing.assert_array_equal(lo, lo2)
    np.testing.assert_array_equal(up, up2)
    assert st["approximated"] + st["excluded"] > 0


def test_high_dimension_is_exact(bk, oracle):
    pts = ds.standard_normal(3_000, 5, 61)
    q = ds.standard_normal(700, 5, 62)
    mass = np.full(len(pts), 1.0 / len(pts))
    desc = bk.KernelDesc(GAUSS, 5, 0.6)
    with bk.DirectKde(pts, q, mass) as kde:
        lo, up, st = kde.eval_bounded(desc, 0, len(q) - 1, 1e-3, 1e-9)
    np.testing.assert_array_equal(lo, up)
    exact = oracle.direct(GAUSS, 0.6, 0.6, f32(pts), f32(q), mass=f32(mass), normalize=True)
    np.testing.assert_allclose(lo, exact, rtol=1e-5)
    assert st["approximated"] == 0 and st["excluded"] == 0


def test_bounded_argument_errors(bk):
    pts, mass = kd_points(1_000, 71, 2)
    desc = bk.KernelDesc(GAUSS, 2, 0.1)
    with bk.DirectKde(pts, pts[:100], mass) as kde:
        with pytest.raises(ValueError):
            kde.eval_bounded(desc, 0, 99, -1.0, 0.0)
        with pytest.raises(IndexError):
            kde.eval_bounded(desc, 0, 100, 1e-3, 0.0)
        with pytest.raises(ValueError):
            kde.eval_bounded(desc, 50, 10, 1e-3, 0.0)


def test_bounded_full_size_test_kde11(bk, oracle):
    # The reference's dual-tree benchmark at N = 1,048,576 (test_kde11.cc:
    # standard-normal points evaluated at themselves, EpanechnikovKernel<2,
    # float> h = 0.1, rel 1e-6, abs 1e-10): on 3 x 512 sampled points the
    # bounds bracket the oracle's exact sum and meet the tolerance; the
    # pruning passes over most (block, tile) pairs as exactly zero.
    n = 1 << 20
    pts = ds.standard_normal(n, 2, 1011)
    pts = pts[ds.kd_order(pts, 64)]
    mass = np.full(n, 1.0 / n)
    h = float(np.float32(0.1))
    desc = bk.KernelDesc(EPAN, 2, h)
    with bk.DirectKde(pts, None, mass) as kde:
        lo, up, st = kde.eval_bounded(desc, 0, n - 1, 1e-6, 1e-10)
    assert st["excluded"] > 0.9 * (st["excluded"] + st["approximated"] + st["evaluated"])
    for i0 in (0, n // 2, n - 512):
        sl = slice(i0, i0 + 512)
        exact = oracle.direct(EPAN, h, h, f32(pts), f32(pts[sl]), mass=f32(mass), normalize=True)
        check_bounds(lo[sl], up[sl], exact, oracle.normalization(EPAN, 2, h, h), 1e-6, 1e-10, False,
                     f"test_kde11 sample at {i0}")
    gap = up - lo
    assert np.all(gap <= np.maximum(1e-10, 1e-6 * lo) * (1 + 1e-5))
