"""GPU parity of workgroup-level tile culling and of the spatial query order.

A workgroup skips every reference tile whose bounding box lies beyond the
kernel's reach from the bounding box of its queries.  For the Epanechnikov
kinds the skipped terms are exactly zero, so results must be bit-identical
with and without the spatial order.  For the Gaussian kinds the skipped
terms of the factored tiles are zero too; those of the difference and Gram
forms are below 2^-100 of the largest mass, so both orders must meet the
oracle at the parity bar (relative 1e-5 above the floor of 1e-10 times an
upper bound of the sum, absolute below it).  FP64 contexts cull only tiles
whose terms are exactly zero and are held to 1e-12 (Gaussian kinds relative
down to 1e-290).  The cases use kd-ordered references and small bandwidths, where
most tiles are out of reach of a workgroup of sorted queries.
"""
from __future__ import annotations

import numpy as np
import pytest

from bbrcit_kde_amd import datasets as ds
from tests.test_gpu_direct import assert_parity, f32

pytestmark = pytest.mark.gpu

EPAN, GAUSS, EPANPROD, GAUSSPROD, EPANCONV1, EPANPRODCONV, GAUSSCONV1, GAUSSPRODCONV = range(1, 9)
CASES = [(GAUSS, (0.03, 0.03)), (GAUSSPROD, (0.025, 0.04)), (GAUSSPRODCONV, (0.025, 0.04)),
         (EPAN, (0.06, 0.06)), (EPANPROD, (0.05, 0.08)), (EPANPRODCONV, (0.05, 0.08))]
EPAN_KINDS = (EPAN, EPANPROD, EPANPRODCONV)


def kd_mixture(n, seed):
    pts, w = ds.gaussian_mixture(n, seed=seed)
    perm = ds.kd_order(pts, 64)
    pts, w = pts[perm], w[perm]
    return pts, w / w.sum()


def evaluate(bk, kde, desc, nq, order):
    kde.set_option(bk.BK_OPT_QUERY_ORDER, order)
    return kde.eval(desc, 0, kde.reference_size() - 1, 0, nq - 1)


@pytest.mark.parametrize("adaptive", [False, True])
@pytest.mark.parametrize("kind,h", CASES)
def test_spatial_order_matches_oracle(bk, oracle, kind, h, adaptive):
    n_ref, n_q = 24_000, 9_000
    pts, mass = kd_mixture(n_ref, 500 + kind)
    q = ds.uniform_box(n_q, -4.0, 4.0, 600 + kind)
    abw = np.random.default_rng(7).uniform(0.7, 1.5, n_ref) if adaptive else None
    desc = bk.KernelDesc(kind, 2, *h)
    with bk.DirectKde(pts, q, mass, abw) as kde:
        given = evaluate(bk, kde, desc, n_q, bk.BK_QUERY_ORDER_GIVEN)
        spatial = evaluate(bk, kde, desc, n_q, bk.BK_QUERY_ORDER_SPATIAL)
        again = evaluate(bk, kde, desc, n_q, bk.BK_QUERY_ORDER_SPATIAL)
    want = oracle.direct(kind, h[0], h[1], f32(pts), f32(q), mass=f32(mass), abw=f32(abw), normalize=True)
    bound = oracle.normalization(kind, 2, *h) * mass.sum()
    assert_parity(spatial, want, bound, what=f"spatial kind {kind} adaptive {adaptive}")
    assert_parity(given, want, bound, what=f"given kind {kind} adaptive {adaptive}")
    np.testing.assert_array_equal(spatial, again)  # deterministic order
    if kind in EPAN_KINDS:
        np.testing.assert_array_equal(spatial, given)
    assert (want > 0).mean() > 0.05  # the case has queries inside the support


@pytest.mark.parametrize("f64", [False, True])
@pytest.mark.parametrize("kind", [GAUSS, EPANPROD])
def test_spatial_order_sub_range_and_tiny_sets(bk, oracle, kind, f64):
    # unaligned reference ranges start inside a culling tile
    pts, mass = kd_mixture(5_000, 77)
    q = ds.uniform_box(3_000, -3.0, 3.0, 78)
    h = (0.04, 0.04) if kind == GAUSS else (0.06, 0.09)
    desc = bk.KernelDesc(kind, 2, *h)
    P, Q, M = (pts, q, mass) if f64 else (f32(pts), f32(q), f32(mass))
    rtol = 1e-12 if f64 else 1e-5
    prec = bk.BK_PREC_F64 if f64 else bk.BK_PREC_F32
    with bk.DirectKde(pts, q, mass, precision=prec) as kde:
        for order in (bk.BK_QUERY_ORDER_SPATIAL, bk.BK_QUERY_ORDER_GIVEN):
            kde.set_option(bk.BK_OPT_QUERY_ORDER, order)
            for (r_i, r_j, q_i, q_j) in [(100, 4321, 250, 2999), (0, 4999, 0, 0), (17, 17, 5, 1400),
                                         (0, 4999, 10, 300), (300, 2900, 0, 2999)]:
                got = kde.unnormalized_eval(desc, r_i, r_j, q_i, q_j)
                want = oracle.direct(kind, h[0], h[1], P[r_i:r_j + 1], Q[q_i:q_j + 1], mass=M[r_i:r_j + 1])
                assert_parity(got, want, M[r_i:r_j + 1].sum(), rtol=rtol,
                              what=f"order {order} range {(r_i, r_j, q_i, q_j)}")


def test_spatial_order_identical_queries(bk, oracle):
    # all queries at one point: the sort's box is degenerate
    pts, mass = kd_mixture(4_000, 91)
    q = np.tile([[0.3, -0.2]], (2_000, 1))
    with bk.DirectKde(pts, q, mass) as kde:
        kde.set_option(bk.BK_OPT_QUERY_ORDER, bk.BK_QUERY_ORDER_SPATIAL)
        got = kde.eval(bk.KernelDesc(EPANPROD, 2, 0.2, 0.3), 0, 3999, 0, 1999)
    want = oracle.direct(EPANPROD, 0.2, 0.3, f32(pts), f32(q), mass=f32(mass), normalize=True)
    assert_parity(got, want, oracle.normalization(EPANPROD, 2, 0.2, 0.3))
    assert np.all(got == got[0])


@pytest.mark.parametrize("kind,h", [(GAUSS, (0.02, 0.02)), (EPANPROD, (0.04, 0.05))])
def test_self_eval_culls_in_tree_order(bk, oracle, kind, h):
    # references in kd order are their own sorted queries: the leave-one-out
    # self evaluation culls without any option
    pts, mass = kd_mixture(30_000, 300 + kind)
    desc = bk.KernelDesc(kind, 2, *h)
    with bk.DirectKde(pts, None, mass) as kde:
        got = kde.self_eval(desc, exclude_self=True)
    want = oracle.direct(kind, h[0], h[1], f32(pts), f32(pts), mass=f32(mass), exclude_self=True, normalize=True)
    assert_parity(got, want, oracle.normalization(kind, 2, *h) * mass.sum(), what=f"loo kind {kind}")


@pytest.mark.parametrize("kind,h", [(EPAN, (0.06, 0.06)), (EPANPROD, (0.05, 0.08)), (EPANPRODCONV, (0.05, 0.08))])
def test_epanechnikov_culling_far_from_origin(bk, oracle, kind, h):
    # Points 5000 units from the origin: the scaled offsets s r reach 1e5, so
    # the culling test's allowance for their rounding is ~1e-10 in u; culled
    # sums must still equal the unculled ones bit for bit and meet the oracle.
    pts, mass = kd_mixture(20_000, 700 + kind)
    pts = pts + np.array([5000.0, -3000.0])
    q = ds.uniform_box(6_000, -4.0, 4.0, 800 + kind) + np.array([5000.0, -3000.0])
    desc = bk.KernelDesc(kind, 2, *h)
    with bk.DirectKde(pts, q, mass) as kde:
        given = evaluate(bk, kde, desc, len(q), bk.BK_QUERY_ORDER_GIVEN)
        spatial = evaluate(bk, kde, desc, len(q), bk.BK_QUERY_ORDER_SPATIAL)
    np.testing.assert_array_equal(spatial, given)
    want = oracle.direct(kind, h[0], h[1], f32(pts), f32(q), mass=f32(mass), normalize=True)
    assert_parity(given, want, oracle.normalization(kind, 2, *h) * mass.sum(), what=f"offset kind {kind}")
    assert (want > 0).mean() > 0.05


@pytest.mark.parametrize("adaptive", [False, True])
@pytest.mark.parametrize("kind,h", CASES)
def test_f64_spatial_order_and_culling(bk, oracle, kind, h, adaptive):
    # FP64 contexts: the same culling on FP64 boxes (every term of a culled
    # tile is exactly zero in FP64: Epanechnikov past its support, Gaussian
    # below 2^-1100), held to the FP64 bar of 1e-12 in both orders; with
    # adaptive bandwidths the reach of a tile follows its largest abw
    n_ref, n_q = 20_000, 7_000
    pts, mass = kd_mixture(n_ref, 900 + kind)
    q = ds.uniform_box(n_q, -4.0, 4.0, 950 + kind)
    abw = np.random.default_rng(17).uniform(0.6, 1.6, n_ref) if adaptive else None
    desc = bk.KernelDesc(kind, 2, *h)
    with bk.DirectKde(pts, q, mass, abw, precision=bk.BK_PREC_F64) as kde:
        given = evaluate(bk, kde, desc, n_q, bk.BK_QUERY_ORDER_GIVEN)
        spatial = evaluate(bk, kde, desc, n_q, bk.BK_QUERY_ORDER_SPATIAL)
    want = oracle.direct(kind, h[0], h[1], pts, q, mass=mass, abw=abw, normalize=True)
    bound = oracle.normalization(kind, 2, *h) * mass.sum()
    assert_parity(spatial, want, bound, rtol=1e-12, what=f"f64 spatial kind {kind} adaptive {adaptive}", kind=kind)
    assert_parity(given, want, bound, rtol=1e-12, what=f"f64 given kind {kind} adaptive {adaptive}", kind=kind)
    if kind in EPAN_KINDS:
        np.testing.assert_array_equal(spatial, given)
    assert (want > 0).mean() > 0.05


CASES_1D = [(GAUSS, (0.02, 0.02)), (EPAN, (0.04, 0.04)), (EPANCONV1, (0.04, 0.04)), (GAUSSCONV1, (0.02, 0.02))]


@pytest.mark.parametrize("f64", [False, True])
@pytest.mark.parametrize("kind,h", CASES_1D)
def test_1d_spatial_order_and_culling(bk, oracle, kind, h, f64):
    # 1-D sets: the Z-curve order is the order along x; the culling and the
    # bars are those of the 2-D cases (bit-identical orders for the
    # Epanechnikov kinds, 1e-5 / 1e-12 against the oracle)
    n_ref, n_q = 20_000, 7_000
    pts, w = ds.gaussian_mixture(n_ref, seed=1100 + kind)
    pts = pts[:, :1]
    perm = ds.kd_order(pts, 64)
    pts, mass = pts[perm], (w / w.sum())[perm]
    q = ds.uniform_box(n_q, -4.0, 4.0, 1150 + kind, dim=1)
    desc = bk.KernelDesc(kind, 1, *h)
    prec = bk.BK_PREC_F64 if f64 else bk.BK_PREC_F32
    with bk.DirectKde(pts, q, mass, precision=prec) as kde:
        given = evaluate(bk, kde, desc, n_q, bk.BK_QUERY_ORDER_GIVEN)
        spatial = evaluate(bk, kde, desc, n_q, bk.BK_QUERY_ORDER_SPATIAL)
    P, Q, M = (pts, q, mass) if f64 else (f32(pts), f32(q), f32(mass))
    want = oracle.direct(kind, h[0], h[1], P, Q, mass=M, normalize=True)
    bound = oracle.normalization(kind, 1, *h) * mass.sum()
    rtol = 1e-12 if f64 else 1e-5
    assert_parity(spatial, want, bound, rtol=rtol, what=f"1-D spatial kind {kind} f64 {f64}", kind=kind)
    assert_parity(given, want, bound, rtol=rtol, what=f"1-D given kind {kind} f64 {f64}", kind=kind)
    if kind in (EPAN, EPANCONV1):
        np.testing.assert_array_equal(spatial, given)
    assert (want > 0).mean() > 0.05


@pytest.mark.parametrize("kind,h", [(GAUSS, (0.02, 0.02)), (EPANPROD, (0.04, 0.05))])
def test_f64_self_eval_culls_in_tree_order(bk, oracle, kind, h):
    pts, mass = kd_mixture(30_000, 310 + kind)
    desc = bk.KernelDesc(kind, 2, *h)
    with bk.DirectKde(pts, None, mass, precision=bk.BK_PREC_F64) as kde:
        got = kde.self_eval(desc, exclude_self=True)
    want = oracle.direct(kind, h[0], h[1], pts, pts, mass=mass, exclude_self=True, normalize=True)
    assert_parity(got, want, oracle.normalization(kind, 2, *h) * mass.sum(), rtol=1e-12, what=f"f64 loo {kind}", kind=kind)


def test_query_order_option_is_checked(bk):
    pts, mass = kd_mixture(1_000, 5)
    with bk.DirectKde(pts, pts, mass) as kde:
        with pytest.raises(ValueError):
            kde.set_option(bk.BK_OPT_QUERY_ORDER, 7)


def test_spatial_order_full_size_is_bit_identical_for_epanechnikov(bk):
    # The headline workload (1M kd-ordered mixture references x 1M uniform
    # queries), EpanechnikovProductKernel2d (0.08, 0.12): every culled term is
    # exactly zero, so the sorted evaluation equals the given-order one bit
    # for bit; the Gaussian kernel (h = 0.05) agrees to the terms below 2^-100
    # of the largest mass that culling drops.
    pts, w = ds.gaussian_mixture(1_000_000, seed=ds.SEED_MIXTURE_1M)
    perm = ds.kd_order(pts, 64)
    pts, w = pts[perm], w[perm]
    mass = w / w.sum()
    q = ds.uniform_box(1_000_000, -4.0, 4.0, ds.SEED_MIXTURE_1M + 1)
    with bk.DirectKde(pts, q, mass) as kde:
        for desc, exact in ((bk.KernelDesc(EPANPROD, 2, 0.08, 0.12), True), (bk.KernelDesc(GAUSS, 2, 0.05), False)):
            given = evaluate(bk, kde, desc, len(q), bk.BK_QUERY_ORDER_GIVEN)
            spatial = evaluate(bk, kde, desc, len(q), bk.BK_QUERY_ORDER_SPATIAL)
            if exact:
                np.testing.assert_array_equal(spatial, given)
            else:
                bound = 1.0 / (2 * np.pi * 0.05 ** 2)
                assert np.all(np.abs(spatial - given) <= 1e-12 * np.maximum(np.abs(given), 1e-10 * bound))
