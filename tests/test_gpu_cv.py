"""GPU parity of the batched leaf-pair evaluation, the leave-one-out self
sums, the cross-validation scores and the adaptive pass (C ABI vs oracle).

LCV and LSCV are FP64 reductions over FP32 per-point sums; each per-point
sum carries <= 1e-5 relative error, so the scores are compared with an
absolute tolerance of 1e-5 * sum_i w_i |term_i| (a bound on the propagated
error).  lambda of the adaptive pass is a power -alpha of the pilot, so its
relative error is <= alpha * 1e-5.  FP64 contexts (the reference's double
kernels) are held to 1e-12 on per-point sums, 1e-10 on the scores and 1e-11 on
the adaptive pass, against the oracle on the FP64 points.
"""
from __future__ import annotations

import numpy as np
import pytest

from bbrcit_kde_amd import datasets as ds
from tests.test_gpu_direct import assert_parity, f32

pytestmark = pytest.mark.gpu

EPAN, GAUSS, EPANPROD, GAUSSPROD = 1, 2, 3, 4


def mixture(n, seed):
    pts, w = ds.gaussian_mixture(n, seed=seed)
    perm = ds.kd_order(pts, 32)
    pts, w = pts[perm], w[perm]
    return pts, w / w.sum()


@pytest.mark.parametrize("kind,h", [(GAUSSPROD, (0.09, 0.13)), (EPANPROD, (0.09, 0.13)), (EPAN, (0.1, 0.1))])
def test_leaf_pairs_match_individual_ranges(bk, oracle, kind, h):
    # unaligned reference and query ranges: the Epanechnikov work items start
    # inside sub-tiles and end in partial ones
    pts, mass = mixture(6000, 31)
    q = ds.uniform_box(4000, -3, 3, 32)
    rng = np.random.default_rng(1)
    pairs = []
    for _ in range(40):
        r_i = int(rng.integers(0, 6000))
        r_j = int(rng.integers(r_i, min(6000, r_i + 3000)))
        q_i = int(rng.integers(0, 4000))
        q_j = int(rng.integers(q_i, min(4000, q_i + 2500)))
        pairs.append((r_i, r_j, q_i, q_j))
    pairs.append((0, 5999, 0, 3999))
    pairs.append((17, 17, 3999, 3999))
    desc = bk.KernelDesc(kind, 2, *h)
    P, Q, M = f32(pts), f32(q), f32(mass)
    with bk.DirectKde(pts, q, mass) as kde:
        outs = kde.eval_leaf_pairs(desc, pairs)
        single = kde.unnormalized_eval(desc, *pairs[5])
    for (r_i, r_j, q_i, q_j), got in zip(pairs, outs):
        want = oracle.direct(kind, *h, P[r_i:r_j + 1], Q[q_i:q_j + 1], mass=M[r_i:r_j + 1])
        assert_parity(got, want, M[r_i:r_j + 1].sum(), what=f"kind {kind} pair {(r_i, r_j, q_i, q_j)}")
    np.testing.assert_allclose(outs[5], single, rtol=1e-6)


@pytest.mark.parametrize("kind", [GAUSS, EPAN, GAUSSPROD, EPANPROD])
def test_self_eval_with_and_without_self(bk, oracle, kind):
    pts, mass = mixture(5000, 40 + kind)
    h = (0.06, 0.08) if kind in (GAUSS, GAUSSPROD) else (0.2, 0.25)
    if kind in (GAUSS, EPAN):
        h = (h[0], h[0])
    desc = bk.KernelDesc(kind, 2, *h)
    P, M = f32(pts), f32(mass)
    with bk.DirectKde(pts, None, mass) as kde:
        full = kde.self_eval(desc, exclude_self=False, normalize=False)
        loo = kde.self_eval(desc, exclude_self=True, normalize=False)
    assert_parity(full, oracle.direct(kind, *h, P, P, mass=M), M.sum(), what="full")
    assert_parity(loo, oracle.direct(kind, *h, P, P, mass=M, exclude_self=True), M.sum(), what="loo")


def test_isolated_point_loo_is_exact(bk):
    # An isolated point's leave-one-out sum is tiny: by subtraction it would
    # cancel in FP32; by exclusion it must keep full relative accuracy.
    pts = np.array([[0.0, 0.0], [0.01, 0.0], [1.5, 1.5], [0.0, 0.013]])
    mass = np.array([0.25, 0.25, 0.25, 0.25])
    desc = bk.KernelDesc(GAUSS, 2, 0.5)
    with bk.DirectKde(pts, None, mass) as kde:
        loo = kde.self_eval(desc, exclude_self=True, normalize=False)
    d = pts[:, None, :] - pts[None, :, :]
    k = 0.25 * np.exp(-0.5 * (d ** 2).sum(-1) / 0.25)
    np.fill_diagonal(k, 0.0)
    want = k.sum(1)
    np.testing.assert_allclose(loo, want, rtol=1e-5)
    assert loo[2] > 0


@pytest.mark.parametrize("kind,dim,h", [(GAUSSPROD, 2, (0.05, 0.07)), (EPANPROD, 2, (0.15, 0.2)),
                                        (GAUSS, 1, (0.05, 0.05)), (EPAN, 1, (0.12, 0.12))])
def test_likelihood_and_lsq_cv(bk, oracle, kind, dim, h):
    pts, w = mixture(8000, 60 + kind)
    pts = pts[:, :dim]
    desc = bk.KernelDesc(kind, dim, *h)
    P, W = f32(pts), w
    with bk.DirectKde(pts, None, w) as kde:
        lcv = kde.likelihood_cv(desc, w)
        lscv = kde.lsq_conv_cv(desc, w)
    want_lcv = oracle.likelihood_cv(kind, *h, P, W, mass=f32(w))
    want_lscv = oracle.lsq_conv_cv(kind, *h, P, W, mass=f32(w))
    if np.isinf(want_lcv):
        # a point with no neighbour inside a compact kernel: log(0), as in the
        # reference (KernelDensityImpl.h:244-248)
        assert lcv == want_lcv
    else:
        assert abs(lcv - want_lcv) <= 1e-5 * (1 + abs(want_lcv))
    assert abs(lscv - want_lscv) <= 1e-5 * abs(want_lscv) + 1e-8


def test_f64_leaf_pairs_and_self_sums(bk, oracle):
    # FP64 contexts (the reference's double kernels): leaf pairs and the
    # self sums with and without the self pair, at the FP64 bar of 1e-12
    pts, mass = mixture(5000, 131)
    q = ds.uniform_box(3000, -3, 3, 132)
    rng = np.random.default_rng(3)
    pairs = []
    for _ in range(20):
        r_i = int(rng.integers(0, 5000))
        r_j = int(rng.integers(r_i, min(5000, r_i + 2500)))
        q_i = int(rng.integers(0, 3000))
        q_j = int(rng.integers(q_i, min(3000, q_i + 2000)))
        pairs.append((r_i, r_j, q_i, q_j))
    pairs.append((0, 4999, 0, 2999))
    for kind, h in ((GAUSSPROD, (0.09, 0.13)), (EPANPROD, (0.2, 0.25))):
        desc = bk.KernelDesc(kind, 2, *h)
        with bk.DirectKde(pts, q, mass, precision=bk.BK_PREC_F64) as kde:
            outs = kde.eval_leaf_pairs(desc, pairs)
        for (r_i, r_j, q_i, q_j), got in zip(pairs, outs):
            want = oracle.direct(kind, *h, pts[r_i:r_j + 1], q[q_i:q_j + 1], mass=mass[r_i:r_j + 1])
            assert_parity(got, want, mass[r_i:r_j + 1].sum(), rtol=1e-12, what=f"f64 kind {kind} pair {r_i, r_j}", kind=kind)
        with bk.DirectKde(pts, None, mass, precision=bk.BK_PREC_F64) as kde:
            full = kde.self_eval(desc, exclude_self=False, normalize=False)
            loo = kde.self_eval(desc, exclude_self=True, normalize=False)
        assert_parity(full, oracle.direct(kind, *h, pts, pts, mass=mass), mass.sum(), rtol=1e-12, what="f64 full", kind=kind)
        assert_parity(loo, oracle.direct(kind, *h, pts, pts, mass=mass, exclude_self=True), mass.sum(),
                      rtol=1e-12, what="f64 loo", kind=kind)


@pytest.mark.parametrize("kind,dim,h", [(GAUSSPROD, 2, (0.05, 0.07)), (EPANPROD, 2, (0.15, 0.2)),
                                        (GAUSS, 1, (0.05, 0.05)), (EPAN, 1, (0.12, 0.12))])
def test_f64_likelihood_and_lsq_cv(bk, oracle, kind, dim, h):
    # per-point sums within 1e-12: the FP64 scores within 1e-10 relative
    pts, w = mixture(6000, 160 + kind)
    pts = pts[:, :dim]
    desc = bk.KernelDesc(kind, dim, *h)
    with bk.DirectKde(pts, None, w, precision=bk.BK_PREC_F64) as kde:
        lcv = kde.likelihood_cv(desc, w)
        lscv = kde.lsq_conv_cv(desc, w)
    want_lcv = oracle.likelihood_cv(kind, *h, pts, w, mass=w)
    want_lscv = oracle.lsq_conv_cv(kind, *h, pts, w, mass=w)
    if np.isinf(want_lcv):
        assert lcv == want_lcv
    else:
        assert abs(lcv - want_lcv) <= 1e-10 * (1 + abs(want_lcv))
    assert abs(lscv - want_lscv) <= 1e-10 * abs(want_lscv) + 1e-12


def test_f64_adapt_matches_oracle(bk, oracle):
    pts, w = mixture(6000, 177)
    desc = bk.KernelDesc(GAUSSPROD, 2, 0.05, 0.05)
    with bk.DirectKde(pts, None, w, precision=bk.BK_PREC_F64) as kde:
        g, lam = kde.adapt(desc, 0.5, w)
        after = kde.self_eval(desc, exclude_self=False, normalize=True)
    g_want, lam_want = oracle.adapt(GAUSSPROD, 0.05, 0.05, pts, w, 0.5)
    assert g == pytest.approx(g_want, rel=1e-11)
    np.testing.assert_allclose(lam, lam_want, rtol=1e-11)
    want_after = oracle.direct(GAUSSPROD, 0.05, 0.05, pts, pts, mass=w * lam ** -2.0, abw=lam, normalize=True)
    assert_parity(after, want_after, oracle.normalization(GAUSSPROD, 2, 0.05, 0.05) * (w * lam ** -2.0).sum(),
                  rtol=1e-12, what="f64 adapted self density", kind=GAUSSPROD)


def test_adapt_matches_oracle(bk, oracle):
    pts, w = mixture(8000, 77)
    desc = bk.KernelDesc(GAUSSPROD, 2, 0.05, 0.05)
    P = f32(pts)
    with bk.DirectKde(pts, None, w) as kde:
        g, lam = kde.adapt(desc, 0.5, w)
        # the references now carry abw = lambda, mass = w lambda^-2
        after = kde.self_eval(desc, exclude_self=False, normalize=True)
    g_want, lam_want = oracle.adapt(GAUSSPROD, 0.05, 0.05, P, f32(w), 0.5)
    assert g == pytest.approx(g_want, rel=1e-5)
    np.testing.assert_allclose(lam, lam_want, rtol=1e-5)
    assert abs((w * np.log(lam)).sum()) < 1e-6
    want_after = oracle.direct(GAUSSPROD, 0.05, 0.05, P, P, mass=f32(w * lam ** -2.0), abw=f32(lam),
                               normalize=True)
    assert_parity(after, want_after, oracle.normalization(GAUSSPROD, 2, 0.05, 0.05) * (w * lam ** -2.0).sum(),
                  what="adapted self density")


def test_loo_rescue_of_underflowed_gaussian_sums(bk, oracle):
    # Sparse 1-D data with a narrow Gaussian: several leave-one-out sums
    # underflow FP32 (below 2^-149 of the mass scale) and are recomputed in
    # FP64 before the logarithm, so the score stays finite and exact.
    rng = np.random.default_rng(12)
    pts = np.sort(rng.uniform(-50, 50, 25))[:, None]
    w = np.full(25, 1 / 25)
    h = 0.3
    desc = bk.KernelDesc(GAUSS, 1, h)
    want = oracle.direct(GAUSS, h, h, f32(pts), f32(pts), mass=f32(w), exclude_self=True)
    tiny = (want > 0) & (want < 1e-40)
    assert tiny.sum() >= 3 and (want > 0).all()  # the case really exercises underflow
    with bk.DirectKde(pts, None, w) as kde:
        loo = kde.self_eval(desc, exclude_self=True, normalize=False)
        lcv = kde.likelihood_cv(desc, w)
    np.testing.assert_allclose(loo, want, rtol=1e-5)
    np.testing.assert_allclose(loo[tiny], want[tiny], rtol=1e-9)
    want_lcv = oracle.likelihood_cv(GAUSS, h, h, f32(pts), w, mass=f32(w))
    assert np.isfinite(want_lcv) and lcv == pytest.approx(want_lcv, rel=1e-6)


def test_adapt_alpha_zero_and_invalid(bk):
    pts, w = mixture(500, 3)
    desc = bk.KernelDesc(GAUSSPROD, 2, 0.1, 0.1)
    with bk.DirectKde(pts, None, w) as kde:
        _, lam = kde.adapt(desc, 0.0, w)
        np.testing.assert_array_equal(lam, 1.0)
        with pytest.raises(ValueError):
            kde.adapt(desc, 1.5, w)
        with pytest.raises(ValueError):
            kde.adapt(desc, -0.1, w)
