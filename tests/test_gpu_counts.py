"""Term counting (BK_OPT_COUNT_TERMS) of the FP32-point kernels: the
instrumented Gaussian kernel and the Epanechnikov kernel count the pair terms
their wavefronts evaluate and return the same results as without counting.

Exact cases: with every tile in the difference form and a bandwidth wide
enough that every query lies within reach of every sub-tile, nothing is passed
over, so the real-query count is n_query x n_ref and the lane count is that of
whole workgroups (ceil(n_query / 256) x 256 lanes per reference, one query per
lane; the padding lanes of the last workgroup run too).
With a narrow bandwidth the count falls below the nominal pairs by the
(sub-)tiles beyond the reach of a whole wavefront (each term left out is below
2^-56 of its mass); the results stay bit-identical to the plain kernel's and
within the parity bar of the FP64 oracle.
The Epanechnikov kernel passes over a 64-reference sub-tile when the
sub-tile's reach box holds none of the wavefront's queries; its count is
checked against a CPU replay: at least every (wavefront, sub-tile) pair with a
nonzero term, at most those with a query within the sub-tile's box widened by
the support.
"""
from __future__ import annotations

import numpy as np
import pytest

from bbrcit_kde_amd import datasets as ds

pytestmark = pytest.mark.gpu

GAUSS = 2


def _case(n_ref, n_q, seed):
    pts, w = ds.gaussian_mixture(n_ref, seed=seed)
    pts = pts[ds.kd_order(pts, 64)]
    q = ds.uniform_box(n_q, -4.0, 4.0, seed + 1)
    return pts, np.full(n_ref, 1.0 / n_ref), q


def test_difference_form_counts_every_pair(bk):
    # h = 2: a term is left out only beyond d = sqrt(56 / g) = 8.8 h = 17.6 with
    # g = log2(e) / (2 h^2), farther than any query lies from any reference
    # here (under 11), so nothing is passed over
    pts, mass, q = _case(20_000, 3_001, 61)
    desc = bk.KernelDesc(GAUSS, 2, 2.0)
    with bk.DirectKde(pts, q, mass) as kde:
        kde.set_option(bk.BK_OPT_GAUSS_FORM, bk.BK_GAUSS_FORM_DIRECT)
        plain = kde.eval(desc, 0, len(pts) - 1, 0, len(q) - 1)
        kde.set_option(bk.BK_OPT_COUNT_TERMS, 1)
        counted = kde.eval(desc, 0, len(pts) - 1, 0, len(q) - 1)
        lanes, real = kde.term_count()
        assert kde.term_count() == (0, 0)  # reading resets
        kde.set_option(bk.BK_OPT_COUNT_TERMS, 0)
    assert np.array_equal(plain, counted)
    assert real == len(q) * len(pts)
    assert lanes == -(-len(q) // 256) * 256 * len(pts)


def test_difference_form_passes_over_subtiles_out_of_reach(bk, oracle):
    from tests.test_gpu_direct import assert_parity, f32

    pts, mass, q = _case(20_000, 3_001, 63)
    desc = bk.KernelDesc(GAUSS, 2, 0.05)
    with bk.DirectKde(pts, q, mass) as kde:
        kde.set_option(bk.BK_OPT_GAUSS_FORM, bk.BK_GAUSS_FORM_DIRECT)
        plain = kde.unnormalized_eval(desc, 0, len(pts) - 1, 0, len(q) - 1)
        kde.set_option(bk.BK_OPT_COUNT_TERMS, 1)
        counted = kde.unnormalized_eval(desc, 0, len(pts) - 1, 0, len(q) - 1)
        lanes, real = kde.term_count()
        kde.set_option(bk.BK_OPT_COUNT_TERMS, 0)
    assert np.array_equal(plain, counted)
    assert 0 < real < len(q) * len(pts)
    want = oracle.direct(GAUSS, 0.05, 0.05, f32(pts), f32(q), mass=f32(mass))
    assert_parity(plain, want, mass.sum(), what="difference form, sub-tiles passed over")


def test_default_forms_count_fewer_terms_and_same_results(bk):
    pts, mass, q = _case(60_000, 20_000, 71)
    desc = bk.KernelDesc(GAUSS, 2, 0.05)
    with bk.DirectKde(pts, q, mass) as kde:
        plain = kde.eval(desc, 0, len(pts) - 1, 0, len(q) - 1)
        kde.set_option(bk.BK_OPT_COUNT_TERMS, 1)
        counted = kde.eval(desc, 0, len(pts) - 1, 0, len(q) - 1)
        lanes, real = kde.term_count()
        # a sub-range counts only its own references and queries
        kde.eval(desc, 100, 8_099, 5, 1_004)
        _, real_sub = kde.term_count()
        with pytest.raises(Exception):
            kde.set_option(bk.BK_OPT_COUNT_TERMS, 0)
            kde.term_count()
    assert np.array_equal(plain, counted)
    nominal = len(q) * len(pts)
    assert 0 < real < nominal
    assert real <= lanes
    assert 0 < real_sub <= 8_000 * 1_000


# ---- Epanechnikov kinds (FP32 points): wavefronts of 64 queries pass over
# 64-reference sub-tiles whose terms are all exactly zero for them ----------
EPANPROD = 3


def _wave_queries(n_q):
    # query indices of each wavefront (workgroups of 256 queries, one per
    # lane; lanes past the last query repeat their workgroup's first query)
    # and its number of lanes that hold a real query
    n_blocks = -(-n_q // 256)
    lane = np.arange(n_blocks * 256)
    idx = np.where(lane < n_q, lane, (lane // 256) * 256)
    return idx.reshape(-1, 64), (lane < n_q).reshape(-1, 64).sum(1)


def _subtile_counts(pts, q, hx, hy, widen):
    # per (wavefront, sub-tile): does some query lie within (hx, hy) * widen
    # of some reference on both axes?  (widen = 1: a nonzero term exists)
    waves, reals = _wave_queries(len(q))
    n = len(pts)
    lanes_lo = real_lo = 0
    lanes_hi = real_hi = 0
    for s0 in range(0, n, 64):
        r = pts[s0:s0 + 64]
        cnt = len(r)
        lo, hi = r.min(0), r.max(0)
        for w, nreal in zip(waves, reals):
            qw = q[w]
            # a query within the widened box (upper bound), a nonzero pair
            # (lower bound)
            near_box = np.any((qw[:, 0] > lo[0] - widen * hx) & (qw[:, 0] < hi[0] + widen * hx) &
                              (qw[:, 1] > lo[1] - widen * hy) & (qw[:, 1] < hi[1] + widen * hy))
            if not near_box:
                continue
            lanes_hi += 64 * cnt
            real_hi += int(nreal) * cnt
            dx = np.abs(qw[:, None, 0] - r[None, :, 0]) < hx
            dy = np.abs(qw[:, None, 1] - r[None, :, 1]) < hy
            if np.any(dx & dy):
                lanes_lo += 64 * cnt
                real_lo += int(nreal) * cnt
    return (lanes_lo, real_lo), (lanes_hi, real_hi)


def test_epanechnikov_counts_every_pair_when_all_in_reach(bk):
    pts, mass, q = _case(20_000, 3_001, 81)
    desc = bk.KernelDesc(EPANPROD, 2, 20.0, 20.0)
    with bk.DirectKde(pts, q, mass) as kde:
        plain = kde.eval(desc, 0, len(pts) - 1, 0, len(q) - 1)
        kde.set_option(bk.BK_OPT_COUNT_TERMS, 1)
        counted = kde.eval(desc, 0, len(pts) - 1, 0, len(q) - 1)
        lanes, real = kde.term_count()
        kde.set_option(bk.BK_OPT_COUNT_TERMS, 0)
    assert np.array_equal(plain, counted)
    assert real == len(q) * len(pts)
    assert lanes == -(-len(q) // 256) * 256 * len(pts)


def test_epanechnikov_passes_over_exactly_the_zero_subtiles(bk, oracle):
    # The count lies between the (wavefront, sub-tile) pairs that hold a
    # nonzero term (these must be evaluated) and those with a query within
    # the sub-tile's box widened by the support and a 1e-4 margin (the reach
    # boxes are wider than the support by 2^-20 and rounding only), and the
    # sums stay bit-identical to those of the spatially sorted evaluation.
    from tests.test_gpu_direct import assert_parity, f32

    pts, mass, q = _case(20_000, 3_001, 83)
    hx, hy = 0.05, 0.08
    desc = bk.KernelDesc(EPANPROD, 2, hx, hy)
    with bk.DirectKde(pts, q, mass) as kde:
        plain = kde.unnormalized_eval(desc, 0, len(pts) - 1, 0, len(q) - 1)
        kde.set_option(bk.BK_OPT_COUNT_TERMS, 1)
        counted = kde.unnormalized_eval(desc, 0, len(pts) - 1, 0, len(q) - 1)
        lanes, real = kde.term_count()
        kde.set_option(bk.BK_OPT_COUNT_TERMS, 0)
        kde.set_option(bk.BK_OPT_QUERY_ORDER, bk.BK_QUERY_ORDER_SPATIAL)
        spatial = kde.unnormalized_eval(desc, 0, len(pts) - 1, 0, len(q) - 1)
    assert np.array_equal(plain, counted)
    assert np.array_equal(plain, spatial)
    P, Q = f32(pts).astype(np.float64), f32(q).astype(np.float64)
    (lanes_lo, real_lo), (lanes_hi, real_hi) = _subtile_counts(P, Q, hx, hy, 1.0 + 1e-4)
    assert 0 < real_lo <= real <= real_hi < len(q) * len(pts)
    assert lanes_lo <= lanes <= lanes_hi
    want = oracle.direct(EPANPROD, hx, hy, f32(pts), f32(q), mass=f32(mass))
    assert_parity(plain, want, mass.sum(), what="Epanechnikov product, sub-tiles passed over")
