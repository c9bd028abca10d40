"""The instrumented FP32 Gaussian kernel (BK_OPT_COUNT_TERMS): it counts the
pair terms its wavefronts evaluate and returns the same results as the plain
kernel.

Exact cases: with every tile in the difference form and a bandwidth wide
enough that every query lies within reach of every sub-tile, nothing is passed
over, so the real-query count is n_query x n_ref and the lane count is that of
whole workgroups (ceil(n_query / 256) x 256 lanes per reference, one query per
lane; the padding lanes of the last workgroup run too).
With a narrow bandwidth the count falls below the nominal pairs by the
(sub-)tiles beyond the reach of a whole wavefront (each term left out is below
2^-56 of its mass); the results stay bit-identical to the plain kernel's and
within the parity bar of the FP64 oracle.
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
