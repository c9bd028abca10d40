"""The instrumented FP32 Gaussian kernel (BK_OPT_COUNT_TERMS): it counts the
pair terms its wavefronts evaluate and returns the same results as the plain
kernel.

Exact cases: with every tile in the difference form nothing is passed over, so
the real-query count is n_query x n_ref and the lane count is that of whole
workgroups (ceil(n_query / 256) x 256 lanes per reference, one query per lane;
the padding lanes of the last workgroup run too).
With the default forms, the count falls below the nominal pairs by the tiles
whose factor is zero for a whole wavefront; the results stay bit-identical.
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
    pts, mass, q = _case(20_000, 3_001, 61)
    desc = bk.KernelDesc(GAUSS, 2, 0.05)
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
