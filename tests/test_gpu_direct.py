"""GPU parity of the direct sums through the C ABI, against the FP64 oracle.

Tolerance: FP32 contexts must agree with the FP64 oracle (run on the same
FP32-rounded points, masses and abw) to a relative error of 1e-5 wherever
the exact value is at least FLOOR_REL times the largest value of the case;
below that floor the absolute error must stay under 1e-5 * FLOOR_REL * max.
FP64 contexts must agree to 1e-12.  Cases follow the reference's tests:
every kernel family, index sub-ranges as in test/test_cukde2.cc:163-257,
the range errors of CudaDirectKde::unnormalized_eval
(include/CudaDirectKde.h:181-190), ragged and single-point sizes.
"""
from __future__ import annotations

import numpy as np
import pytest

from bbrcit_kde_amd import datasets as ds

pytestmark = pytest.mark.gpu

# Floor: a fraction of an upper bound of the sum (C * sum of masses, U <= 1).
FLOOR_REL = 1e-10
RTOL_F32 = 1e-5

EPAN, GAUSS, EPANPROD, GAUSSPROD, EPANCONV1, EPANPRODCONV, GAUSSCONV1, GAUSSPRODCONV = range(1, 9)
KINDS = [(EPAN, 1), (EPAN, 2), (GAUSS, 1), (GAUSS, 2), (EPANPROD, 2), (GAUSSPROD, 2), (EPANCONV1, 1),
         (EPANPRODCONV, 2), (GAUSSCONV1, 1), (GAUSSPRODCONV, 2)]
BW = {EPAN: (0.35, 0.35), GAUSS: (0.12, 0.12), EPANPROD: (0.3, 0.45), GAUSSPROD: (0.1, 0.16),
      EPANCONV1: (0.3, 0.3), EPANPRODCONV: (0.3, 0.45), GAUSSCONV1: (0.12, 0.12), GAUSSPRODCONV: (0.1, 0.16)}


# Gaussian kinds in FP64 contexts (rtol <= FP64_RTOL_MAX) are held to their
# tolerance relative to every value down to FP64_REL_FLOOR, not only above
# FLOOR_REL * bound: FP64 tile culling skips only tiles whose terms are exactly
# zero, and a Gaussian term keeps its relative precision however small.  The
# compact kinds keep the floor: near the edge of the support a term is the
# difference 1 - u^2 (2 - u for the convolutions) of nearly equal numbers, and
# its relative error grows without bound as it goes to zero in any FP64
# evaluation, the oracle's included.
FP64_RTOL_MAX = 1e-11
FP64_REL_FLOOR = 1e-290
GAUSS_KINDS = (GAUSS, GAUSSPROD, GAUSSCONV1, GAUSSPRODCONV)


def assert_parity(got, want, bound, rtol=RTOL_F32, what="", kind=None):
    """bound: an upper bound of every exact value (C * sum of masses)."""
    got = np.asarray(got)
    want = np.asarray(want)
    fp64_tails = rtol <= FP64_RTOL_MAX and kind in GAUSS_KINDS
    floor = FP64_REL_FLOOR if fp64_tails else FLOOR_REL * bound
    above = np.abs(want) >= floor
    if above.any():
        rel = np.abs(got[above] - want[above]) / np.abs(want[above])
        assert rel.max() <= rtol, f"{what}: max rel err {rel.max():.3e} at {np.argmax(rel)}"
    if (~above).any():
        assert np.abs(got[~above] - want[~above]).max() <= rtol * max(floor, 1e-300), f"{what}: below-floor"


def make_case(n_ref, n_q, dim, seed, adaptive):
    pts, w = ds.gaussian_mixture(n_ref, seed=seed)
    q = ds.uniform_box(n_q, -3.0, 3.0, seed + 1)
    pts, q = pts[:, :dim], q[:, :dim]
    mass = w / w.sum()
    rng = np.random.default_rng(seed + 2)
    abw = rng.uniform(0.6, 1.8, n_ref) if adaptive else None
    return pts, q, mass, abw


def f32(a):
    return None if a is None else ds.round_f32(a)


@pytest.mark.parametrize("adaptive", [False, True])
@pytest.mark.parametrize("kind,dim", KINDS)
def test_direct_f32_matches_oracle(bk, oracle, kind, dim, adaptive):
    pts, q, mass, abw = make_case(3001, 2050, dim, 100 + kind * 10 + dim, adaptive)
    hx, hy = BW[kind]
    desc = bk.KernelDesc(kind, dim, hx, hy)
    with bk.DirectKde(pts, q, mass, abw) as kde:
        got = kde.eval(desc, 0, len(pts) - 1, 0, len(q) - 1)
    want = oracle.direct(kind, hx, hy, f32(pts), f32(q), mass=f32(mass), abw=f32(abw), normalize=True)
    bound = oracle.normalization(kind, dim, hx, hy) * mass.sum()
    assert_parity(got, want, bound, what=f"kind {kind} dim {dim} adaptive {adaptive}")


@pytest.mark.parametrize("kind,dim", KINDS)
def test_direct_f64_matches_oracle(bk, oracle, kind, dim):
    pts, q, mass, abw = make_case(1500, 700, dim, 300 + kind, True)
    hx, hy = BW[kind]
    desc = bk.KernelDesc(kind, dim, hx, hy)
    with bk.DirectKde(pts, q, mass, abw, precision=bk.BK_PREC_F64) as kde:
        got = kde.unnormalized_eval(desc, 0, len(pts) - 1, 0, len(q) - 1)
    want = oracle.direct(kind, hx, hy, pts, q, mass=mass, abw=abw)
    assert_parity(got, want, mass.sum(), rtol=1e-12, what=f"f64 kind {kind}", kind=kind)


def test_index_subranges_like_reference_test(bk, oracle):
    # test/test_cukde2.cc:163-257 evaluates five (ref, query) sub-ranges.
    pts, q, mass, _ = make_case(5000, 3000, 2, 17, False)
    desc = bk.KernelDesc(GAUSS, 2, 0.1)
    P, Q, M = f32(pts), f32(q), f32(mass)
    ranges = [(0, 4999, 0, 2999), (0, 0, 0, 2999), (1234, 4321, 17, 17), (4999, 4999, 2999, 2999),
              (250, 3333, 1000, 2600)]
    with bk.DirectKde(pts, q, mass) as kde:
        for r_i, r_j, q_i, q_j in ranges:
            got = kde.unnormalized_eval(desc, r_i, r_j, q_i, q_j)
            want = oracle.direct(GAUSS, 0.1, 0.1, P[r_i:r_j + 1], Q[q_i:q_j + 1], mass=M[r_i:r_j + 1])
            assert got.shape == (q_j - q_i + 1,)
            assert_parity(got, want, M[r_i:r_j + 1].sum(), what=f"range {(r_i, r_j, q_i, q_j)}")


def test_range_errors_match_reference(bk):
    pts, q, mass, _ = make_case(100, 50, 2, 5, False)
    desc = bk.KernelDesc(GAUSS, 2, 0.1)
    with bk.DirectKde(pts, q, mass) as kde:
        assert kde.reference_size() == 100 and kde.query_size() == 50
        with pytest.raises(IndexError):
            kde.unnormalized_eval(desc, 0, 100, 0, 10)  # j_r >= n_ref: out_of_range
        with pytest.raises(IndexError):
            kde.unnormalized_eval(desc, 0, 99, 0, 50)  # j_q >= n_query
        with pytest.raises(ValueError):
            kde.unnormalized_eval(desc, 10, 9, 0, 10)  # i_r > j_r: invalid_argument
        with pytest.raises(ValueError):
            kde.unnormalized_eval(desc, 0, 9, 20, 10)
        with pytest.raises(bk.BkError):
            kde.unnormalized_eval(bk.KernelDesc(EPANPROD, 1, 0.1), 0, 9, 0, 9)  # kind/dim mismatch
        with pytest.raises(ValueError):
            kde.unnormalized_eval(bk.KernelDesc(GAUSS, 2, -0.1), 0, 9, 0, 9)


@pytest.mark.parametrize("n_ref,n_q", [(1, 1), (1, 777), (255, 1), (257, 1025), (1024, 1024), (70001, 33)])
def test_ragged_sizes(bk, oracle, n_ref, n_q):
    pts, q, mass, abw = make_case(n_ref, n_q, 2, 40 + n_ref % 97, True)
    desc = bk.KernelDesc(GAUSSPROD, 2, 0.08, 0.12)
    with bk.DirectKde(pts, q, mass, abw) as kde:
        got = kde.unnormalized_eval(desc, 0, n_ref - 1, 0, n_q - 1)
    want = oracle.direct(GAUSSPROD, 0.08, 0.12, f32(pts), f32(q), mass=f32(mass), abw=f32(abw))
    assert_parity(got, want, mass.sum(), what=f"ragged {n_ref}x{n_q}")


def test_results_are_deterministic(bk):
    pts, q, mass, _ = make_case(20000, 5000, 2, 9, False)
    desc = bk.KernelDesc(GAUSS, 2, 0.05)
    with bk.DirectKde(pts, q, mass) as kde:
        a = kde.eval(desc, 0, len(pts) - 1, 0, len(q) - 1)
        b = kde.eval(desc, 0, len(pts) - 1, 0, len(q) - 1)
    assert np.array_equal(a, b)


def test_invariants_linearity_permutation(bk):
    pts, q, mass, _ = make_case(8000, 2000, 2, 21, False)
    rng = np.random.default_rng(0)
    m1 = rng.random(len(pts))
    m2 = rng.random(len(pts))
    desc = bk.KernelDesc(GAUSSPROD, 2, 0.07, 0.11)
    n, nq = len(pts), len(q)
    with bk.DirectKde(pts, q, m1 + m2) as kde:
        s12 = kde.unnormalized_eval(desc, 0, n - 1, 0, nq - 1)
    with bk.DirectKde(pts, q, m1) as kde:
        s1 = kde.unnormalized_eval(desc, 0, n - 1, 0, nq - 1)
    with bk.DirectKde(pts, q, m2) as kde:
        s2 = kde.unnormalized_eval(desc, 0, n - 1, 0, nq - 1)
    perm = rng.permutation(n)
    with bk.DirectKde(pts[perm], q, (m1 + m2)[perm]) as kde:
        sp = kde.unnormalized_eval(desc, 0, n - 1, 0, nq - 1)
    big = s12 > 1e-10 * s12.max()
    np.testing.assert_allclose((s1 + s2)[big], s12[big], rtol=3e-6)
    np.testing.assert_allclose(sp[big], s12[big], rtol=3e-6)


@pytest.mark.parametrize("form", ["auto", "direct"])
@pytest.mark.parametrize("kind,dim", [(GAUSS, 3), (GAUSS, 8), (GAUSS, 9), (GAUSS, 12), (GAUSS, 16), (EPAN, 5),
                                      (EPAN, 11), (EPAN, 16)])
def test_general_dimension(bk, oracle, kind, dim, form):
    # D >= 8 runs the Gram form (FP64 MFMA dot products) by default and the
    # FP64 difference form with BK_HIGH_DIM_DIRECT; D < 8 always the latter.
    rng = np.random.default_rng(dim)
    pts = rng.standard_normal((2001, dim))
    q = rng.standard_normal((701, dim))
    mass = np.full(len(pts), 1.0 / len(pts))
    abw = rng.uniform(0.8, 1.25, len(pts))
    h = 0.5 if kind == GAUSS else 2.0
    desc = bk.KernelDesc(kind, dim, h)
    with bk.DirectKde(pts, q, mass, abw) as kde:
        if form == "direct":
            kde.set_option(bk.BK_OPT_HIGH_DIM_FORM, bk.BK_HIGH_DIM_DIRECT)
        got = kde.eval(desc, 0, len(pts) - 1, 0, len(q) - 1)
        sub = kde.eval(desc, 17, 1500, 3, 650)
    want = oracle.direct(kind, h, h, f32(pts), f32(q), mass=f32(mass), abw=f32(abw), normalize=True)
    bound = oracle.normalization(kind, dim, h, h) * mass.sum()
    assert_parity(got, want, bound, what=f"D={dim} {form}")
    want_sub = oracle.direct(kind, h, h, f32(pts[17:1501]), f32(q[3:651]), mass=f32(mass[17:1501]),
                             abw=f32(abw[17:1501]), normalize=True)
    assert_parity(sub, want_sub, bound, what=f"D={dim} {form} sub-range")


@pytest.mark.parametrize("dim", [8, 16])
def test_gram_leave_one_out_and_forms_agree(bk, oracle, dim):
    rng = np.random.default_rng(100 + dim)
    pts = rng.standard_normal((3000, dim))
    mass = np.full(len(pts), 1.0 / len(pts))
    desc = bk.KernelDesc(GAUSS, dim, 0.5)
    with bk.DirectKde(pts, None, mass) as kde:
        loo = kde.self_eval(desc, exclude_self=True, normalize=True)
        kde.set_option(bk.BK_OPT_HIGH_DIM_FORM, bk.BK_HIGH_DIM_DIRECT)
        loo_direct = kde.self_eval(desc, exclude_self=True, normalize=True)
    want = oracle.direct(GAUSS, 0.5, 0.5, f32(pts), f32(pts), mass=f32(mass), exclude_self=True, normalize=True)
    bound = oracle.normalization(GAUSS, dim, 0.5, 0.5)
    assert_parity(loo, want, bound, what=f"Gram LOO D={dim}")
    assert_parity(loo_direct, want, bound, what=f"direct LOO D={dim}")


@pytest.mark.parametrize("kind", [GAUSS, EPAN])
def test_gram_far_from_origin(bk, oracle, kind):
    # The Gram form adds |r|^2, r.q and |q|^2 parts of the exponent; points
    # 40 units from the origin make each part ~1e4 times their sum, which the
    # FP64 parts must still resolve to the parity bar (adaptive bandwidths,
    # uneven masses, D = 12 and 16).
    for dim in (12, 16):
        rng = np.random.default_rng(300 + 10 * kind + dim)
        pts = 40.0 + rng.standard_normal((2500, dim))
        q = 40.0 + 1.2 * rng.standard_normal((900, dim))
        mass = rng.uniform(0.5, 1.5, len(pts))
        mass /= mass.sum()
        abw = rng.uniform(0.8, 1.25, len(pts))
        h = 0.8 if kind == GAUSS else 3.0
        desc = bk.KernelDesc(kind, dim, h)
        with bk.DirectKde(pts, q, mass, abw) as kde:
            got = kde.eval(desc, 0, len(pts) - 1, 0, len(q) - 1)
        want = oracle.direct(kind, h, h, f32(pts), f32(q), mass=f32(mass), abw=f32(abw), normalize=True)
        bound = oracle.normalization(kind, dim, h, h) * mass.sum()
        assert (want > 1e-10 * bound).mean() > 0.5  # most sums lie above the floor
        assert_parity(got, want, bound, what=f"offset D={dim} kind {kind}")


def test_mixture_1m_refs_sampled_queries(bk, oracle):
    # Full-size reference set of the headline workload (1M mixture points,
    # kd order), checked on sampled query ranges against the oracle.
    pts, w = ds.gaussian_mixture(1_000_000, seed=ds.SEED_MIXTURE_1M)
    perm = ds.kd_order(pts, 64)
    pts, w = pts[perm], w[perm]
    mass = w / w.sum()
    q = ds.uniform_box(1_000_000, -4.0, 4.0, ds.SEED_MIXTURE_1M + 1)
    desc = bk.KernelDesc(GAUSS, 2, 0.05)
    with bk.DirectKde(pts, q, mass) as kde:
        for q_i in (0, 500_000, 999_488):
            got = kde.eval(desc, 0, len(pts) - 1, q_i, q_i + 511)
            want = oracle.direct(GAUSS, 0.05, 0.05, f32(pts), f32(q[q_i:q_i + 512]), mass=f32(mass),
                                 normalize=True)
            assert_parity(got, want, oracle.normalization(GAUSS, 2, 0.05, 0.05), what=f"1M sample at {q_i}")


@pytest.mark.parametrize("kind,hx,hy", [(GAUSS, 0.05, 0.05), (GAUSSPROD, 0.04, 0.07), (GAUSSCONV1, 0.05, 0.05)])
def test_gaussian_forms_agree_with_oracle(bk, oracle, kind, hx, hy):
    # The tile-centred form (compact tiles) and the difference form must both
    # meet the tolerance; kd-ordered mixture points give mostly compact tiles.
    dim = 1 if kind == GAUSSCONV1 else 2
    pts, w = ds.gaussian_mixture(60000, seed=91)
    perm = ds.kd_order(pts, 64)
    pts, w = pts[perm][:, :dim], w[perm]
    mass = w / w.sum()
    rng = np.random.default_rng(5)
    abw = rng.uniform(0.7, 1.4, len(pts))
    q = ds.uniform_box(3000, -2.5, 2.5, 92)[:, :dim]
    desc = bk.KernelDesc(kind, dim, hx, hy)
    want = oracle.direct(kind, hx, hy, f32(pts), f32(q), mass=f32(mass), abw=f32(abw))
    with bk.DirectKde(pts, q, mass, abw) as kde:
        auto = kde.unnormalized_eval(desc, 0, len(pts) - 1, 0, len(q) - 1)
        kde.set_option(bk.BK_OPT_GAUSS_FORM, bk.BK_GAUSS_FORM_DIRECT)
        direct = kde.unnormalized_eval(desc, 0, len(pts) - 1, 0, len(q) - 1)
        with pytest.raises(ValueError):
            kde.set_option(bk.BK_OPT_GAUSS_FORM, 7)
    assert_parity(auto, want, mass.sum(), what="tile-centred form")
    assert_parity(direct, want, mass.sum(), what="difference form")


@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_reference_cpu_case_10k(bk, oracle, reference, precision):
    """The reference's own CPU-runnable case: 10,000 bimodal references (means
    (-1,-1), (1,1), sigma 0.5, unit weights), 10,000 uniform queries on
    [-3,3]^2, GaussianKernel h = 0.1.  Against the reference's
    KernelDensity<2,GaussianKernel<2,double>>::direct_eval (FP64) and the FP64
    oracle on the same (FP32-rounded for the FP32 context) inputs."""
    pts = ds.bimodal_gaussian(10_000, ds.SEED_BIMODAL_10K)
    q = ds.uniform_box(10_000, -3.0, 3.0, ds.SEED_BIMODAL_10K + 1)
    w = np.full(len(pts), 1.0 / len(pts))
    prec = bk.BK_PREC_F64 if precision == "f64" else bk.BK_PREC_F32
    with bk.DirectKde(pts, q, w, precision=prec) as kde:
        got = kde.eval(bk.KernelDesc(GAUSS, 2, 0.1), 0, len(pts) - 1, 0, len(q) - 1)
    want_ref = reference.direct_eval(GAUSS, 0, 0.1, 0.1, pts, np.ones(len(pts)), q, leaf_nmax=1024)
    bound = oracle.normalization(GAUSS, 2, 0.1, 0.1)
    if precision == "f64":
        assert_parity(got, want_ref, bound, rtol=1e-12, what="FP64 vs reference", kind=GAUSS)
    else:
        want = oracle.direct(GAUSS, 0.1, 0.1, f32(pts), f32(q), mass=f32(w), normalize=True)
        assert_parity(got, want, bound, what="FP32 vs oracle")
        assert_parity(got, want_ref, bound, rtol=2e-5, what="FP32 vs reference (FP64 inputs)")


@pytest.mark.parametrize("dim", [1, 2, 3])
def test_lexico_order_matches_numpy(bk, dim):
    # the device order that starts Kdtree construction: lexicographic by
    # coordinate 0, then 1, ...; exact duplicates (and -0 against +0) keep
    # their index order, as numpy's stable lexsort does
    rng = np.random.default_rng(40 + dim)
    n = 300_001
    pts = rng.normal(size=(n, dim))
    pts[::97] = pts[5]                      # exact duplicates
    pts[1000:1300, 0] = np.round(pts[1000:1300, 0], 1)  # ties in the first coordinate
    pts[7, 0], pts[8, 0] = -0.0, 0.0
    pts[20:30] = np.array([[-0.0] * dim])
    pts[30:40] = np.array([[0.0] * dim])
    got = bk.lexico_order(pts)
    keys = pts + 0.0                        # -0 -> +0
    want = np.lexsort(tuple(keys[:, d] for d in reversed(range(dim))))
    np.testing.assert_array_equal(got, want)
    assert len(bk.lexico_order(np.empty((0, dim)))) == 0
    lib = bk.load_library()
    assert lib.bk_lexico_order(0, 17, 1, None, None) == bk.BK_ERR_INVALID_ARG
    assert lib.bk_lexico_order(0, dim, 10, None, None) == bk.BK_ERR_INVALID_ARG
