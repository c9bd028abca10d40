"""GPU parity of the batched cross-validation grid (bk_cv_grid_*).

The grid replaces the reference's grid_search bandwidth loop
(grid_search/grid_search.cc:366-402): one likelihood CV
(KernelDensityImpl.h:215-252) and one convolution LSCV
(KernelDensityImpl.h:142-209) per bandwidth pair.  Bars:
  * leave-one-out densities: relative error <= 1e-5 against the FP64 oracle
    on the same FP32-rounded points, masses and abw (floor 1e-10 C sum(m));
  * LCV = sum_i w_i log f_i: each log f_i carries at most the density's
    relative error, so |dLCV| <= 1e-5 sum_i w_i;
  * LSCV = A - 2B with A = sum_i w_i f*_i and B = sum_i w_i f_i (both >= 0):
    |dLSCV| <= 1e-5 (A + 2B).
Every tiling variant and the graph replay must give bit-identical scores.
"""
from __future__ import annotations

import numpy as np
import pytest

from bbrcit_kde_amd import datasets as ds
from tests.test_gpu_direct import assert_parity, f32

pytestmark = pytest.mark.gpu

EPANPROD, GAUSSPROD = 3, 4
EPANPROD_CONV, GAUSSPROD_CONV = 6, 8


def mixture(n, seed):
    pts, w = ds.gaussian_mixture(n, seed=seed)
    return pts, w / w.sum()


def masses(w, abw):
    return f32(w if abw is None else w / abw ** 2)


def oracle_scores(oracle, kind, hx, hy, pts, w, abw):
    """(lcv, lscv, A + 2B bound, LOO densities) from the FP64 oracle."""
    P, M, A = f32(pts), masses(w, abw), f32(abw)
    f = oracle.direct(kind, hx, hy, P, P, mass=M, abw=A, exclude_self=True, normalize=True)
    conv = GAUSSPROD_CONV if kind == GAUSSPROD else EPANPROD_CONV
    fc = oracle.direct(conv, hx, hy, P, P, mass=M, abw=A, exclude_self=False, normalize=True)
    lcv = oracle.likelihood_cv(kind, hx, hy, P, w, mass=M, abw=A)
    lscv = oracle.lsq_conv_cv(kind, hx, hy, P, w, mass=M, abw=A)
    bound = float((w * fc).sum() + 2.0 * (w * f).sum())
    # the scores restate sum_i w_i log f_i and A - 2B from the same vectors
    assert lscv == pytest.approx(float((w * fc).sum() - 2.0 * (w * f).sum()), rel=1e-9, abs=1e-12)
    return lcv, lscv, bound, f


def check_scores(got_lcv, got_lscv, want_lcv, want_lscv, bound, w, what):
    if np.isinf(want_lcv):
        assert got_lcv == want_lcv, what
    else:
        assert abs(got_lcv - want_lcv) <= 1e-5 * w.sum(), f"{what}: lcv {got_lcv} vs {want_lcv}"
    assert abs(got_lscv - want_lscv) <= 1e-5 * bound, f"{what}: lscv {got_lscv} vs {want_lscv}"


def norm_c(kind, hx, hy):
    return 1.0 / (2 * np.pi * hx * hy) if kind == GAUSSPROD else 0.5625 / (hx * hy)


@pytest.mark.parametrize("adaptive", [False, True])
@pytest.mark.parametrize("kind", [GAUSSPROD, EPANPROD])
def test_grid_matches_oracle(bk, oracle, kind, adaptive):
    sets = [mixture(700, 501), mixture(1900, 502)]
    rng = np.random.default_rng(7)
    abws = [rng.uniform(0.6, 1.7, len(p)) if adaptive else None for p, _ in sets]
    hx = np.array([0.05, 0.11, 0.3])
    hy = np.array([0.04, 0.07, 0.16, 0.25, 0.5])  # 3 x 5: ragged against every tiling
    with bk.CvGrid(kind, [p for p, _ in sets], [w for _, w in sets],
                   abw=abws if adaptive else None) as cv:
        lcv, lscv = cv.run(hx, hy)
        for d, (pts, w) in enumerate(sets):
            for ix, iy in [(0, 0), (1, 3), (2, 4), (2, 1)]:
                want_lcv, want_lscv, bound, f = oracle_scores(oracle, kind, hx[ix], hy[iy], pts, w, abws[d])
                what = f"set {d} h=({hx[ix]},{hy[iy]})"
                check_scores(lcv[d, ix, iy], lscv[d, ix, iy], want_lcv, want_lscv, bound, w, what)
                M = masses(w, abws[d])
                assert_parity(cv.loo(d, ix, iy), f, norm_c(kind, hx[ix], hy[iy]) * M.sum(), what=what + " loo")


def test_uneven_datasets_full_log_grid(bk, oracle):
    """Four uneven datasets x the 8 x 8 log grid h_k = 0.01 * 50^(k/7)."""
    sizes = [500, 1300, 2900, 6100]
    sets = [mixture(n, 1004 * 10 + k) for k, n in enumerate(sizes)]
    h = 0.01 * 50.0 ** (np.arange(8) / 7.0)
    with bk.CvGrid(GAUSSPROD, [p for p, _ in sets], [w for _, w in sets]) as cv:
        lcv, lscv = cv.run(h, h)
        assert lcv.shape == (4, 8, 8)
        for d, (pts, w) in enumerate(sets):
            for ix, iy in [(0, 0), (0, 7), (3, 4), (7, 0), (7, 7), (5, 2)]:
                want_lcv, want_lscv, bound, _ = oracle_scores(oracle, GAUSSPROD, h[ix], h[iy], pts, w, None)
                check_scores(lcv[d, ix, iy], lscv[d, ix, iy], want_lcv, want_lscv, bound, w,
                             f"set {d} ({ix},{iy})")


@pytest.mark.parametrize("kind", [GAUSSPROD, EPANPROD])
def test_variants_and_graph_are_bit_identical(bk, kind):
    sets = [mixture(3000, 601), mixture(900, 602)]
    hx = np.geomspace(0.03, 0.4, 6)
    hy = np.geomspace(0.02, 0.5, 9)
    results = []
    with bk.CvGrid(kind, [p for p, _ in sets], [w for _, w in sets]) as cv:
        for variant in range(bk.BK_CVG_VARIANTS):
            for graph in (0, 1):
                cv.set_option(bk.BK_CVG_VARIANT, variant)
                cv.set_option(bk.BK_CVG_GRAPH, graph)
                results.append(cv.run(hx, hy))
                assert cv.get_option(bk.BK_CVG_VARIANT) == variant
        cv.set_option(bk.BK_CVG_VARIANT, -1)
        cv.set_option(bk.BK_CVG_GRAPH, 1)
        results.append(cv.run(hx, hy))
        assert 0 <= cv.get_option(bk.BK_CVG_VARIANT) < bk.BK_CVG_VARIANTS
    for lcv, lscv in results[1:]:
        np.testing.assert_array_equal(lcv, results[0][0])
        np.testing.assert_array_equal(lscv, results[0][1])


def test_graph_replays_track_new_bandwidths(bk, oracle):
    """Replays of one captured graph with new bandwidth values, eager runs in between."""
    pts, w = mixture(2500, 611)
    grids = [np.array([0.05, 0.1, 0.2, 0.4]), np.array([0.07, 0.13, 0.19, 0.31]),
             np.array([0.05, 0.1, 0.2, 0.4])]
    with bk.CvGrid(GAUSSPROD, [pts], [w]) as cv, bk.CvGrid(GAUSSPROD, [pts], [w]) as eager:
        eager.set_option(bk.BK_CVG_GRAPH, 0)
        out = []
        for h in grids:
            g = cv.run(h, h[::-1])
            e = eager.run(h, h[::-1])
            np.testing.assert_array_equal(g[0], e[0])
            np.testing.assert_array_equal(g[1], e[1])
            out.append(g)
        np.testing.assert_array_equal(out[0][0], out[2][0])
        want_lcv, want_lscv, bound, _ = oracle_scores(oracle, GAUSSPROD, 0.13, 0.19, pts, w, None)
        check_scores(out[1][0][0, 1, 1], out[1][1][0, 1, 1], want_lcv, want_lscv, bound, w, "replay")


def test_reference_fixture_scores(bk):
    """LCV/LSCV of the reference's own implementation (tests/golden/cv_ref.npz)."""
    z = np.load(__import__("pathlib").Path(__file__).parent / "golden" / "cv_ref.npz")
    pts, w = z["pts"], z["w"] / z["w"].sum()
    for i, (kind, dim, hx, hy) in enumerate(z["cases"]):
        if int(kind) not in (GAUSSPROD, EPANPROD):
            continue
        with bk.CvGrid(int(kind), [pts], [w]) as cv:
            lcv, lscv = cv.run([hx], [hy])
        # the fixture is the reference on FP64 points; FP32 rounding of the
        # points moves the scores by ~1e-7 relative
        assert lcv[0, 0, 0] == pytest.approx(float(z[f"lcv_{i}"]), rel=1e-5, abs=1e-6)
        assert lscv[0, 0, 0] == pytest.approx(float(z[f"lscv_{i}"]), rel=1e-5, abs=1e-7)


def test_underflowed_sums_are_rescued(bk, oracle):
    rng = np.random.default_rng(12)
    pts = rng.uniform(-50, 50, (25, 2))
    w = np.full(25, 1.0 / 25)
    hx = np.array([1.5, 3.0])  # nearest neighbours 7-25 bandwidths away
    with bk.CvGrid(GAUSSPROD, [pts], [w]) as cv:
        lcv, lscv = cv.run(hx, hx)
        _, _, rescued = cv.last_timing()
    assert rescued > 0
    for ix in range(2):
        for iy in range(2):
            want_lcv, want_lscv, bound, f = oracle_scores(oracle, GAUSSPROD, hx[ix], hx[iy], pts, w, None)
            assert np.isfinite(want_lcv)
            check_scores(lcv[0, ix, iy], lscv[0, ix, iy], want_lcv, want_lscv, bound, w, f"({ix},{iy})")


@pytest.mark.parametrize("n", [20_000, 33_333, 200_000])
def test_agrees_with_single_bandwidth_path_at_scale(bk, n):
    """One dataset, split over many work items: the grid's scores equal
    bk_likelihood_cv / bk_lsq_conv_cv (an independent kernel path)."""
    pts, w = mixture(n, 1004)
    h = np.array([0.01, 0.05, 0.2])
    with bk.CvGrid(GAUSSPROD, [pts], [w]) as cv:
        lcv, lscv = cv.run(h, h)
    with bk.DirectKde(pts, None, w) as kde:
        for ix, iy in [(0, 0), (1, 2), (2, 1)]:
            desc = bk.KernelDesc(GAUSSPROD, 2, h[ix], h[iy])
            assert lcv[0, ix, iy] == pytest.approx(kde.likelihood_cv(desc, w), abs=2e-6)
            want = kde.lsq_conv_cv(desc, w)
            assert lscv[0, ix, iy] == pytest.approx(want, rel=2e-6, abs=1e-9)


def test_errors(bk):
    pts, w = mixture(100, 3)
    with pytest.raises(bk.BkError):
        bk.CvGrid(2, [pts], [w])  # GaussianKernel: not a product kernel
    with pytest.raises(ValueError):
        bk.CvGrid(GAUSSPROD, [pts, pts[:0]], [w, w[:0]])  # empty dataset
    with bk.CvGrid(GAUSSPROD, [pts], [w]) as cv:
        with pytest.raises(bk.BkError):
            cv.loo(0, 0, 0)  # no run yet
        with pytest.raises(ValueError):
            cv.run([0.1], [0.0])
        with pytest.raises(ValueError):
            cv.run(np.full(65, 0.1), [0.1])
        with pytest.raises(ValueError):
            cv.set_option(bk.BK_CVG_VARIANT, bk.BK_CVG_VARIANTS)
        cv.run([0.1, 0.2], [0.3])
        with pytest.raises(IndexError):
            cv.loo(0, 2, 0)
        with pytest.raises(IndexError):
            cv.loo(1, 0, 0)


def test_unit_count(bk):
    """BK_CVG_COUNT_UNITS: the Epanechnikov kernel skips only stages whose
    terms are all zero, so its count lies between the units inside the
    convolution kernel's support (|dx| < 2 hx and |dy| < 2 hy, which it must
    evaluate) and all units.  The Gaussian kernel skips only stages whose
    products are all below 2^-152 (each would leave its FP32 sum unchanged):
    with small-bandwidth sub-grids (the 4 x 4 tiling) it skips some, and its
    scores stay bit-identical to every other tiling's.  Counting does not
    change the scores."""
    sets = [mixture(3000, 621), mixture(900, 622)]
    hx = np.geomspace(0.03, 0.4, 6)
    hy = np.geomspace(0.02, 0.5, 9)
    total = sum(len(p) ** 2 for p, _ in sets) * len(hx) * len(hy)
    inside = 0
    for p, _ in sets:
        P = f32(p)
        dx = np.abs(P[:, None, 0] - P[None, :, 0])
        dy = np.abs(P[:, None, 1] - P[None, :, 1])
        for a in hx:
            mx = dx < 2.0 * a * (1 - 1e-9)
            for b in hy:
                inside += int(np.count_nonzero(mx & (dy < 2.0 * b * (1 - 1e-9))))
    for kind in (GAUSSPROD, EPANPROD):
        with bk.CvGrid(kind, [p for p, _ in sets], [w for _, w in sets]) as cv:
            base = cv.run(hx, hy)
            with pytest.raises(bk.BkError):
                cv.unit_count()
            cv.set_option(bk.BK_CVG_COUNT_UNITS, 1)
            counted = cv.run(hx, hy)
            units = cv.unit_count()
            again = cv.run(hx, hy)
            assert cv.unit_count() == units
            cv.set_option(bk.BK_CVG_COUNT_UNITS, 0)
            after = cv.run(hx, hy)
        for r in (counted, again, after):
            np.testing.assert_array_equal(r[0], base[0])
            np.testing.assert_array_equal(r[1], base[1])
        print(f"[cvgrid] kind {kind}: {units} of {total} units evaluated ({units / total:.1%}), "
              f"{inside} inside the support")
        if kind == GAUSSPROD:
            assert 0 < units <= total
        else:
            assert inside <= units < total


def test_gaussian_stages_skipped_without_changing_a_bit(bk):
    """The Gaussian kernel's 4 x 4 tiling skips stages whose products are all
    below 2^-152 (small bandwidths, points far apart in Z order); its LOO
    sums and scores must equal bit for bit those of the 8 x 8 tiling, which
    tests no stage, and of the 4 x 8 one."""
    rng = np.random.default_rng(651)
    pts = rng.uniform(-4.0, 4.0, (40_000, 2))
    w = np.full(len(pts), 1.0 / len(pts))
    hx = np.geomspace(0.004, 0.03, 4)
    hy = np.geomspace(0.005, 0.025, 4)
    total = len(pts) ** 2 * len(hx) * len(hy)
    with bk.CvGrid(GAUSSPROD, [pts], [w]) as cv:
        cv.set_option(bk.BK_CVG_VARIANT, 0)
        base = cv.run(hx, hy)
        loo0 = cv.loo(0, 1, 2)
        cv.set_option(bk.BK_CVG_VARIANT, 1)
        mid = cv.run(hx, hy)
        cv.set_option(bk.BK_CVG_VARIANT, 2)
        cv.set_option(bk.BK_CVG_COUNT_UNITS, 1)
        small = cv.run(hx, hy)
        units = cv.unit_count()
        loo2 = cv.loo(0, 1, 2)
    for r in (mid, small):
        np.testing.assert_array_equal(r[0], base[0])
        np.testing.assert_array_equal(r[1], base[1])
    np.testing.assert_array_equal(loo2, loo0)
    print(f"[cvgrid] Gaussian 4x4 sub-grids: {units} of {total} units evaluated ({units / total:.1%})")
    assert 0 < units < total // 4


@pytest.mark.parametrize("kind", [GAUSSPROD, EPANPROD])
def test_autotuner_cache(bk, tmp_path, kind):
    """Without a cache entry a run uses the default schedule; bk_cv_grid_tune
    times every schedule on whole runs and records the fastest, which later
    runs and plans of the same shape take from the cache in memory or from
    the saved file; a file of another version is ignored and rewritten.
    Every schedule gives the same bits."""
    sets = [mixture(2500, 631), mixture(700, 632)]
    hx = np.geomspace(0.03, 0.4, 5)
    hy = np.geomspace(0.02, 0.5, 7)
    path = tmp_path / "tune.txt"
    path.write_text("bbrcit-hip-tune 1 0\nstale\t0 0\n")
    try:
        bk.tune_set_cache(path)
        assert bk.tune_entry_count() == 0
        runs = []
        with bk.CvGrid(kind, [p for p, _ in sets], [w for _, w in sets]) as cv:
            runs.append(cv.run(hx, hy))
            assert cv.get_option(bk.BK_CVG_TUNE_SOURCE) == 1
            timings = cv.tune(hx, hy)
            assert len(timings) == bk.BK_CVG_VARIANTS * (3 if kind == GAUSSPROD else 2)
            assert all(ms > 0 for _, _, ms in timings)
            best = min(timings, key=lambda t: t[2])
            runs.append(cv.run(hx, hy))
            assert cv.get_option(bk.BK_CVG_TUNE_SOURCE) == 2
            assert cv.get_option(bk.BK_CVG_VARIANT) == best[0]
            sched = (cv.get_option(bk.BK_CVG_VARIANT), cv.get_option(bk.BK_CVG_UNROLL))
            cv.set_option(bk.BK_CVG_VARIANT, 0)
            with pytest.raises(bk.BkError):
                cv.tune(hx, hy)
            cv.set_option(bk.BK_CVG_VARIANT, -1)
        lines = path.read_text().splitlines()
        assert lines[0] == "bbrcit-hip-tune 1 1" and len(lines) == 2
        key, vals = lines[1].split("\t")
        assert key.startswith("cvgrid/") and f"/kind{kind}/" in key and key.endswith("/grid5x7/sets2500,700")
        assert [int(v) for v in vals.split()] == [best[0], best[1]]
        print(f"[tune] kind {kind}: {key} -> {vals}; timings {timings}")
        # a new plan of the same shape, the cache reloaded from the file only
        bk.tune_set_cache(path)
        assert bk.tune_entry_count() == 1
        with bk.CvGrid(kind, [p for p, _ in sets], [w for _, w in sets]) as cv:
            runs.append(cv.run(hx, hy))
            assert cv.get_option(bk.BK_CVG_TUNE_SOURCE) == 2
            assert (cv.get_option(bk.BK_CVG_VARIANT), cv.get_option(bk.BK_CVG_UNROLL)) == sched
            # forced tilings with their default unroll: same bits
            for v in range(bk.BK_CVG_VARIANTS):
                cv.set_option(bk.BK_CVG_VARIANT, v)
                runs.append(cv.run(hx, hy))
                assert cv.get_option(bk.BK_CVG_TUNE_SOURCE) == 0
            with pytest.raises(ValueError):
                cv.set_option(bk.BK_CVG_TUNE_SOURCE, 1)
        for lcv, lscv in runs[1:]:
            np.testing.assert_array_equal(lcv, runs[0][0])
            np.testing.assert_array_equal(lscv, runs[0][1])
    finally:
        bk.tune_set_cache(None)
