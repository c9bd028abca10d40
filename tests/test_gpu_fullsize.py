"""Benchmark workloads at their full sizes, pinned to the FP64 oracle on
sampled points.

The smaller parity tests cover every kernel and edge case; these run the
exact inputs the benchmark times and check them where they are timed:
  * the adaptive pass on 1M mixture points (pilot GaussianProductKernel2d
    0.05 / 0.05 with the self term, lambda_i = (f_i / g)^-0.5, g the FP64
    geometric mean), the adapted density at 1M uniform queries and the
    leave-one-out sums of the adapted density (reference
    include/KernelDensityImpl.h:254-359, 215-252);
  * the batched CV grid over the four datasets of 50k, 200k, 500k and 1M
    points: leave-one-out densities of 512 points per dataset at three
    bandwidth pairs of the 8 x 8 log grid (GaussianProductKernel2d) and one
    (EpanechnikovProductKernel2d), grid_search.cc:366-402;
  * the d = 16 Gaussian, 200k standard-normal references x 256 of the 200k
    queries.
Tolerance: 1e-5 relative where the exact value is at least 1e-10 of C * sum m
(an upper bound of every value), 1e-5 of that floor below it; derived
quantities (g, lambda from the returned pilot) to 1e-12.
"""
from __future__ import annotations

import numpy as np
import pytest

from bbrcit_kde_amd import datasets as ds

pytestmark = pytest.mark.gpu

GPROD, EPROD, GAUSS = 4, 3, 2
RTOL = 1e-5
FLOOR_REL = 1e-10


def _check(got, want, bound, what):
    floor = FLOOR_REL * bound
    above = np.abs(want) >= floor
    if above.any():
        rel = np.abs(got[above] - want[above]) / np.abs(want[above])
        assert rel.max() <= RTOL, f"{what}: max rel err {rel.max():.3e}"
    if (~above).any():
        assert np.abs(got[~above] - want[~above]).max() <= RTOL * floor, f"{what}: below floor"
    return float(np.mean(~above))


def _sample(n, k, seed):
    return np.sort(np.random.default_rng(seed).choice(n, size=k, replace=False))


def test_adaptive_pass_on_1m_points(bk, oracle):
    pts, w = ds.gaussian_mixture(1_000_000, seed=ds.SEED_MIXTURE_1M)
    perm = ds.kd_order(pts, 64)
    pts, w = pts[perm], w[perm]
    mass = w / w.sum()
    q = ds.uniform_box(1_000_000, -4.0, 4.0, ds.SEED_MIXTURE_1M + 1)
    h = 0.05
    desc = bk.KernelDesc(GPROD, 2, h, h)
    C = 1.0 / (2.0 * np.pi * h * h)
    P, M = ds.round_f32(pts), ds.round_f32(mass)
    idx = _sample(len(pts), 512, 1)
    with bk.DirectKde(pts, None, mass) as kde:
        print("[fullsize] adaptive pilot and pass", flush=True)
        pilot = kde.self_eval(desc, exclude_self=False, normalize=True)
        g, lam = kde.adapt(desc, 0.5, mass)
        loo = kde.self_eval(desc, exclude_self=True, normalize=True)
        lcv = kde.likelihood_cv(desc, mass)
    # the pilot at sampled points against the oracle (self term included)
    want = oracle.direct(GPROD, h, h, P, P[idx], mass=M, normalize=True)
    _check(pilot[idx], want, C * M.sum(), "pilot")
    # g and lambda follow from the pilot exactly as the reference defines them
    g_host = np.exp(np.sum(mass * np.log(pilot)))
    assert abs(g - g_host) <= 1e-12 * g_host
    np.testing.assert_allclose(lam, (pilot / g_host) ** -0.5, rtol=1e-12)
    np.testing.assert_allclose(lam[idx], (want / g_host) ** -0.5, rtol=RTOL)
    # the adapted density's leave-one-out sums (the context holds FP32 masses and abw)
    a_mass = mass * lam ** -2.0
    A32, AM32 = ds.round_f32(lam), ds.round_f32(a_mass)
    want_loo = oracle.direct_loo_at(GPROD, h, h, P, idx, mass=AM32, abw=A32, normalize=True)
    _check(loo[idx], want_loo, C * AM32.sum(), "adapted leave-one-out")
    assert abs(lcv - np.sum(mass * np.log(loo))) <= 1e-10 * abs(lcv)
    # the adapted density at 1M uniform queries
    print("[fullsize] adaptive evaluation", flush=True)
    with bk.DirectKde(pts, q, a_mass, lam) as ev:
        got = ev.eval(desc, 0, len(pts) - 1, 0, len(q) - 1)
    jdx = _sample(len(q), 512, 2)
    want_q = oracle.direct(GPROD, h, h, P, ds.round_f32(q[jdx]), mass=AM32, abw=A32, normalize=True)
    frac = _check(got[jdx], want_q, C * AM32.sum(), "adaptive evaluation")
    print(f"[fullsize] adaptive: g {g:.6g}, lambda in [{lam.min():.3g}, {lam.max():.3g}], "
          f"{frac:.1%} of sampled queries below the floor", flush=True)


CV_SIZES = (50_000, 200_000, 500_000, 1_000_000)
CV_SEED = 10040
CV_GRID = 0.01 * 50.0 ** (np.arange(8) / 7.0)


@pytest.mark.parametrize("kind,cases", [(GPROD, ((0, 0), (3, 5), (7, 2))), (EPROD, ((4, 4),))])
def test_cv_grid_on_the_four_benchmark_datasets(bk, oracle, kind, cases):
    sets = []
    for k, n in enumerate(CV_SIZES):
        p, w = ds.gaussian_mixture(n, seed=CV_SEED + k)
        sets.append((p, w / w.sum()))
    with bk.CvGrid(kind, [p for p, _ in sets], [w for _, w in sets]) as cv:
        print(f"[fullsize] cv grid kind {kind}", flush=True)
        lcv, lscv = cv.run(CV_GRID, CV_GRID)
        loos = {(d, ix, iy): cv.loo(d, ix, iy) for d in range(len(sets)) for ix, iy in cases}
    for d, (p, w) in enumerate(sets):
        P, W = ds.round_f32(p), ds.round_f32(w)
        idx = _sample(len(p), 512, 10 + d)
        for ix, iy in cases:
            hx, hy = CV_GRID[ix], CV_GRID[iy]
            C = oracle.normalization(kind, 2, hx, hy)
            want = oracle.direct_loo_at(kind, hx, hy, P, idx, mass=W, normalize=True)
            _check(loos[(d, ix, iy)][idx], want, C * W.sum(), f"set {d} kind {kind} ({ix}, {iy})")
            if kind == GPROD:
                # the LCV score is the weighted log of these sums, reduced in FP64
                with np.errstate(divide="ignore"):
                    host = np.sum(w * np.log(loos[(d, ix, iy)]))
                if np.isfinite(host):
                    assert abs(lcv[d, ix, iy] - host) <= 1e-9 * max(1.0, abs(host))
                else:
                    assert lcv[d, ix, iy] == host


def test_gauss16_200k(bk, oracle):
    n, dim, h = 200_000, 16, 0.5
    R = ds.standard_normal(n, dim, 1004)
    Q = ds.standard_normal(n, dim, 1005)
    m = np.full(n, 1.0 / n)
    jdx = _sample(n, 256, 3)
    with bk.DirectKde(R, Q, m) as kde:
        print("[fullsize] d=16 Gram form", flush=True)
        got = kde.eval(bk.KernelDesc(GAUSS, dim, h), 0, n - 1, 0, n - 1)
    want = oracle.direct(GAUSS, h, h, ds.round_f32(R), ds.round_f32(Q[jdx]), mass=ds.round_f32(m), normalize=True)
    C = oracle.normalization(GAUSS, dim, h, h)
    _check(got[jdx], want, C * ds.round_f32(m).sum(), "d=16")
