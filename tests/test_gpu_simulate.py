"""Batched sampling on the GPU (KernelDensity::simulate(e, n, out) over
bk_sampler_draw) against the reference's simulate.

  * Bit-identical to the reference's KernelDensity::simulate called once per
    sample with the same std::mt19937_64 seed (tests/golden/simulate_ref.npz,
    written by the reference itself), for every samplable kernel, FP32 and
    FP64 kernels (include/KernelDensityImpl.h:99-138).
  * At 1M samples of a 200k-point density: bit-identical to the host's
    single-point calls, and the sample moments match the density's exact
    mean and covariance (sum_i w_i x_i; the mixture of the points' kernels,
    Gaussian variance (h a)^2 per axis) within their standard errors.
"""
from __future__ import annotations

import time

import numpy as np
import pytest

from bbrcit_kde_amd import datasets as ds
from tests.test_cpp_api import GOLDEN, api, api_simulate  # noqa: F401  (fixture)

pytestmark = pytest.mark.gpu


def test_batched_simulate_matches_reference_fixture(api):  # noqa: F811
    f = np.load(GOLDEN / "simulate_ref.npz")
    pts, w = f["pts"], f["w"]
    for i, (kind, dim, kf, hx, hy) in enumerate(f["cases"]):
        kind, dim, kf = int(kind), int(dim), int(kf)
        got = api_simulate(api, kind, dim, kf, hx, hy, pts[:, :dim], w, 100 + i, 2000, batched=True)
        np.testing.assert_array_equal(got, f[f"sim_{i}"], err_msg=f"case {i}")


def test_batched_simulate_at_1m_samples(api):  # noqa: F811
    pts, w = ds.gaussian_mixture(200_000, seed=91)
    n, h = 1_000_000, 0.05
    t0 = time.perf_counter()
    gpu = api_simulate(api, 4, 2, 0, h, h, pts, w, 5, n, batched=True, leaf_nmax=256)
    t1 = time.perf_counter()
    host = api_simulate(api, 4, 2, 0, h, h, pts, w, 5, n, batched=False, leaf_nmax=256)
    t2 = time.perf_counter()
    print(f"[simulate] 1M samples of a 200k-point density: batched {t1 - t0:.3f} s, host calls {t2 - t1:.3f} s "
          "(both include building the KernelDensity)")
    np.testing.assert_array_equal(gpu, host)
    # moments of the density: the weighted mean of the points, and their
    # weighted covariance plus h^2 I (GaussianProductKernel2d, abw 1)
    p = w / w.sum()
    mean = p @ pts
    cov = (pts - mean).T @ ((pts - mean) * p[:, None]) + h * h * np.eye(2)
    got_mean = gpu.mean(axis=0)
    se = np.sqrt(np.diag(cov) / n)
    assert np.all(np.abs(got_mean - mean) <= 5 * se)
    got_cov = np.cov(gpu.T)
    assert np.allclose(got_cov, cov, rtol=0.02, atol=0.01 * np.abs(cov).max())
