"""Why the surrogate cannot meet 1e-3 in the data region (CPU analysis).

The surrogate's workload is the weighted GaussianKernel density of 1M points
of the seeded 3-component mixture at h = 0.05 (masses w / sum w, w uniform on
(0.5, 1.5) and independent of the points).  Its expectation is the mixture
convolved with the kernel, in closed form
    E f(q) = sum_k (1/3) N(q; mu_k, Sigma_k + h^2 I),
the smoothest model of the density there is.  The exact sums (the FP64
oracle) at 3,000 queries of the data region differ from it by the sampling
noise of 1M points: a median of ~1% relative where E f > 0.1 and ~3% where
E f > 1e-3, with only a few percent of the queries within 1e-3.  The noise
varies over a length ~h, so a network would have to learn ~10^4 independent
bumps over the data region to reach 1e-3 there; the served path therefore
uses the network only where it meets 1e-3 (far from the data) and the exact
sum elsewhere.
"""
from __future__ import annotations

import numpy as np

from bbrcit_kde_amd import datasets as ds

GAUSS = 2


def mixture_expectation(q, means, covs, h):
    out = np.zeros(len(q))
    for mu, cov in zip(means, covs):
        c = cov + h * h * np.eye(2)
        inv = np.linalg.inv(c)
        d = q - mu
        e = np.einsum("ij,jk,ik->i", d, inv, d)
        out += np.exp(-0.5 * e) / (2.0 * np.pi * np.sqrt(np.linalg.det(c))) / len(means)
    return out


def test_kde_noise_exceeds_the_surrogate_tolerance_in_the_data_region(oracle):
    n, h = 1_000_000, 0.05
    pts, w = ds.gaussian_mixture(n, seed=ds.SEED_MIXTURE_1M)
    means, covs = ds.gaussian_mixture_components(ds.SEED_MIXTURE_1M)
    mass = w / w.sum()
    # 3,000 queries of the data region (E f > 1e-3), drawn uniformly on [-4, 4]^2
    cand = ds.uniform_box(40_000, -4.0, 4.0, 77)
    ef_c = mixture_expectation(cand, means, covs, h)
    q = cand[ef_c > 1e-3][:3000]
    assert len(q) == 3000
    ef = mixture_expectation(q, means, covs, h)
    f = oracle.direct(GAUSS, h, h, pts, q, mass=mass, normalize=True)
    rel = np.abs(f - ef) / ef
    dense = ef > 0.1
    med_dense, med_all = float(np.median(rel[dense])), float(np.median(rel))
    within = float(np.mean(rel <= 1e-3))
    print(f"[surrogate noise] E f > 0.1: median |f - E f| / E f = {med_dense:.3%} ({dense.sum()} queries); "
          f"E f > 1e-3: median {med_all:.3%}; within 1e-3: {within:.2%}")
    # the closed form is the density's expectation: no bias beyond the noise
    assert abs(np.median((f - ef) / ef)) < 0.5 * med_all
    assert med_dense > 3e-3 and med_all > 1e-2
    assert within < 0.1
