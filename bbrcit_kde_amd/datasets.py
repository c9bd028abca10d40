"""Seeded synthetic point sets of the benchmark workloads.

All generators use numpy's PCG64 (``numpy.random.default_rng(seed)``), whose
output stream is fixed by the seed, so every run and every machine sees the
same FP64 points.  The FP32 path rounds them once on upload; checkers that
compare against it use :func:`round_f32` to see the same numbers.
"""
from __future__ import annotations

import numpy as np

# Seeds: 1000 + workload number.
SEED_BIMODAL_10K = 1001
SEED_MIXTURE_1M = 1002
SEED_ADAPTIVE_1M = 1003
SEED_CV_BASE = 1004
SEED_D16 = 1014
SEED_SURROGATE = 1005


def round_f32(a: np.ndarray) -> np.ndarray:
    """The FP32-rounded values, as double (what the FP32 path computes on)."""
    return np.asarray(a, dtype=np.float64).astype(np.float32).astype(np.float64)


def bimodal_gaussian(n: int, seed: int = SEED_BIMODAL_10K) -> np.ndarray:
    """Equal mixture of isotropic N((-1,-1), 0.5^2 I) and N((1,1), 0.5^2 I)."""
    rng = np.random.default_rng(seed)
    comp = rng.random(n) < 0.5
    z = rng.standard_normal((n, 2)) * 0.5
    mean = np.where(comp[:, None], -1.0, 1.0)
    return z + mean


def uniform_box(n: int, lo: float, hi: float, seed: int, dim: int = 2) -> np.ndarray:
    rng = np.random.default_rng(seed)
    return rng.uniform(lo, hi, size=(n, dim))


def _mixture_components(rng, n_components, mean_box, eig_range):
    means = rng.uniform(-mean_box, mean_box, size=(n_components, 2))
    lam = rng.uniform(eig_range[0], eig_range[1], size=(n_components, 2))
    theta = rng.uniform(0.0, np.pi, size=n_components)
    covs = []
    for k in range(n_components):
        c, s = np.cos(theta[k]), np.sin(theta[k])
        r = np.array([[c, -s], [s, c]])
        covs.append(r @ np.diag(lam[k]) @ r.T)
    return means, covs


def gaussian_mixture_components(seed: int, n_components: int = 3, mean_box: float = 2.0, eig_range=(0.05, 0.5)):
    """(means (k, 2), covariances [k x (2, 2)]) of gaussian_mixture(n, seed)'s
    components (equal proportions), for any n."""
    return _mixture_components(np.random.default_rng(seed), n_components, mean_box, eig_range)


def gaussian_mixture(n: int, seed: int, n_components: int = 3, mean_box: float = 2.0,
                     eig_range=(0.05, 0.5)):
    """Points of an equal-proportion 2-D Gaussian mixture and their weights.

    Component means are uniform on [-mean_box, mean_box]^2; each covariance
    is R(theta) diag(l1, l2) R(theta)^T with l1, l2 uniform on eig_range and
    theta uniform on [0, pi).  Weights are uniform on (0.5, 1.5).
    Returns (points (n, 2), weights (n,)).
    """
    rng = np.random.default_rng(seed)
    means, covs = _mixture_components(rng, n_components, mean_box, eig_range)
    chol = [np.linalg.cholesky(c) for c in covs]
    comp = rng.integers(0, n_components, size=n)
    z = rng.standard_normal((n, 2))
    pts = np.empty((n, 2))
    for k in range(n_components):
        sel = comp == k
        pts[sel] = z[sel] @ chol[k].T + means[k]
    w = rng.uniform(0.5, 1.5, size=n)
    return pts, w


def standard_normal(n: int, dim: int, seed: int) -> np.ndarray:
    rng = np.random.default_rng(seed)
    return rng.standard_normal((n, dim))


def kd_order(points: np.ndarray, leaf_nmax: int = 64) -> np.ndarray:
    """Permutation that puts points in kd-tree leaf order.

    Lower-median splits cycling through the dimensions, the partition rule
    of the reference's Kdtree (include/Kdtree.h:429-471).  Consecutive
    points of the result are spatially close, which is the order the
    reference's KernelDensity hands its points to the device.
    """
    pts = np.asarray(points)
    n, dim = pts.shape
    perm = np.arange(n)
    stack = [(0, n, 0)]
    while stack:
        i, j, d = stack.pop()
        if j - i <= leaf_nmax:
            continue
        m = i + (j - 1 - i) // 2
        seg = perm[i:j]
        part = np.argpartition(pts[seg, d], m - i, kind="introselect")
        perm[i:j] = seg[part]
        stack.append((i, m + 1, (d + 1) % dim))
        stack.append((m + 1, j, (d + 1) % dim))
    return perm
