"""Why the d = 16 Gram form keeps FP64 matrix products (CPU analysis).

The Gaussian term's base-2 exponent in Gram form is
    e = -g (|q|^2 + |r|^2) + 2 g q.r,   g = log2(e) / (2 h^2),
and at the d = 16 benchmark workload (200k standard-normal references in kd
order, standard-normal queries, h = 0.5) the dot product carries
|2 g q.r| up to ~100 exponent units, so its rounding moves the term by
ln 2 * 2 g * (error of q.r).  An FP32 matrix product (FP32 products, FP32
accumulation over k) is emulated here with sequential FP32 accumulation, in
two forms: uncentred, and centred on each 256-reference tile's box centre c
(q - c and r - c rounded to FP32), the norms taken exactly in FP64 in both
(a best case).  256 sampled queries' sums are compared with FP64 sums: both
forms reach ~1.1e-5 - 1.4e-5 relative, above the 1e-5 bar, because kd tiles
of 256 points in 16 dimensions are not compact (a kd split halves ~10 of the
16 axes once), so centring shrinks |r - c| little.
"""
from __future__ import annotations

import numpy as np

from bbrcit_kde_amd import datasets as ds

F32 = np.float32


def fp32_dots(A, B):
    """A (m, D) . B (n, D) per pair with FP32 products and sequential FP32
    accumulation over the D axis: (n, m)."""
    acc = np.zeros((B.shape[0], A.shape[0]), F32)
    for k in range(A.shape[1]):
        acc = (acc + B[:, k:k + 1].astype(F32) * A[None, :, k].astype(F32)).astype(F32)
    return acc.astype(np.float64)


def test_fp32_gram_form_misses_the_parity_bar_at_d16():
    n, dim, h, tile = 200_000, 16, 0.5, 256
    R = ds.round_f32(ds.standard_normal(n, dim, 1004))
    Q = ds.round_f32(ds.standard_normal(n, dim, 1005))
    R = R[ds.kd_order(R, tile)]
    Qs = Q[np.random.default_rng(1).choice(n, 256, replace=False)]
    g = 0.5 * np.log2(np.e) / (h * h)
    tid = np.arange(n) // tile
    starts = np.arange(0, n, tile)
    C = 0.5 * (np.minimum.reduceat(R, starts, axis=0) + np.maximum.reduceat(R, starts, axis=0))
    Rc = (R - C[tid]).astype(F32)
    rn = np.sum(R.astype(np.float64) ** 2, 1)
    rcn = np.sum(Rc.astype(np.float64) ** 2, 1)
    err_u, err_c = [], []
    for b in range(0, len(Qs), 16):
        qb = Qs[b:b + 16]
        # FP64 sums from the exact differences, per query
        exact = np.array([np.sum(np.exp2(-g * np.sum((R - q) ** 2, 1))) for q in qb])
        uncentred = -g * (rn[None, :] + np.sum(qb.astype(np.float64) ** 2, 1)[:, None]) + 2 * g * fp32_dots(R, qb)
        err_u += list(np.abs(np.exp2(uncentred).sum(1) - exact) / exact)
        cen = np.empty_like(uncentred)
        for i, q in enumerate(qb):
            Qc = (q[None, :] - C[tid]).astype(F32)
            acc = np.zeros(n, F32)
            for k in range(dim):
                acc = (acc + Rc[:, k] * Qc[:, k]).astype(F32)
            cen[i] = -g * (rcn + np.sum(Qc.astype(np.float64) ** 2, 1)) + 2 * g * acc.astype(np.float64)
        err_c += list(np.abs(np.exp2(cen).sum(1) - exact) / exact)
    err_u, err_c = np.array(err_u), np.array(err_c)
    print(f"[gram16 fp32] uncentred max {err_u.max():.3e} median {np.median(err_u):.3e}; "
          f"tile-centred max {err_c.max():.3e} median {np.median(err_c):.3e}")
    assert err_u.max() > 1e-5 and err_c.max() > 1e-5
