"""Parity in the tails of the density, where sums are many orders of
magnitude below an upper bound of the sum.

  * FP64 contexts: a point 15 bandwidths from every other one, evaluated as
    a query in the given and the spatial query order and as a leave-one-out
    point of a kd-ordered set, held to 1e-12 relative with no floor.  FP64
    Gaussian culling drops at most 2^-50 of each workgroup's sums (its
    threshold follows a lower bound of those sums), so such sums are the
    oracle's to FP64 rounding and the likelihood cross-validation score stays
    finite.
  * FP32 Gram tiles (per-reference scales, tile radius rho^2 in (3, 4] in
    base-2 exponent units): queries whose sums sit just above the parity
    floor (1e-10 of C * sum m) are held to 1e-5 relative.  The worst
    relative error is printed.
  * FP64 Gaussian, compact tiles in their centred form and wide ones from
    the differences: a mixture 5000 units from the origin, queries near the
    data and beyond it, held to 1e-12 relative down to the FP64 floor.
  * D = 3 Gaussian (FP64 distances of the FP32 points): queries 10 to 16
    from the centre of a standard normal sample, sums far below the FP32
    range (down to ~1e-150), held to 1e-5 relative with no floor (the
    integer part of the base-2 exponent is applied in FP64).
"""
from __future__ import annotations

import numpy as np
import pytest

from bbrcit_kde_amd import datasets as ds
from tests.test_gpu_direct import assert_parity, f32

pytestmark = pytest.mark.gpu

EPAN, GAUSS, EPANPROD, GAUSSPROD = 1, 2, 3, 4
LOG2E = 1.4426950408889634


def rel_err(got, want):
    got, want = np.asarray(got), np.asarray(want)
    assert np.all(want > 0), "every checked value must be positive"
    return np.abs(got - want) / want


def isolated_set(n, seed, hx):
    """kd mixture plus one point 15 hx beyond the largest x of the others."""
    pts, w = ds.gaussian_mixture(n, seed=seed)
    far = pts[np.argmax(pts[:, 0])] + np.array([15.0 * hx, 0.0])
    return np.vstack([pts, far]), np.append(w, w.mean()), far


def outliers(pts, hx, hy):
    """Queries 15 to 18 bandwidths beyond the extreme points in x and y."""
    out = []
    for k, (axis, sign) in enumerate([(0, 1), (0, -1), (1, 1), (1, -1)]):
        i = np.argmax(sign * pts[:, axis])
        step = np.zeros(2)
        step[axis] = sign * (15.0 + k) * (hx if axis == 0 else hy)
        out.append(pts[i] + step)
    return np.array(out)


@pytest.mark.parametrize("kind,h", [(GAUSS, (0.02, 0.02)), (GAUSSPROD, (0.02, 0.03))])
def test_f64_isolated_point_in_given_and_spatial_order(bk, oracle, kind, h):
    pts, w, _ = isolated_set(20_000, 1200 + kind, h[0])
    perm = ds.kd_order(pts, 64)
    pts, mass = pts[perm], (w / w.sum())[perm]
    far_q = outliers(pts, *h)
    q = np.vstack([ds.uniform_box(3_000, -4.0, 4.0, 1250 + kind), far_q])
    desc = bk.KernelDesc(kind, 2, *h)
    with bk.DirectKde(pts, q, mass, precision=bk.BK_PREC_F64) as kde:
        kde.set_option(bk.BK_OPT_QUERY_ORDER, bk.BK_QUERY_ORDER_GIVEN)
        given = kde.eval(desc, 0, len(pts) - 1, 0, len(q) - 1)
        kde.set_option(bk.BK_OPT_QUERY_ORDER, bk.BK_QUERY_ORDER_SPATIAL)
        spatial = kde.eval(desc, 0, len(pts) - 1, 0, len(q) - 1)
    want = oracle.direct(kind, h[0], h[1], pts, q, mass=mass, normalize=True)
    bound = oracle.normalization(kind, 2, *h) * mass.sum()
    assert_parity(given, want, bound, rtol=1e-12, what=f"given kind {kind}", kind=kind)
    assert_parity(spatial, want, bound, rtol=1e-12, what=f"spatial kind {kind}", kind=kind)
    tail = slice(len(q) - len(far_q), len(q))
    print(f"[tails] kind {kind}: outlier sums {want[tail]} (bound {bound:.3g})")
    assert np.all(want[tail] < 1e-40 * bound)
    assert rel_err(given[tail], want[tail]).max() <= 1e-12
    assert rel_err(spatial[tail], want[tail]).max() <= 1e-12


@pytest.mark.parametrize("kind,h", [(GAUSS, (0.02, 0.02)), (GAUSSPROD, (0.02, 0.03))])
def test_f64_isolated_point_leave_one_out(bk, oracle, kind, h):
    pts, w, far = isolated_set(20_000, 1300 + kind, h[0])
    perm = ds.kd_order(pts, 64)
    pts, w = pts[perm], w[perm]
    mass = w / w.sum()
    i_far = int(np.flatnonzero(np.all(pts == far, axis=1))[0])
    desc = bk.KernelDesc(kind, 2, *h)
    with bk.DirectKde(pts, None, mass, precision=bk.BK_PREC_F64) as kde:
        loo = kde.self_eval(desc, exclude_self=True)
        lcv = kde.likelihood_cv(desc, mass)
    want = oracle.direct_loo_at(kind, h[0], h[1], pts, np.array([i_far]), mass=mass, normalize=True)
    print(f"[tails] kind {kind}: isolated leave-one-out sum {want[0]:.6e}, got {loo[i_far]:.6e}")
    assert 0 < want[0] < 1e-40
    assert rel_err(loo[[i_far]], want).max() <= 1e-12
    idx = np.append(np.random.default_rng(3).choice(len(pts), 2_000, replace=False), i_far)
    want_s = oracle.direct_loo_at(kind, h[0], h[1], pts, idx, mass=mass, normalize=True)
    assert_parity(loo[idx], want_s, oracle.normalization(kind, 2, *h) * mass.sum(), rtol=1e-12,
                  what=f"loo kind {kind}", kind=kind)
    # the score is the weighted log of these sums, finite with the isolated point
    host = np.sum(mass * np.log(loo))
    assert np.isfinite(lcv) and abs(lcv - host) <= 1e-12 * abs(host)


def gram_tiles(n_tiles, tile, seed, h, rho2_lo=3.2, rho2_hi=3.95):
    """Tiles of `tile` references whose radius rho^2 = max W gx |r - c|^2
    (base-2 exponent units, W = 1/a^2, c the box centre) is a target drawn
    from (rho2_lo, rho2_hi), with scales a varying inside every tile; tile
    centres 2 apart, so the reach of one tile covers no other."""
    rng = np.random.default_rng(seed)
    gx = 0.5 * LOG2E / (h * h)
    side = int(np.ceil(np.sqrt(n_tiles)))
    pts, abw = [], []
    for t in range(n_tiles):
        c = np.array([t % side, t // side], float) * 2.0 - side
        a = rng.uniform(0.8, 1.25, tile)
        a[:4] = a.max()
        target = rng.uniform(rho2_lo, rho2_hi)
        # every point within radius a sqrt(target / gx) of c (W gx |r - c|^2
        # <= target); the first four, with the largest scale, at that radius
        # on the axes, so the box is centred at c and rho^2 = target
        r = a * np.sqrt(target / gx)
        ang = rng.uniform(0, 2 * np.pi, tile)
        rad = r * np.sqrt(rng.uniform(0, 1, tile))
        off = np.stack([rad * np.cos(ang), rad * np.sin(ang)], 1)
        R = r[0]
        off[:4] = [[R, 0.0], [-R, 0.0], [0.0, R], [0.0, -R]]
        pts.append(c + off)
        abw.append(a)
    return np.vstack(pts), np.concatenate(abw)


def tile_rho2(pts, abw, tile, h):
    """rho^2 and the scale spread of each aligned tile, as the kernel's
    preparation computes them from the FP32 points."""
    P, A = f32(pts), f32(abw)
    gx = 0.5 * LOG2E / (h * h)
    out = []
    for t in range(len(P) // tile):
        p, a = P[t * tile:(t + 1) * tile], A[t * tile:(t + 1) * tile]
        lo = p.min(0).astype(np.float32)
        hi = p.max(0).astype(np.float32)
        c = (np.float32(0.5) * (lo + hi)).astype(np.float32).astype(float)
        W = (1.0 / (a.astype(float) ** 2)).astype(np.float32).astype(float)
        r2 = W * gx * ((p[:, 0] - c[0]) ** 2 + (p[:, 1] - c[1]) ** 2)
        out.append((r2.max(), W.min() != W.max()))
    return out


def test_gram_tiles_near_rho2_4_above_the_floor(bk, oracle):
    h, tile, n_tiles = 0.05, 256, 64
    pts, abw = gram_tiles(n_tiles, tile, 77, h)
    rho = tile_rho2(pts, abw, tile, h)
    assert all(v for _, v in rho), "every tile has varying scales"
    assert all(3.0 < r <= 4.0 for r, _ in rho), f"rho^2 out of (3, 4]: {[r for r, _ in rho]}"
    mass = np.full(len(pts), 1.0 / len(pts))
    # queries around every tile at distances whose sums reach from well above
    # the floor (1e-10 of the bound) to below it
    gx = 0.5 * LOG2E / (h * h)
    rng = np.random.default_rng(78)
    side = int(np.ceil(np.sqrt(n_tiles)))
    qs = []
    for t in range(n_tiles):
        c = np.array([t % side, t // side], float) * 2.0 - side
        e = rng.uniform(10.0, 80.0, 96)        # base-2 exponents from the tile centre
        d = np.sqrt(e / gx)
        ang = rng.uniform(0, 2 * np.pi, 96)
        qs.append(c + np.stack([d * np.cos(ang), d * np.sin(ang)], 1))
    q = np.vstack(qs)
    desc = bk.KernelDesc(GAUSS, 2, h)
    with bk.DirectKde(pts, q, mass, abw) as kde:
        got = kde.eval(desc, 0, len(pts) - 1, 0, len(q) - 1)
    want = oracle.direct(GAUSS, h, h, f32(pts), f32(q), mass=f32(mass), abw=f32(abw), normalize=True)
    bound = oracle.normalization(GAUSS, 2, h) * f32(mass).sum()
    assert_parity(got, want, bound, what="Gram tiles rho^2 in (3, 4]")
    near = (want >= 1e-10 * bound) & (want < 1e-8 * bound)
    assert near.sum() >= 500, f"only {near.sum()} queries just above the floor"
    worst = rel_err(got[near], want[near]).max()
    print(f"[tails] Gram rho^2 in (3, 4]: {near.sum()} queries with sums in [1e-10, 1e-8) of the bound, "
          f"worst relative error {worst:.3e}; all above the floor {rel_err(got[want >= 1e-10 * bound], want[want >= 1e-10 * bound]).max():.3e}")
    assert worst <= 1e-5


def test_d3_far_queries_without_floor(bk, oracle):
    n, dim, h = 5_000, 3, 0.5
    R = ds.standard_normal(n, dim, 1401)
    rng = np.random.default_rng(1402)
    u = rng.normal(size=(512, dim))
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    Q = u * rng.uniform(10.0, 16.0, (512, 1))
    m = np.full(n, 1.0 / n)
    with bk.DirectKde(R, Q, m) as kde:
        got = kde.eval(bk.KernelDesc(GAUSS, dim, h), 0, n - 1, 0, len(Q) - 1)
    want = oracle.direct(GAUSS, h, h, f32(R), f32(Q), mass=f32(m), normalize=True)
    print(f"[tails] d=3: sums from {want.min():.3e} to {want.max():.3e}")
    assert want.min() > 1e-290 and want.min() < 1e-100
    assert (want < 1e-45).mean() > 0.5  # below the smallest FP32 denormal
    assert rel_err(got, want).max() <= 1e-5


@pytest.mark.parametrize("h", [(0.03, 0.03), (0.02, 0.035)])
def test_f64_gaussian_centred_and_direct_tiles_far_from_origin(bk, oracle, h):
    # FP64 Gaussian tiles compact enough (W |r - c|^2_rho <= 100) are summed in
    # their centred form, the others from the differences; a mixture 5000
    # units from the origin has both (its sparse tails are wide), and the
    # centred parts cancel nothing of the offset.  Held to 1e-12 relative
    # down to the FP64 tail floor, queries near the data and beyond it.
    kind = GAUSS if h[0] == h[1] else GAUSSPROD
    pts, w = ds.gaussian_mixture(30_000, seed=1300)
    pts = pts[ds.kd_order(pts, 64)] + np.array([5000.0, -3000.0])
    mass = w / w.sum()
    rng = np.random.default_rng(1301)
    q = np.vstack([pts[rng.choice(len(pts), 1500, replace=False)] + rng.normal(0.0, 0.05, (1500, 2)),
                   ds.uniform_box(1500, -4.0, 4.0, 1302) + np.array([5000.0, -3000.0])])
    desc = bk.KernelDesc(kind, 2, *h)
    with bk.DirectKde(pts, q, mass, precision=bk.BK_PREC_F64) as kde:
        got = kde.unnormalized_eval(desc, 0, len(pts) - 1, 0, len(q) - 1)
    want = oracle.direct(kind, h[0], h[1], pts, q, mass=mass)
    assert_parity(got, want, mass.sum(), rtol=1e-12, what=f"f64 offset kind {kind}", kind=kind)
