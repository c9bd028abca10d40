"""The C++ drop-in API (bbrcit_kde_amd/include) through tests/cpp/api_driver.cc.

Host-only checks (CPU): the Kdtree orders and merges points exactly as the
reference's Kdtree (include/Kdtree.h:364-471), and the kernel classes give the
reference's normalisations and kernel values.  GPU checks: KernelDensity's
direct, dual-tree, cross-validation and adaptive results against the golden
fixtures written by the reference itself (tests/golden/), with FP64 kernels
to 1e-9 relative and float kernels to 1e-5.
"""
from __future__ import annotations

import ctypes
from pathlib import Path

import numpy as np
import pytest

from bbrcit_kde_amd import datasets as ds

GOLDEN = Path(__file__).resolve().parent / "golden"
DRIVER = Path(__file__).resolve().parent / "cpp" / "libapi_driver.so"

_D = ctypes.c_double
_DP = ctypes.POINTER(ctypes.c_double)
_SZ = ctypes.c_size_t
_I = ctypes.c_int

EPAN, GAUSS, EPANPROD, GAUSSPROD, EPANCONV1, EPANPRODCONV, GAUSSCONV1, GAUSSPRODCONV = range(1, 9)


def _p(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.float64).ctypes.data_as(_DP)


@pytest.fixture(scope="module")
def api():
    if not DRIVER.exists():
        pytest.skip("tests/cpp/libapi_driver.so not built")
    lib = ctypes.CDLL(str(DRIVER))
    lib.api_last_error.restype = ctypes.c_char_p
    lib.api_tree_points.restype = ctypes.c_long
    lib.api_tree_points.argtypes = [_I, _I, _I, _D, _D, _DP, _DP, _SZ, _I, _DP, _DP, _DP]
    lib.api_kernel.argtypes = [_I, _I, _I, _D, _D, _DP, _DP, _D, _DP, _DP]
    lib.api_direct_eval.argtypes = [_I, _I, _I, _D, _D, _DP, _DP, _SZ, _I, _DP, _SZ, _DP]
    lib.api_tree_eval.restype = ctypes.c_long
    lib.api_tree_eval.argtypes = [_I, _I, _I, _D, _D, _DP, _DP, _SZ, _I, _DP, _SZ, _D, _D, _I, _DP, _DP, _DP]
    for f in ("api_likelihood_cv", "api_lsq_conv_cv"):
        getattr(lib, f).argtypes = [_I, _I, _I, _D, _D, _DP, _DP, _SZ, _I, _DP]
    lib.api_lsq_numint_cv.argtypes = [_I, _I, _D, _D, _DP, _DP, _SZ, _I, _D, _D, _I, _D, _D, _I, _DP]
    lib.api_adapt.restype = ctypes.c_long
    lib.api_adapt.argtypes = [_I, _I, _I, _D, _D, _DP, _DP, _SZ, _I, _D, _DP, _DP, _DP, _DP, _SZ, _DP, _DP]
    lib.api_hip_direct_range.argtypes = [_I, _I, _I, _D, _D, _DP, _SZ, _DP, _SZ, _SZ, _SZ, _SZ, _SZ, _I, _DP]
    lib.api_cv_grid.argtypes = [_I, _DP, _DP, _SZ, _I, _DP, _DP, _SZ, _I, _D, _DP, _DP, _DP, _DP]
    lib.api_grid_layouts.argtypes = [_DP, _SZ, _DP, _SZ, _D, _D, _I, _D, _D, _I, _DP, _DP]
    lib.api_marginal_eval.argtypes = [_I, _I, _D, _D, _DP, _DP, _SZ, _I, _D, _I, _DP, _SZ, _DP, _DP, _SZ, _DP]
    lib.api_simulate.argtypes = [_I, _I, _I, _D, _D, _DP, _DP, _SZ, _I, ctypes.c_ulonglong, _SZ, _I, _DP]
    lib.api_tree_dump.restype = ctypes.c_long
    lib.api_tree_dump.argtypes = [_DP, _DP, _SZ, _I, _I, _D, _DP, _SZ, _DP]
    return lib


def api_simulate(api, kind, dim, kf, hx, hy, pts, w, seed, n, batched, leaf_nmax=2):
    out = np.empty((n, dim))
    rc = api.api_simulate(kind, dim, kf, hx, hy, _p(pts), _p(w), len(pts), leaf_nmax, seed, n, int(batched), _p(out))
    assert rc == 0, _err(api)
    return out


def _err(api):
    return api.api_last_error().decode()


# ---- host only ------------------------------------------------------------------------
def test_simulate_matches_reference_fixture_on_the_host(api):
    # n single-point KernelDensity::simulate calls with std::mt19937_64(seed),
    # against the same calls of the reference (tests/golden/simulate_ref.npz):
    # bit-identical (same draws in the same order, same arithmetic)
    f = np.load(GOLDEN / "simulate_ref.npz")
    pts, w = f["pts"], f["w"]
    for i, (kind, dim, kf, hx, hy) in enumerate(f["cases"]):
        kind, dim, kf = int(kind), int(dim), int(kf)
        got = api_simulate(api, kind, dim, kf, hx, hy, pts[:, :dim], w, 100 + i, 2000, batched=False)
        np.testing.assert_array_equal(got, f[f"sim_{i}"], err_msg=f"case {i}")



def test_kdtree_order_matches_reference(api, reference):
    rng = np.random.default_rng(3)
    pts = rng.normal(size=(3000, 2))
    pts[100:130] = pts[5]  # exact duplicates are merged
    w = rng.uniform(0.5, 1.5, 3000)
    for leaf in (2, 16, 1024):
        c_ref, w_ref, m_ref, _ = reference.tree_points(GAUSS, 0, 0.3, 0.3, pts, w, leaf_nmax=leaf)
        n = len(pts)
        c, ww, m = np.empty((n, 2)), np.empty(n), np.empty(n)
        k = api.api_tree_points(GAUSS, 2, 0, 0.3, 0.3, _p(pts), _p(w), n, leaf, _p(c), _p(ww), _p(m))
        assert k == len(c_ref), _err(api)
        np.testing.assert_array_equal(c[:k], c_ref)
        np.testing.assert_allclose(ww[:k], w_ref, rtol=1e-15)
        np.testing.assert_allclose(m[:k], m_ref, rtol=1e-15)


def test_kdtree_parallel_build_matches_reference(api, reference):
    # 300,000 points take the multi-threaded build (sorted chunks merged,
    # sibling subtrees partitioned concurrently): same order as the
    # reference's single-threaded Kdtree
    rng = np.random.default_rng(4)
    n = 300_000
    pts = rng.normal(size=(n, 2))
    w = rng.uniform(0.5, 1.5, n)
    for leaf in (32, 4096):
        c_ref, w_ref, m_ref, _ = reference.tree_points(EPAN, 1, 0.1, 0.1, pts, w, leaf_nmax=leaf)
        c, ww, m = np.empty((n, 2)), np.empty(n), np.empty(n)
        k = api.api_tree_points(EPAN, 2, 1, 0.1, 0.1, _p(pts), _p(w), n, leaf, _p(c), _p(ww), _p(m))
        assert k == len(c_ref) == n, _err(api)
        np.testing.assert_array_equal(c, c_ref)
        np.testing.assert_array_equal(ww, w_ref)
        np.testing.assert_array_equal(m, m_ref)


@pytest.mark.gpu
def test_kdtree_device_sorted_build_matches_reference(api, reference):
    # with a HIP device the build's lexicographic sort runs on it
    # (bk_lexico_order); the tree's point order and merged duplicates must
    # still be the reference's
    rng = np.random.default_rng(8)
    n = 400_000
    pts = rng.normal(size=(n, 2))
    pts[::1000] = pts[3]  # exact duplicates, merged
    w = rng.uniform(0.5, 1.5, n)
    for leaf in (32, 4096):
        c_ref, w_ref, m_ref, _ = reference.tree_points(EPAN, 1, 0.1, 0.1, pts, w, leaf_nmax=leaf)
        c, ww, m = np.empty((n, 2)), np.empty(n), np.empty(n)
        k = api.api_tree_points(EPAN, 2, 1, 0.1, 0.1, _p(pts), _p(w), n, leaf, _p(c), _p(ww), _p(m))
        assert k == len(c_ref), _err(api)
        np.testing.assert_array_equal(c[:k], c_ref)
        np.testing.assert_allclose(ww[:k], w_ref, rtol=1e-14)
        np.testing.assert_allclose(m[:k], m_ref, rtol=1e-14)


@pytest.mark.gpu
def test_kdtree_device_build_equals_host_build(api):
    # the device steps of a large build (sort, runs of duplicates and root
    # box in bk_lexico_runs, the duplicates' attribute merge in bk_run_merge,
    # the partition over records, node attributes on host threads, and the
    # refresh after adapt_density) give the host build's tree bit for bit:
    # points, merged attributes, splits, cells and every node's attributes
    rng = np.random.default_rng(12)
    n = 400_000
    pts = rng.normal(size=(n, 2))
    pts[::997] = pts[11]  # exact duplicates, merged
    pts[5::1999, 0] = pts[7, 0]  # equal x, distinct y
    w = rng.uniform(0.5, 1.5, n)
    for leaf, alpha in ((32, 0.0), (4096, 0.0), (256, 0.5)):
        outs = []
        for use_device in (1, 0):
            P, N = np.zeros((n, 6)), np.zeros((100_000, 13))
            k = api.api_tree_dump(_p(pts), _p(w), n, leaf, use_device, alpha, _p(P), 100_000, _p(N))
            assert 0 < k <= 100_000, _err(api)
            outs.append((P, N[:k]))
        np.testing.assert_array_equal(outs[0][0], outs[1][0], err_msg=f"points, leaf {leaf}, alpha {alpha}")
        np.testing.assert_array_equal(outs[0][1], outs[1][1], err_msg=f"nodes, leaf {leaf}, alpha {alpha}")


@pytest.mark.parametrize("kind,dim", [(EPAN, 1), (EPAN, 2), (GAUSS, 1), (GAUSS, 2), (EPANPROD, 2), (GAUSSPROD, 2),
                                      (EPANCONV1, 1), (EPANPRODCONV, 2), (GAUSSCONV1, 1), (GAUSSPRODCONV, 2)])
@pytest.mark.parametrize("kf", [0, 1])
def test_kernel_classes_match_reference(api, reference, kind, dim, kf):
    rng = np.random.default_rng(kind * 10 + dim)
    hx, hy = 0.37, 0.61
    for _ in range(20):
        p, q = rng.normal(size=dim), rng.normal(size=dim)
        a = float(rng.uniform(0.5, 2.0))
        n_ref, u_ref = reference.kernel(kind, dim, kf, hx, hy, p, q, a)
        n_out, u_out = _D(), _D()
        assert api.api_kernel(kind, dim, kf, hx, hy, _p(p), _p(q), a, ctypes.byref(n_out), ctypes.byref(u_out)) == 0
        rtol = 1e-14 if kf == 0 else 2e-6
        assert n_out.value == pytest.approx(n_ref, rel=rtol)
        assert u_out.value == pytest.approx(u_ref, rel=rtol, abs=1e-6 if kf else 1e-15)


# ---- GPU ------------------------------------------------------------------------------------
@pytest.mark.gpu
@pytest.mark.parametrize("i", range(10))
@pytest.mark.parametrize("kf", [0, 1])
def test_direct_eval_matches_reference_fixture(api, oracle, i, kf):
    f = np.load(GOLDEN / "direct_ref.npz")
    kind, dim, hx, hy = f["cases"][i]
    kind, dim = int(kind), int(dim)
    pts, w, q = f["pts"][:, :dim], f["w"], f["q"][:, :dim]
    want = f[f"direct_{i}_kf{kf}"]
    out = np.empty(len(q))
    assert api.api_direct_eval(kind, dim, kf, hx, hy, _p(pts), _p(w), len(pts), 2, _p(q), len(q), _p(out)) == 0, \
        _err(api)
    big = np.abs(want) > 1e-10 * np.abs(want).max()
    if kf == 0:
        # KernelDensity<D, Kernel<double>, double>: FP64 end to end
        np.testing.assert_allclose(out[big], want[big], rtol=1e-12)
    else:
        # Kernel<float>: the reference evaluates each term in FP32, which near a
        # compact kernel's support edge is itself only ~1e-5 accurate; the bar
        # is the FP64 value on the same FP32-rounded inputs, and the reference's
        # float result agrees with it to its own FP32 accuracy.
        hx32, hy32 = float(np.float32(hx)), float(np.float32(hy))
        exact = oracle.direct(kind, hx32, hy32, ds.round_f32(pts), ds.round_f32(q),
                              mass=ds.round_f32(w / w.sum()), normalize=True)
        np.testing.assert_allclose(out[big], exact[big], rtol=1e-5)
        np.testing.assert_allclose(out[big], want[big], rtol=1e-4)


@pytest.mark.gpu
@pytest.mark.parametrize("kind,hx,hy,tag", [(GAUSS, 0.3, 0.3, "gauss"), (EPANPROD, 0.4, 0.55, "epanprod")])
def test_dual_tree_matches_reference(api, oracle, kind, hx, hy, tag):
    # Same query tree order as the reference; bounds bracket the exact
    # density and meet the tolerances (rel 1e-3, abs 1e-6).
    f = np.load(GOLDEN / "tree_ref.npz")
    pts, w, q = f["pts"], f["w"], f["q"]
    nq = len(q)
    qo, lo, up = np.empty((nq, 2)), np.empty(nq), np.empty(nq)
    k = api.api_tree_eval(kind, 2, 0, hx, hy, _p(pts), _p(w), len(pts), 16, _p(q), nq, 1e-3, 1e-6, 16, _p(qo),
                          _p(lo), _p(up))
    assert k == len(f[f"q_{tag}"]), _err(api)
    np.testing.assert_array_equal(qo[:k], f[f"q_{tag}"])
    exact = oracle.direct(kind, hx, hy, pts, qo[:k], mass=w / w.sum(), normalize=True)
    slack = 1e-12 * np.maximum(exact, 1.0)
    assert np.all(lo[:k] <= exact + slack) and np.all(exact <= up[:k] + slack)
    mid = 0.5 * (lo[:k] + up[:k])
    assert np.all(np.abs(mid - exact) <= np.maximum(1e-6, 1e-3 * exact) + slack)
    ref_mid = 0.5 * (f[f"lower_{tag}"] + f[f"upper_{tag}"])
    assert np.all(np.abs(mid - ref_mid) <= 2 * np.maximum(1e-6, 1e-3 * exact) + slack)


@pytest.mark.gpu
def test_dual_tree_large_leaves_float_kernel(api, oracle):
    # the configuration of the reference's GPU benchmark (test/test_kde11.cc:
    # float Epanechnikov kernel, leaf 4096, rel 1e-6, abs 1e-10) at N = 20000
    rng = np.random.default_rng(11)
    pts = rng.normal(size=(20000, 2))
    w = np.ones(20000)
    nq = len(pts)
    qo, lo, up = np.empty((nq, 2)), np.empty(nq), np.empty(nq)
    k = api.api_tree_eval(EPAN, 2, 1, 0.1, 0.1, _p(pts), _p(w), nq, 4096, _p(pts), nq, 1e-6, 1e-10, 4096, _p(qo),
                          _p(lo), _p(up))
    assert k == nq, _err(api)
    exact = oracle.direct(EPAN, float(np.float32(0.1)), None, ds.round_f32(pts), ds.round_f32(qo),
                          mass=ds.round_f32(w / w.sum()), normalize=True)
    mid = 0.5 * (lo + up)
    np.testing.assert_allclose(mid, exact, rtol=2e-5, atol=1e-10)


@pytest.mark.gpu
@pytest.mark.parametrize("i", range(4))
def test_cv_scores_match_reference_fixture(api, i):
    f = np.load(GOLDEN / "cv_ref.npz")
    kind, dim, hx, hy = f["cases"][i]
    kind, dim = int(kind), int(dim)
    pts, w = f["pts"][:, :dim], f["w"]
    lcv, lscv = _D(), _D()
    assert api.api_likelihood_cv(kind, dim, 0, hx, hy, _p(pts), _p(w), len(pts), 2, ctypes.byref(lcv)) == 0, _err(api)
    assert api.api_lsq_conv_cv(kind, dim, 0, hx, hy, _p(pts), _p(w), len(pts), 2, ctypes.byref(lscv)) == 0, _err(api)
    tol = 1e-6 if (kind, dim) == (GAUSS, 1) else 1e-9  # float-narrowed 1-D double Gaussian normalisation
    assert lcv.value == pytest.approx(float(f[f"lcv_{i}"]), rel=tol)
    assert lscv.value == pytest.approx(float(f[f"lscv_{i}"]), rel=tol)


@pytest.mark.gpu
def test_numint_lscv_matches_reference_fixture(api):
    f = np.load(GOLDEN / "numint_ref.npz")
    out = _D()
    assert api.api_lsq_numint_cv(GAUSSPROD, 0, 0.3, 0.45, _p(f["pts"]), _p(f["w"]), len(f["w"]), 2, -4.0, 4.0, 100,
                                 -4.0, 4.0, 100, ctypes.byref(out)) == 0, _err(api)
    assert out.value == pytest.approx(float(f["numint_gaussprod"]), rel=1e-9)


@pytest.mark.gpu
def test_adapt_density_matches_reference_fixture(api, oracle):
    f = np.load(GOLDEN / "adapt_ref.npz")
    pts, w = f["pts"], f["w"]
    n = len(pts)
    q = ds.uniform_box(300, -2.5, 2.5, 5)
    for alpha in (0.0, 0.5, 1.0):
        tag = str(alpha).replace(".", "p")
        c, a, m = np.empty((n, 2)), np.empty(n), np.empty(n)
        vals, lcv = np.empty(len(q)), _D()
        k = api.api_adapt(GAUSSPROD, 2, 0, 0.3, 0.45, _p(pts), _p(w), n, 2, alpha, _p(c), _p(a), _p(m), _p(q), len(q),
                          _p(vals), ctypes.byref(lcv))
        assert k == n, _err(api)
        np.testing.assert_array_equal(c, f[f"coords_{tag}"])
        np.testing.assert_allclose(a, f[f"abw_{tag}"], rtol=1e-9)
        np.testing.assert_allclose(m, f[f"mass_{tag}"], rtol=1e-9)
        want = oracle.direct(GAUSSPROD, 0.3, 0.45, c, q, mass=m, abw=a, normalize=True)
        np.testing.assert_allclose(vals, want, rtol=1e-12, atol=1e-300)
        assert lcv.value == pytest.approx(oracle.likelihood_cv(GAUSSPROD, 0.3, 0.45, c, f[f"w_{tag}"], mass=m, abw=a),
                                          rel=1e-10)


@pytest.mark.gpu
@pytest.mark.parametrize("kf", [0, 1])
def test_hip_direct_kde_ranges_and_errors(api, oracle, kf):
    rng = np.random.default_rng(2)
    pts = rng.normal(size=(3000, 2))
    q = rng.uniform(-3, 3, (1200, 2))
    out = np.empty(1200)
    for r_i, r_j, q_i, q_j in ((0, 2999, 0, 1199), (17, 1717, 100, 900), (2999, 2999, 1199, 1199)):
        for norm in (0, 1):
            st = api.api_hip_direct_range(GAUSSPROD, 2, kf, 0.2, 0.3, _p(pts), 3000, _p(q), 1200, r_i, r_j, q_i, q_j,
                                          norm, _p(out))
            assert st == 0, _err(api)
            P = ds.round_f32(pts) if kf else pts
            Q = ds.round_f32(q) if kf else q
            want = oracle.direct(GAUSSPROD, 0.2, 0.3, P[r_i:r_j + 1], Q[q_i:q_j + 1], mass=np.full(r_j - r_i + 1, 1 / 3000),
                                 normalize=bool(norm))
            got = out[:q_j - q_i + 1]
            big = want > 1e-10 * want.max() if want.max() > 0 else np.zeros(len(want), bool)
            np.testing.assert_allclose(got[big], want[big], rtol=1e-12 if kf == 0 else 1e-5)
    assert api.api_hip_direct_range(GAUSSPROD, 2, kf, 0.2, 0.3, _p(pts), 3000, _p(q), 1200, 0, 3000, 0, 10, 0,
                                    _p(out)) == -102  # std::out_of_range
    assert api.api_hip_direct_range(GAUSSPROD, 2, kf, 0.2, 0.3, _p(pts), 3000, _p(q), 1200, 5, 4, 0, 10, 0,
                                    _p(out)) == -101  # std::invalid_argument


def test_bandwidth_grid_layouts(api):
    """make_crossed_vector_pairs: v1 outer, v2 inner; generate_2dgrid: end
    included, y outer (grid_search/cv_utils_impl.h:15-25, 43-59)."""
    v1, v2 = np.array([0.1, 0.2]), np.array([0.3, 0.4, 0.5])
    crossed = np.empty((6, 2))
    gen = np.empty((6, 2))
    assert api.api_grid_layouts(_p(v1), 2, _p(v2), 3, 0.0, 1.0, 2, 0.0, 1.0, 1, _p(crossed), _p(gen)) == 0
    np.testing.assert_array_equal(crossed, [[0.1, 0.3], [0.1, 0.4], [0.1, 0.5], [0.2, 0.3], [0.2, 0.4], [0.2, 0.5]])
    np.testing.assert_array_equal(gen, [[0.0, 0.0], [0.5, 0.0], [1.0, 0.0], [0.0, 1.0], [0.5, 1.0], [1.0, 1.0]])


@pytest.mark.gpu
@pytest.mark.parametrize("adapt", [0, 1])
@pytest.mark.parametrize("kind", [EPANPROD, GAUSSPROD])
def test_grid_search_matches_per_bandwidth_calls(api, kind, adapt):
    """The batched grid equals the per-bandwidth KernelDensity cross-validations
    grid_search makes (FP64 device sums) within the FP32 grid's bound, and the
    reference's own scores at the fixture bandwidths."""
    f = np.load(GOLDEN / "cv_ref.npz")
    pts, w = f["pts"], f["w"]
    case = [i for i, c in enumerate(f["cases"]) if int(c[0]) == kind][0]
    hx0, hy0 = float(f["cases"][case][2]), float(f["cases"][case][3])
    pairs = [(hx0, hy0), (0.2, 0.25), (0.2, 0.6), (0.5, 0.25), (hx0, 0.6)]
    hx = np.array([a for a, _ in pairs])
    hy = np.array([b for _, b in pairs])
    out = [np.empty(len(pairs)) for _ in range(4)]
    assert api.api_cv_grid(kind, _p(pts), _p(w), len(pts), 16, _p(hx), _p(hy), len(pairs), adapt, 0.5,
                           *[_p(o) for o in out]) == 0, _err(api)
    lcv_g, lscv_g, lcv_l, lscv_l = out
    finite = np.isfinite(lcv_l)
    np.testing.assert_array_equal(np.isfinite(lcv_g), finite)
    np.testing.assert_allclose(lcv_g[finite], lcv_l[finite], rtol=0, atol=1e-5)
    np.testing.assert_allclose(lscv_g, lscv_l, rtol=1e-5, atol=1e-7)
    if not adapt:
        assert lcv_g[0] == pytest.approx(float(f[f"lcv_{case}"]), rel=1e-5, abs=1e-6)
        assert lscv_g[0] == pytest.approx(float(f[f"lscv_{case}"]), rel=1e-5, abs=1e-7)


@pytest.mark.gpu
@pytest.mark.parametrize("kind,hx,hy,kf", [(GAUSSPROD, 0.3, 0.45, 0), (EPANPROD, 0.4, 0.55, 0),
                                           (GAUSSPROD, 0.3, 0.45, 1)])
@pytest.mark.parametrize("alpha", [0.0, 0.5])
@pytest.mark.parametrize("dim_idx", [0, 1])
def test_marginal_density_is_the_integral_of_the_2d_density(api, oracle, kind, hx, hy, kf, alpha, dim_idx):
    # MarginalDensityAssociator (include/MarginalDensityAssociator.h:13-128):
    # the marginal along one axis equals the 2-D density integrated over the
    # other axis (fine trapezoid rule on the GPU's own 2-D values), and without
    # adaptation it is the 1-D density of that coordinate (oracle)
    f = np.load(GOLDEN / "tree_ref.npz")
    pts, w = f["pts"], f["w"]
    t = np.linspace(-2.5, 2.5, 21)
    other = np.linspace(-25.0, 25.0, 25001)  # adapted bandwidth factors reach about 5
    q2 = np.array([(a, b) if dim_idx == 0 else (b, a) for a in t for b in other])
    m1, d2 = np.empty(len(t)), np.empty(len(q2))
    assert api.api_marginal_eval(kind, kf, hx, hy, _p(pts), _p(w), len(pts), 16, alpha, dim_idx, _p(t), len(t),
                                 _p(m1), _p(q2), len(q2), _p(d2)) == 0, _err(api)
    integral = np.trapezoid(d2.reshape(len(t), len(other)), other, axis=1)
    tol = 2e-5 if kf else 1e-9
    rtol = (1e-4 if kind == EPANPROD else 1e-7) + tol  # trapezoid error at the Epanechnikov support kinks
    np.testing.assert_allclose(m1, integral, rtol=rtol, atol=1e-9)
    if alpha == 0.0:
        h = hx if dim_idx == 0 else hy
        kind1 = GAUSS if kind == GAUSSPROD else EPAN
        x1 = pts[:, dim_idx:dim_idx + 1]
        want = oracle.direct(kind1, h, h, ds.round_f32(x1) if kf else x1, t[:, None], mass=w / w.sum(),
                             normalize=True)
        if kind1 == GAUSS and not kf:
            # GaussianTraits<1,double>::normalization takes h as float
            # (include/Kernels/KernelTraits.h:143), kept as the reference has it
            want = want * (h / float(np.float32(h)))
        np.testing.assert_allclose(m1, want, rtol=tol, atol=1e-12)
