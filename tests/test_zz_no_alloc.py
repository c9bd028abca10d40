"""bk_eval_range_device_noalloc: the range evaluation that allocates nothing.

Its results are compared, through the C ABI, with bk_eval_range_device on
the same context and range, on calls in a row whose inputs shrink (the
scratch array keeps the previous call's contents, which must not leak into
the next one), and with the FP64 oracle on a sample of queries.  The two
entry points run the same kernels on the same tile-aligned stages; they may
differ only in the split count (the no-alloc variant lowers it until the
partial sums fit its scratch), i.e. in the FP64 order in which the split
sums are combined, so they must agree to 1e-12 relative.  Against the oracle
the bar is that of every FP32 context (1e-5 relative above 1e-10 of C * sum m)
and 1e-12 for FP64 contexts.
"""
from __future__ import annotations

import numpy as np
import pytest

from bbrcit_kde_amd import datasets as ds

pytestmark = pytest.mark.gpu

EPAN, GAUSS, EPANPROD, GAUSSPROD = 1, 2, 3, 4
FLOOR_REL = 1e-10


def _dev_pair(torch, n):
    return (torch.full((n,), float("nan"), dtype=torch.float64, device="cuda"),
            torch.full((n,), float("nan"), dtype=torch.float64, device="cuda"))


def _run_both(torch, kde, desc, r_i, r_j, q_i, q_j):
    n = q_j - q_i + 1
    want_dev, got_dev = _dev_pair(torch, n)
    torch.cuda.synchronize()
    kde.eval_device(desc, r_i, r_j, q_i, q_j, want_dev.data_ptr(), normalize=True)
    kde.eval_device_noalloc(desc, r_i, r_j, q_i, q_j, got_dev.data_ptr(), normalize=True)
    torch.cuda.synchronize()
    return got_dev.cpu().numpy(), want_dev.cpu().numpy()


def _check_oracle(oracle, kind, hx, hy, pts, q, mass, abw, r_i, r_j, q_i, q_j, got, rtol, f32, sample=300):
    rng = np.random.default_rng(5)
    idx = np.sort(rng.choice(q_j - q_i + 1, size=min(sample, q_j - q_i + 1), replace=False))
    P = pts[r_i:r_j + 1]
    M = mass[r_i:r_j + 1]
    A = None if abw is None else abw[r_i:r_j + 1]
    Q = q[q_i:q_j + 1][idx]
    if f32:
        P, M, Q = ds.round_f32(P), ds.round_f32(M), ds.round_f32(Q)
        A = None if A is None else ds.round_f32(A)
    want = oracle.direct(kind, hx, hy, P, Q, mass=M, abw=A, normalize=True)
    bound = oracle.normalization(kind, pts.shape[1], hx, hy) * M.sum()
    floor = FLOOR_REL * bound
    g = got[idx]
    above = np.abs(want) >= floor
    assert above.any()
    rel = np.abs(g[above] - want[above]) / np.abs(want[above])
    assert rel.max() <= rtol, f"max rel err vs oracle {rel.max():.3e}"
    if (~above).any():
        assert np.abs(g[~above] - want[~above]).max() <= rtol * floor


def _agree(got, want):
    assert np.all(np.isfinite(got)), "non-finite output (a query was not written)"
    scale = np.maximum(np.abs(want), 1e-300)
    rel = np.abs(got - want) / scale
    assert rel.max() <= 1e-12, f"noalloc vs eval_range_device: max rel diff {rel.max():.3e}"


def test_noalloc_matches_eval_range_device_on_shrinking_calls(bk, oracle):
    torch = pytest.importorskip("torch")
    # call 1: the 1M x 1M headline shape's kernel on 400k mixture refs (kd order)
    # x 300k uniform queries, FP32 Gaussian (partial sums of several splits)
    pts, w = ds.gaussian_mixture(400_000, seed=41)
    pts = pts[ds.kd_order(pts, leaf_nmax=256)]
    w = np.ones(len(pts))
    mass = w / w.sum()
    q = ds.uniform_box(300_000, -4.0, 4.0, seed=42)
    print("[noalloc] call 1: Gaussian 400k x 300k", flush=True)
    with bk.DirectKde(pts, q, mass) as kde:
        desc = bk.KernelDesc(GAUSS, 2, 0.05)
        got, want = _run_both(torch, kde, desc, 0, len(pts) - 1, 0, len(q) - 1)
        _agree(got, want)
        _check_oracle(oracle, GAUSS, 0.05, 0.05, pts, q, mass, None, 0, len(pts) - 1, 0, len(q) - 1, got, 1e-5,
                      True)
        # call 2 in a row, smaller: a sub-range of the same sets with the
        # Epanechnikov product kernel (other features, fewer queries)
        print("[noalloc] call 2: Epanechnikov product, sub-ranges", flush=True)
        desc = bk.KernelDesc(EPANPROD, 2, 0.08, 0.12)
        r_i, r_j, q_i, q_j = 1000, 150_999, 777, 40_776
        got, want = _run_both(torch, kde, desc, r_i, r_j, q_i, q_j)
        _agree(got, want)
        _check_oracle(oracle, EPANPROD, 0.08, 0.12, pts, q, mass, None, r_i, r_j, q_i, q_j, got, 1e-5, True)
        # call 3: one query, one reference
        got, want = _run_both(torch, kde, desc, 5, 5, 9, 9)
        _agree(got, want)


def test_noalloc_other_paths_and_smaller_second_call(bk, oracle):
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(7)
    # FP64 context with adaptive bandwidths, then a smaller FP64 context
    for n_ref, n_q in ((60_000, 20_000), (3_000, 1_000)):
        print(f"[noalloc] FP64 Gaussian product {n_ref} x {n_q}", flush=True)
        pts, w = ds.gaussian_mixture(n_ref, seed=n_ref)
        mass = w / w.sum()
        abw = rng.uniform(0.6, 1.8, n_ref)
        q = ds.uniform_box(n_q, -3.0, 3.0, seed=n_q)
        with bk.DirectKde(pts, q, mass, abw, precision=bk.BK_PREC_F64) as kde:
            desc = bk.KernelDesc(GAUSSPROD, 2, 0.1, 0.16)
            got, want = _run_both(torch, kde, desc, 0, n_ref - 1, 0, n_q - 1)
            _agree(got, want)
            _check_oracle(oracle, GAUSSPROD, 0.1, 0.16, pts, q, mass, abw, 0, n_ref - 1, 0, n_q - 1, got, 1e-12,
                          False, sample=100)
    # D = 16 (Gram form on MFMA) and D = 3 (general form)
    for dim, n_ref, n_q in ((16, 20_000, 3_000), (3, 5_000, 700)):
        print(f"[noalloc] Gaussian D={dim} {n_ref} x {n_q}", flush=True)
        pts = rng.standard_normal((n_ref, dim))
        q = rng.standard_normal((n_q, dim))
        mass = np.full(n_ref, 1.0 / n_ref)
        with bk.DirectKde(pts, q, mass) as kde:
            desc = bk.KernelDesc(GAUSS, dim, 0.5)
            got, want = _run_both(torch, kde, desc, 0, n_ref - 1, 0, n_q - 1)
            _agree(got, want)
            _check_oracle(oracle, GAUSS, 0.5, 0.5, pts, q, mass, None, 0, n_ref - 1, 0, n_q - 1, got, 1e-5, True,
                          sample=64)


def test_noalloc_rejects_reference_sets_beyond_its_scratch(bk):
    torch = pytest.importorskip("torch")
    # FP64 features take 64.6 bytes per reference: 1.1M references exceed 64 MiB
    n = 1_100_000
    pts = np.random.default_rng(3).uniform(-1.0, 1.0, (n, 2))
    q = pts[:1000]
    out = torch.zeros(1000, dtype=torch.float64, device="cuda")
    with bk.DirectKde(pts, q, precision=bk.BK_PREC_F64) as kde:
        with pytest.raises(ValueError, match="scratch"):
            kde.eval_device_noalloc(bk.KernelDesc(GAUSS, 2, 0.1), 0, n - 1, 0, 999, out.data_ptr())
        # the context still works through the allocating entry point
        kde.eval_device(bk.KernelDesc(GAUSS, 2, 0.1), 0, n - 1, 0, 999, out.data_ptr())
        torch.cuda.synchronize()
        assert np.all(out.cpu().numpy() > 0.0)
