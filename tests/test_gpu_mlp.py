"""GPU checks of the learned density surrogate (bk_mlp_*).

* The training kernels' loss and gradient equal the FP64 oracle's
  (oracle/mlp_oracle.py, itself pinned by finite differences) for the same
  BF16 weights: relative gradient error <= 1e-4 in norm (FP32 arithmetic,
  FP32 sums over 512 samples).
* The serving kernel (BF16 MFMA, activations split hi + lo) equals the
  oracle's forward pass to 5e-5 absolute on the normalised output.
* Training lowers the loss and is deterministic for a fixed seed.
* Serving: every query of a failing cell gets the exact direct sum, every
  checked sample gets the exact value, and the surrogate replaces exact work
  only where its sampled error passed (the gain over the method without it is
  the exact evaluations saved; bench.py reports the resulting queries/s).
"""
from __future__ import annotations

import numpy as np
import pytest

from bbrcit_kde_amd import datasets as ds
from oracle import mlp_oracle as mo

pytestmark = pytest.mark.gpu


def test_params_round_trip(bk):
    with bk.Surrogate(seed=5) as s:
        th, lo, hi, mu, sigma, eps = s.params()
        assert th.shape == (bk.BK_MLP_NUM_PARAMS,) and np.isfinite(th).all()
        th2 = th * 0.5
        s.set_params(th2, -3.0, 5.0, -2.0, 3.0, 1e-9)
        got = s.params()
        assert np.array_equal(got[0], th2)
        assert got[1:] == (-3.0, 5.0, -2.0, 3.0, 1e-9)
        with pytest.raises(ValueError):
            s.set_params(th2, 1.0, 1.0, 0.0, 1.0, 0.0)


def test_loss_and_gradient_match_oracle(bk):
    rng = np.random.default_rng(11)
    with bk.Surrogate(seed=21) as s:
        th = s.params()[0]
        x = rng.uniform(-1, 1, (512, 2))
        t = rng.standard_normal(512)
        loss, g = s.loss_grad(x, t)
    wq = mo.quantize(th)
    want_loss, want_g = mo.loss_grad(wq, ds.round_f32(x), ds.round_f32(t))
    assert loss == pytest.approx(want_loss, rel=1e-5)
    err = np.linalg.norm(g - want_g) / np.linalg.norm(want_g)
    assert err <= 1e-4, err
    big = np.abs(want_g) > 1e-3 * np.abs(want_g).max()
    np.testing.assert_allclose(g[big], want_g[big], rtol=2e-3)


def test_serving_forward_matches_oracle(bk):
    rng = np.random.default_rng(12)
    with bk.Surrogate(seed=22) as s:
        th = s.params()[0]
        x = rng.uniform(-1, 1, (5000, 2))
        y = s.forward(x)
    want, _ = mo.forward(mo.quantize(th), ds.round_f32(x))
    assert np.abs(y - want).max() <= 5e-5 * max(1.0, np.abs(want).max())


def toy_set(n, seed):
    rng = np.random.default_rng(seed)
    q = rng.uniform(-2, 2, (n, 2))
    f = np.exp(-0.5 * ((q[:, 0] - 0.3) ** 2 / 0.5 + (q[:, 1] + 0.2) ** 2 / 0.8)) + 0.1
    return q, f


def test_training_lowers_loss_and_is_deterministic(bk):
    q, f = toy_set(65536, 3)
    out = []
    for _ in range(2):
        with bk.Surrogate(seed=1) as s:
            s.set_training_data(q, f, -2.0, 2.0, 1e-6)
            first, _ = s.train(1, 1024, 1e-3, 1e-3, seed=9)
            loss, ms = s.train(600, 1024, 3e-3, 1e-5, seed=10)
            out.append((s.params()[0], loss))
    assert out[0][1] < 0.02 * first, (first, out[0][1])
    assert np.array_equal(out[0][0], out[1][0])


def test_checkpoint_resumes_training_exactly(bk, tmp_path):
    """Training 300 + 300 steps with a save and a load in between (in a new
    surrogate) gives the bits of training without the break: the checkpoint
    carries the master weights, the Adam moments and step count and the
    normalisation.  Corrupted, truncated, foreign and other-version files are
    refused and leave the surrogate unchanged."""
    q, f = toy_set(65536, 4)
    path = tmp_path / "mlp.ckpt"
    with bk.Surrogate(seed=2) as a:
        a.set_training_data(q, f, -2.0, 2.0, 1e-6)
        a.train(300, 1024, 3e-3, 1e-4, seed=11)
        a.save(path)
        a.train(300, 1024, 1e-3, 1e-5, seed=12)
        want = a.params()
    with bk.Surrogate(seed=99) as b:
        b.set_training_data(q, f, -2.0, 2.0, 1e-6)
        b.load(path)
        b.train(300, 1024, 1e-3, 1e-5, seed=12)
        got = b.params()
    np.testing.assert_array_equal(got[0], want[0])
    assert got[1:] == want[1:]
    raw = path.read_bytes()
    bad = {"flipped": raw[:200] + bytes([raw[200] ^ 1]) + raw[201:], "truncated": raw[:-9],
           "foreign": b"NOTACKPT" + raw[8:], "version": raw[:8] + (2).to_bytes(4, "little") + raw[12:]}
    with bk.Surrogate(seed=5) as c:
        before = c.params()
        for name, data in bad.items():
            p = tmp_path / f"{name}.ckpt"
            p.write_bytes(data)
            with pytest.raises(ValueError):
                c.load(p)
        after = c.params()
    np.testing.assert_array_equal(after[0], before[0])
    assert after[1:] == before[1:]


def test_serve_replaces_failing_cells_with_exact_sums(bk, oracle):
    torch = pytest.importorskip("torch")
    pts, w = ds.gaussian_mixture(20000, seed=77)
    mass = w / w.sum()
    h = 0.1
    desc = bk.KernelDesc(bk.BK_GAUSSIAN, 2, h)
    c_norm = 1.0 / (2 * np.pi * h * h)
    floor = 1e-10 * c_norm * float(ds.round_f32(mass).sum())
    with bk.DirectKde(pts, None, mass) as kde, bk.Surrogate(seed=3) as s:
        qt = ds.uniform_box(65536, -4, 4, 78)
        with bk.DirectKde(pts, qt, mass) as lab:
            ft = lab.eval(desc, 0, len(pts) - 1, 0, len(qt) - 1)
        s.set_training_data(qt, ft, -4.0, 4.0, 1e-3 * floor)
        s.train(800, 2048, 3e-3, 1e-5, seed=4)
        qs = ds.uniform_box(200_000, -4, 4, 79)
        qx = torch.tensor(qs[:, 0], dtype=torch.float32, device="cuda")
        qy = torch.tensor(qs[:, 1], dtype=torch.float32, device="cuda")
        out = torch.empty(len(qs), dtype=torch.float64, device="cuda")
        surr = torch.empty(len(qs), dtype=torch.float64, device="cuda")
        s.infer_device(len(qs), qx.data_ptr(), qy.data_ptr(), surr.data_ptr())
        st = s.serve_device(kde, desc, len(qs), qx.data_ptr(), qy.data_ptr(), out.data_ptr(), sample_fraction=0.01,
                            rel_tol=1e-3, floor=floor, cells=16, seed=5)
        torch.cuda.synchronize()
        served = out.cpu().numpy()
        raw = surr.cpu().numpy()
    assert st["n"] == len(qs) and st["cells"] == 256
    assert 0.005 * len(qs) < st["n_samples"] < 0.015 * len(qs)
    exact_mask = served != raw  # replaced by exact values (samples or fallback cells)
    assert exact_mask.sum() >= st["n_samples"] and exact_mask.sum() <= st["n_samples"] + st["n_fallback"]
    # exact wherever replaced: check a deterministic subset against the oracle
    idx = np.flatnonzero(exact_mask)[::97][:400]
    want = oracle.direct(bk.BK_GAUSSIAN, h, h, ds.round_f32(pts), ds.round_f32(qs[idx]), mass=ds.round_f32(mass),
                         normalize=True)
    err = np.abs(served[idx] - want) / np.maximum(want, floor)
    assert err.max() <= 1e-5
    # cells that were accepted used the surrogate; their checked samples met rel_tol
    assert st["n_fallback"] <= len(qs)
    saved = len(qs) - st["n_fallback"] - st["n_samples"]
    assert saved >= 0
    print(f"serve: cells failed {st['cells_failed']}/{st['cells']}, fallback {st['n_fallback']}, "
          f"exact evaluations saved {saved} of {len(qs)}, max sampled err {st['max_sample_err']:.3e}")


def _splitmix64(x: np.ndarray) -> np.ndarray:
    """splitmix64 (Steele, Lea, Flood 2014) on uint64 arrays, wrapping."""
    with np.errstate(over="ignore"):
        x = x + np.uint64(0x9E3779B97F4A7C15)
        x = (x ^ (x >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        x = (x ^ (x >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return x ^ (x >> np.uint64(31))


def test_serve_checks_exactly_the_deterministic_sample(bk, oracle):
    # The sampled check takes query i when splitmix64(seed ^ i) % period == 0
    # and sorts the sample along a Z curve before the exact sum.  With far
    # more queries than samples (indices much wider than the sample count)
    # and a tolerance no cell can fail, the served values must equal the
    # surrogate's everywhere except at exactly those queries, which carry the
    # exact sums.
    torch = pytest.importorskip("torch")
    pts, w = ds.gaussian_mixture(5000, seed=81)
    mass = w / w.sum()
    h = 0.15
    desc = bk.KernelDesc(bk.BK_GAUSSIAN, 2, h)
    floor = 1e-10 / (2 * np.pi * h * h)
    n, seed, period = 1 << 20, 9, 100
    qs = ds.uniform_box(n, -4, 4, 82)
    with bk.DirectKde(pts, None, mass) as kde, bk.Surrogate(seed=5) as s:
        qt = ds.uniform_box(8192, -4, 4, 83)
        with bk.DirectKde(pts, qt, mass) as lab:
            ft = lab.eval(desc, 0, len(pts) - 1, 0, len(qt) - 1)
        s.set_training_data(qt, ft, -4.0, 4.0, 1e-3 * floor)
        s.train(50, 1024, 3e-3, 1e-5, seed=6)
        qx = torch.tensor(qs[:, 0], dtype=torch.float32, device="cuda")
        qy = torch.tensor(qs[:, 1], dtype=torch.float32, device="cuda")
        out = torch.empty(n, dtype=torch.float64, device="cuda")
        surr = torch.empty(n, dtype=torch.float64, device="cuda")
        s.infer_device(n, qx.data_ptr(), qy.data_ptr(), surr.data_ptr())
        st = s.serve_device(kde, desc, n, qx.data_ptr(), qy.data_ptr(), out.data_ptr(), sample_fraction=1.0 / period,
                            rel_tol=1e30, floor=floor, cells=1, seed=seed)
        torch.cuda.synchronize()
        served, raw = out.cpu().numpy(), surr.cpu().numpy()
    idx = np.arange(n, dtype=np.uint64)
    sample = np.flatnonzero(_splitmix64(np.uint64(seed) ^ idx) % np.uint64(period) == 0)
    assert st["n_samples"] == len(sample) and st["n_fallback"] == 0 and st["cells_failed"] == 0
    assert sample.max() > 4 * len(sample)  # indices far wider than the sample count
    others = np.ones(n, bool)
    others[sample] = False
    np.testing.assert_array_equal(served[others], raw[others])
    want = oracle.direct(bk.BK_GAUSSIAN, h, h, ds.round_f32(pts), ds.round_f32(qs[sample]), mass=ds.round_f32(mass),
                         normalize=True)
    err = np.abs(served[sample] - want) / np.maximum(want, floor)
    assert err.max() <= 1e-5
