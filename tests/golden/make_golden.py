"""Write the golden fixtures of tests/golden/ from the reference itself.

The values come from the reference's own CPU implementation
(dchao34/bbrcit_kde, compiled from its sources by oracle/Makefile into
oracle/_ref/libbbrcit_ref.so).  Inputs are seeded synthetic points; each
fixture stores inputs and expected outputs only.

    make -C oracle && python tests/golden/make_golden.py             # every fixture
    make -C oracle && python tests/golden/make_golden.py high_dim    # direct_hd_ref.npz only
    make -C oracle && python tests/golden/make_golden.py simulate    # simulate_ref.npz only
"""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

from oracle.oracle import Reference  # noqa: E402

OUT = Path(__file__).resolve().parent

# (kind, dim, hx, hy) cases of the direct sum; kinds as include/bbrcit_hip.h.
DIRECT_CASES = [
    (1, 1, 0.35, 0.35), (1, 2, 0.45, 0.45), (2, 1, 0.25, 0.25), (2, 2, 0.3, 0.3),
    (3, 2, 0.4, 0.55), (4, 2, 0.3, 0.45), (5, 1, 0.35, 0.35), (6, 2, 0.4, 0.55),
    (7, 1, 0.25, 0.25), (8, 2, 0.3, 0.45),
]
CV_CASES = [(1, 1, 0.35, 0.35), (2, 1, 0.25, 0.25), (3, 2, 0.4, 0.55), (4, 2, 0.3, 0.45)]


def inputs(seed=2024, n=600, nq=200):
    rng = np.random.default_rng(seed)
    comp = rng.random(n) < 0.4
    pts = np.where(comp[:, None], rng.normal([-1.0, 0.5], [0.4, 0.6], (n, 2)),
                   rng.normal([0.8, -0.3], [0.5, 0.3], (n, 2)))
    w = rng.uniform(0.5, 1.5, n)
    q = rng.uniform(-2.5, 2.5, (nq, 2))
    return pts, w, q


def grid_search_case():
    """Bandwidth grid ends of the grid_search fixture."""
    return {"bw_x0": 0.2, "bw_x1": 0.6, "bw_y0": 0.25, "bw_y1": 0.7}


# (kind, dim, h) of the generic GaussianKernel<D> / EpanechnikovKernel<D> at D > 2
HIGH_DIM_CASES = [(2, 3, 0.7), (1, 3, 1.5), (2, 8, 0.7), (1, 8, 3.0), (2, 16, 0.7), (1, 16, 5.0)]


def write_high_dim(ref) -> None:
    """direct_eval of the generic kernels at D = 3, 8, 16 (FP64 kernels):
    direct_hd_ref.npz (300 standard-normal references, 50 queries)."""
    rng = np.random.default_rng(3)
    out = {"cases": np.array(HIGH_DIM_CASES, dtype=np.float64)}
    for i, (kind, dim, h) in enumerate(HIGH_DIM_CASES):
        p = rng.normal(size=(300, dim))
        w = rng.uniform(0.5, 1.5, 300)
        q = rng.normal(size=(50, dim)) * 1.2
        out[f"pts_{i}"], out[f"w_{i}"], out[f"q_{i}"] = p, w, q
        out[f"direct_{i}"] = ref.direct_eval(kind, 0, h, h, p, w, q)
    np.savez_compressed(OUT / "direct_hd_ref.npz", **out)


# (kind, dim, kernel_float, hx, hy) of the samplable kernels
SIMULATE_CASES = [(1, 1, 0, 0.3, 0.3), (1, 1, 1, 0.3, 0.3), (2, 1, 0, 0.2, 0.2), (2, 2, 0, 0.25, 0.25),
                  (2, 2, 1, 0.25, 0.25), (3, 2, 0, 0.3, 0.45), (3, 2, 1, 0.3, 0.45), (4, 2, 0, 0.2, 0.35),
                  (4, 2, 1, 0.2, 0.35)]


def write_simulate(ref) -> None:
    """KernelDensity::simulate, 2000 samples per case from std::mt19937_64(seed):
    simulate_ref.npz (the same 600 weighted points as direct_ref.npz)."""
    pts, w, _ = inputs()
    out = {"pts": pts, "w": w, "cases": np.array(SIMULATE_CASES, dtype=np.float64)}
    for i, (kind, dim, kf, hx, hy) in enumerate(SIMULATE_CASES):
        out[f"sim_{i}"] = ref.simulate(kind, kf, hx, hy, pts[:, :dim], w, seed=100 + i, n_samples=2000)
    np.savez_compressed(OUT / "simulate_ref.npz", **out)


def main() -> int:
    ref = Reference()
    if sys.argv[1:] == ["high_dim"]:
        write_high_dim(ref)
        return 0
    if sys.argv[1:] == ["simulate"]:
        write_simulate(ref)
        return 0
    write_high_dim(ref)
    write_simulate(ref)
    pts, w, q = inputs()
    out = {"pts": pts, "w": w, "q": q, "cases": np.array(DIRECT_CASES, dtype=np.float64)}
    for i, (kind, dim, hx, hy) in enumerate(DIRECT_CASES):
        for kf in (0, 1):
            out[f"direct_{i}_kf{kf}"] = ref.direct_eval(kind, kf, hx, hy, pts[:, :dim], w, q[:, :dim])
    np.savez_compressed(OUT / "direct_ref.npz", **out)

    # cross-validation scores with tight dual-tree tolerances (near exact)
    cv = {"pts": pts, "w": w, "cases": np.array(CV_CASES, dtype=np.float64)}
    for i, (kind, dim, hx, hy) in enumerate(CV_CASES):
        cv[f"lcv_{i}"] = ref.likelihood_cv(kind, 0, hx, hy, pts[:, :dim], w, rel=1e-12, abs_err=1e-15)
        cv[f"lscv_{i}"] = ref.lsq_conv_cv(kind, 0, hx, hy, pts[:, :dim], w, rel=1e-12, abs_err=1e-15)
    np.savez_compressed(OUT / "cv_ref.npz", **cv)

    # adaptive densities: tree-ordered points with adapted abw and mass
    ad = {"pts": pts, "w": w}
    for alpha in (0.0, 0.5, 1.0):
        c, ww, a, m = ref.adapt(4, 0, 0.3, 0.45, pts, w, alpha, rel=1e-12, abs_err=1e-15)
        tag = str(alpha).replace(".", "p")
        ad[f"coords_{tag}"], ad[f"w_{tag}"], ad[f"abw_{tag}"], ad[f"mass_{tag}"] = c, ww, a, m
    np.savez_compressed(OUT / "adapt_ref.npz", **ad)

    # dual-tree evaluation with its default-style tolerances: permuted
    # queries and the lower/upper bounds of each
    tr = {"pts": pts, "w": w, "q": q}
    for kind, hx, hy, tag in ((2, 0.3, 0.3, "gauss"), (3, 0.4, 0.55, "epanprod")):
        qo, lo, up = ref.tree_eval(kind, 0, hx, hy, pts, w, q, rel=1e-3, abs_err=1e-6, leaf_nmax=16,
                                   q_leaf_nmax=16)
        tr[f"q_{tag}"], tr[f"lower_{tag}"], tr[f"upper_{tag}"] = qo, lo, up
    np.savez_compressed(OUT / "tree_ref.npz", **tr)

    # least-squares CV by numerical integration on a 100x100 grid
    ni = {"pts": pts, "w": w}
    ni["numint_gaussprod"] = ref.lsq_numint_cv(4, 0, 0.3, 0.45, pts, w, -4.0, 4.0, 100, -4.0, 4.0, 100,
                                               rel=1e-12, abs_err=1e-15)
    np.savez_compressed(OUT / "numint_ref.npz", **ni)
    # the steps of grid_search/grid_search.cc with the reference's own
    # functions: lsq_conv scores over an automatic 4 x 4 bandwidth grid
    # (EpanechnikovProductKernel2d<float>, double tree, rel 1e-6, abs 1e-8,
    # leaves of 32), the density at the best pair over a 21 x 21 plotting
    # grid (dual tree), and the adaptive density (alpha 0.5) there: the
    # reference's adapted masses and abw, summed exactly by the oracle
    from oracle.oracle import Oracle

    gs = grid_search_case()
    grid_bw = [(gs["bw_x0"] + i * (gs["bw_x1"] - gs["bw_x0"]) / 3, gs["bw_y0"] + j * (gs["bw_y1"] - gs["bw_y0"]) / 3)
               for j in range(4) for i in range(4)]
    scores = np.array([ref.lsq_conv_cv(3, 1, hx, hy, pts, w, rel=1e-6, abs_err=1e-8, leaf_nmax=32)
                       for hx, hy in grid_bw])
    bx, by = grid_bw[int(np.argmin(scores))]
    g = np.linspace(-2.5, 2.5, 21)
    qgrid = np.array([(x, y) for y in g for x in g])
    qo, lo, up = ref.tree_eval(3, 1, bx, by, pts, w, qgrid, rel=1e-6, abs_err=1e-8, leaf_nmax=32, q_leaf_nmax=32)
    order = np.lexsort((qo[:, 0], qo[:, 1]))
    c, _, a, m = ref.adapt(3, 1, bx, by, pts, w, 0.5, rel=1e-6, abs_err=1e-8, leaf_nmax=32)
    adaptive = Oracle().direct(3, bx, by, c, qgrid, mass=m, abw=a, normalize=True)
    np.savez_compressed(OUT / "grid_search_ref.npz", pts=pts, w=w, bw=np.array(grid_bw), lscv=scores,
                        best=np.array([bx, by]), qgrid=qgrid, kde=0.5 * (lo + up)[order], akde=adaptive)
    print("fixtures written to", OUT)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
