"""GPU parity of the grid_search tool against the reference's pipeline.

The fixture (tests/golden/grid_search_ref.npz, written by make_golden.py from
the reference's own CPU implementation) holds the steps of the reference's
grid_search/grid_search.cc on 600 weighted points:
  * lsq_conv scores of EpanechnikovProductKernel2d<float> over an automatic
    4 x 4 bandwidth grid (dual tree, rel 1e-6, abs 1e-8),
  * the density at the best pair over a 21 x 21 plotting grid,
  * the adaptive density (alpha 0.5) over the same grid.
The tool writes its numbers with the stream's default 6 significant digits,
so scores and densities are compared at 1e-4 relative (the reference's own
float-kernel sums differ from exact sums by up to 2e-5 at the support edge),
with an absolute floor of 1e-6 (the dual tree's tolerance of 1e-8 rounds to
that in the files).
"""
from __future__ import annotations

import re
import subprocess
from pathlib import Path

import numpy as np
import pytest

from bbrcit_kde_amd.build import GRID_SEARCH_BIN

pytestmark = pytest.mark.gpu

GOLDEN = Path(__file__).resolve().parent / "golden" / "grid_search_ref.npz"
RTOL, ATOL = 1e-4, 1e-6


def write_inputs(tmp: Path, pts, w, **overrides) -> Path:
    with open(tmp / "points.dat", "w") as f:
        for (x, y), ww in zip(pts, w):
            f.write(f"{float(x):.17g} {float(y):.17g} {float(ww):.17g}\n")
    cfg = {
        "rel_tol": "1e-6", "abs_tol": "1e-8", "cuda_device_number": "0", "gpu_block_size": "128",
        "refpt_max_leaf_size": "32", "qgrid_max_leaf_size": "32", "use_adaptive_cv": "false",
        "cv_method": "lsq_conv", "skip_cross_validation": "false", "input_refpts_fname": str(tmp / "points.dat"),
        "output_scatter_fname": str(tmp / "scatter.csv"), "base_bwx": "0.3", "base_bwy": "0.35", "alpha": "0.5",
        "adapted_base_bwx": "0.3", "adapted_base_bwy": "0.35", "use_manual_bwgrid": "false",
        "output_gridsearch_fname": str(tmp / "gridsearch.csv"), "manual_bwx": "0.2 0.4", "manual_bwy": "0.3 0.5",
        "start_bwx": "0.2", "end_bwx": "0.6", "steps_bwx": "3", "start_bwy": "0.25", "end_bwy": "0.7",
        "steps_bwy": "3", "start_qix": "-3.0", "end_qix": "3.0", "steps_qix": "40", "start_qiy": "-3.0",
        "end_qiy": "3.0", "steps_qiy": "40", "use_gridsearch_best": "true", "start_qx": "-2.5", "end_qx": "2.5",
        "steps_qx": "20", "start_qy": "-2.5", "end_qy": "2.5", "steps_qy": "20",
        "output_eval_fname": str(tmp / "kde.csv"), "output_adaptive_eval_fname": str(tmp / "akde.csv"),
    }
    cfg.update(overrides)
    path = tmp / "run.cfg"
    with open(path, "w") as f:
        f.write("# grid_search configuration\n")
        for k, v in cfg.items():
            f.write(f"{k} = {v}\n")
    return path


def run_tool(cfg: Path, *extra: str) -> subprocess.CompletedProcess:
    assert GRID_SEARCH_BIN.exists(), "grid_search is not built"
    return subprocess.run([str(GRID_SEARCH_BIN), *extra, str(cfg)], capture_output=True, text=True, timeout=300)


def read_grid_file(path: Path):
    lines = path.read_text().strip().split("\n")
    gx = np.array(lines[0].split(), dtype=float)
    gy = np.array(lines[1].split(), dtype=float)
    vals = np.array([np.array(l.split(), dtype=float) for l in lines[2:]])
    return gx, gy, vals


def close(got, want):
    return np.abs(got - want) <= RTOL * np.abs(want) + ATOL


def test_grid_search_matches_reference_pipeline(tmp_path):
    f = np.load(GOLDEN)
    cfg = write_inputs(tmp_path, f["pts"], f["w"])
    res = run_tool(cfg)
    assert res.returncode == 0, res.stderr
    # 1. scatter data: the points as read
    sc = np.loadtxt(tmp_path / "scatter.csv")
    np.testing.assert_allclose(sc, f["pts"], rtol=1e-5, atol=1e-6)
    # 3. scores over the automatic grid, y outer and x inner
    gx, gy, scores = read_grid_file(tmp_path / "gridsearch.csv")
    np.testing.assert_allclose(gx, np.linspace(0.2, 0.6, 4), rtol=1e-5)
    np.testing.assert_allclose(gy, np.linspace(0.25, 0.7, 4), rtol=1e-5)
    assert scores.shape == (4, 4)
    assert np.all(close(scores.ravel(), f["lscv"])), (scores.ravel(), f["lscv"])
    assert f"best bandwidth: ({f['best'][0]:.6g}, {f['best'][1]:.6g})" in res.stdout
    # 4. density at the best pair; 5. adaptive density
    _, _, kde = read_grid_file(tmp_path / "kde.csv")
    assert kde.shape == (21, 21)
    assert np.all(close(kde.ravel(), f["kde"]))
    _, _, akde = read_grid_file(tmp_path / "akde.csv")
    assert np.all(close(akde.ravel(), f["akde"]))


def test_grid_search_adaptive_numint_and_manual_grid(tmp_path, oracle):
    # use_adaptive_cv forces lsq_numint; a manual grid writes no score file;
    # the plotting-grid density then uses the base bandwidths
    f = np.load(GOLDEN)
    cfg = write_inputs(tmp_path, f["pts"], f["w"], use_adaptive_cv="true", use_manual_bwgrid="true")
    res = run_tool(cfg, "--steps_qx=10", "--steps_qy", "10")
    assert res.returncode == 0, res.stderr
    assert "numerical integration" in res.stdout
    scored = re.findall(r"^  \(([-\d.e+]+), ([-\d.e+]+)\) ([-\d.e+]+)$", res.stdout, flags=re.M)
    assert [(float(a), float(b)) for a, b, _ in scored] == [(0.2, 0.3), (0.2, 0.5), (0.4, 0.3), (0.4, 0.5)]
    assert all(np.isfinite(float(c)) for _, _, c in scored)
    assert not (tmp_path / "gridsearch.csv").exists()
    _, _, kde = read_grid_file(tmp_path / "kde.csv")
    assert kde.shape == (11, 11)
    g = np.linspace(-2.5, 2.5, 11)
    q = np.array([(x, y) for y in g for x in g])
    w = f["w"] / f["w"].sum()
    want = oracle.direct(3, 0.3, 0.35, f["pts"], q, mass=w, normalize=True)
    assert np.all(close(kde.ravel(), want))


def test_grid_search_reports_bad_configuration(tmp_path):
    f = np.load(GOLDEN)
    cfg = write_inputs(tmp_path, f["pts"][:50], f["w"][:50], cv_method="fft")
    res = run_tool(cfg)
    assert res.returncode == 1 and "unrecognized cv method" in res.stderr
    cfg = write_inputs(tmp_path, f["pts"][:50], f["w"][:50])
    text = cfg.read_text().replace("rel_tol = 1e-6\n", "")
    cfg.write_text(text)
    res = run_tool(cfg)
    assert res.returncode == 1 and "rel_tol" in res.stderr
