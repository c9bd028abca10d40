"""CPU-side checks of the C ABI: the library loads, exports every function
include/bbrcit_hip.h declares, and reports errors without a GPU.  No compute
call is made here."""
from __future__ import annotations

import ctypes
import re
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]


def declared_functions():
    names = []
    for h in sorted((ROOT / "include").glob("*.h")):
        text = h.read_text()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names += re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(bk_\w+)\s*\(", text, flags=re.M)
    return sorted(set(names))


def test_header_declares_the_boundary():
    names = declared_functions()
    for required in ("bk_create", "bk_destroy", "bk_last_error", "bk_upload", "bk_eval_range",
                     "bk_eval_leaf_pairs", "bk_likelihood_cv", "bk_lsq_conv_cv", "bk_adapt"):
        assert required in names


def test_library_exports_every_declared_symbol(bk):
    lib = ctypes.CDLL(str(bk.library_path()))
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, f"symbols declared but not exported: {missing}"


def test_library_is_gfx950_code_object():
    data = (ROOT / "bbrcit_kde_amd" / "lib" / "libbbrcit_hip.so").read_bytes()
    assert b"gfx950" in data


def _gpu_present(bk) -> bool:
    lib = bk.load_library()
    ctx = ctypes.c_void_p()
    st = lib.bk_create(ctypes.byref(ctx), 0, 0)
    if st == 0:
        lib.bk_destroy(ctx)
        return True
    return False


def test_create_without_gpu_fails_cleanly(bk):
    if _gpu_present(bk):
        pytest.skip("a GPU is present")
    lib = bk.load_library()
    ctx = ctypes.c_void_p()
    st = lib.bk_create(ctypes.byref(ctx), 0, 0)
    assert st in (bk.BK_ERR_HIP, bk.BK_ERR_INVALID_ARG)
    assert lib.bk_last_error().decode()
    assert ctx.value is None


def test_null_arguments_are_rejected(bk):
    lib = bk.load_library()
    assert lib.bk_create(None, 0, 0) == bk.BK_ERR_INVALID_ARG
    assert "null" in lib.bk_last_error().decode()
    assert lib.bk_set_stream(None, None) == bk.BK_ERR_INVALID_ARG
    n = ctypes.c_size_t()
    assert lib.bk_size(None, 0, ctypes.byref(n)) == bk.BK_ERR_INVALID_ARG
    assert lib.bk_destroy(None) == bk.BK_OK
    d = bk.KernelDesc(bk.BK_GAUSSIAN, 2, 0.1)
    out = np.zeros(1)
    assert lib.bk_eval_range(None, ctypes.byref(d), 0, 0, 0, 0, out.ctypes.data_as(bk._DBL_P), 0) == \
        bk.BK_ERR_INVALID_ARG
    assert lib.bk_create(ctypes.byref(ctypes.c_void_p()), 0, 7) == bk.BK_ERR_INVALID_ARG
    # the device sort checks its arguments before touching a device
    assert lib.bk_lexico_order(0, 17, 4, None, None) == bk.BK_ERR_INVALID_ARG
    assert lib.bk_lexico_order(0, 2, 4, None, None) == bk.BK_ERR_INVALID_ARG
    assert lib.bk_lexico_order(0, 2, 0, None, None) == bk.BK_OK  # nothing to order


def test_python_view_raises_without_library(monkeypatch, tmp_path):
    import bbrcit_kde_amd as pkg

    monkeypatch.setattr(pkg, "LIB_PATH", tmp_path / "missing.so")
    monkeypatch.setattr(pkg, "_LIB", None)
    with pytest.raises(RuntimeError, match="missing"):
        pkg.load_library()


def test_kd_order_is_a_compact_permutation():
    from bbrcit_kde_amd import datasets as ds

    pts, _ = ds.gaussian_mixture(20000, seed=11)
    perm = ds.kd_order(pts, leaf_nmax=64)
    assert np.array_equal(np.sort(perm), np.arange(len(pts)))
    # consecutive blocks of 256 points are far more compact than random ones
    blk = pts[perm][: (len(pts) // 256) * 256].reshape(-1, 256, 2)
    spread_kd = np.median(blk.max(1) - blk.min(1))
    rnd = pts[: (len(pts) // 256) * 256].reshape(-1, 256, 2)
    spread_rnd = np.median(rnd.max(1) - rnd.min(1))
    assert spread_kd < 0.25 * spread_rnd


def test_tune_cache_file_is_versioned(bk, tmp_path):
    """bk_tune_set_cache loads entries from a file of this version only (no GPU call)."""
    good = tmp_path / "tune.txt"
    good.write_text("bbrcit-hip-tune 1 1\ncvgrid/gfx950/256cu/kind4/fixed/grid8x8/sets100\t0 1\n"
                    "cvgrid/gfx950/256cu/kind3/fixed/grid4x4/sets100,200\t1 0\nmalformed line without tab\n")
    try:
        bk.tune_set_cache(good)
        assert bk.tune_entry_count() == 2
        old = tmp_path / "old.txt"
        old.write_text("bbrcit-hip-tune 1 0\ncvgrid/x\t0 0 0\n")
        bk.tune_set_cache(old)
        assert bk.tune_entry_count() == 0
        bk.tune_set_cache(tmp_path / "missing.txt")
        assert bk.tune_entry_count() == 0
    finally:
        bk.tune_set_cache(None)
    assert bk.tune_entry_count() == 0
