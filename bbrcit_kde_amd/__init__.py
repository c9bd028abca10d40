"""bbrcit_kde_amd: MI355X-native direct kernel-density sums.

This package is a thin Python view of the C ABI in ``include/bbrcit_hip.h``
(the library ``lib/libbbrcit_hip.so``).  The product interface is the C ABI
and the header-only C++ API under ``include/`` of this package; this module
exists so tests and benchmarks can drive the same entry points from Python.

:class:`DirectKde` mirrors the reference's ``CudaDirectKde``
(include/CudaDirectKde.h:28-123): construct it over reference and query
points, then evaluate closed index ranges.  Range errors raise the Python
counterparts of the reference's exceptions: ``IndexError`` for
``std::out_of_range`` and ``ValueError`` for ``std::invalid_argument``.

There is no CPU fallback: if the library is missing, loading it raises.
"""
from __future__ import annotations

import ctypes
from pathlib import Path
from typing import Iterable, Optional, Sequence

import numpy as np

from .build import LIB_PATH

__all__ = [
    "BK_EPANECHNIKOV", "BK_GAUSSIAN", "BK_EPANECHNIKOV_PRODUCT_2D", "BK_GAUSSIAN_PRODUCT_2D",
    "BK_EPANECHNIKOV_CONV_1D", "BK_EPANECHNIKOV_PRODUCT_CONV_2D", "BK_GAUSSIAN_CONV_1D",
    "BK_GAUSSIAN_PRODUCT_CONV_2D", "BK_PREC_F32", "BK_PREC_F64", "BkError", "KernelDesc",
    "DirectKde", "CvGrid", "Surrogate", "load_library", "library_path", "lexico_order", "tune_set_cache",
    "tune_entry_count",
]

# ---- constants (mirror include/bbrcit_hip.h) ---------------------------------
BK_OK = 0
BK_ERR_INVALID_ARG = -1
BK_ERR_OUT_OF_RANGE = -2
BK_ERR_HIP = -3
BK_ERR_STATE = -4
BK_ERR_UNSUPPORTED = -5
BK_ERR_ALLOC = -6

BK_OPT_GAUSS_FORM = 1
BK_GAUSS_FORM_AUTO = 0
BK_GAUSS_FORM_DIRECT = 1
BK_OPT_HIGH_DIM_FORM = 2
BK_HIGH_DIM_AUTO = 0
BK_HIGH_DIM_DIRECT = 1
BK_OPT_QUERY_ORDER = 3
BK_QUERY_ORDER_GIVEN = 0
BK_QUERY_ORDER_SPATIAL = 1
BK_OPT_COUNT_TERMS = 4

BK_MLP_NUM_PARAMS = 33537

BK_CVG_VARIANT = 1
BK_CVG_GRAPH = 2
BK_CVG_COUNT_UNITS = 3
BK_CVG_TUNE_SOURCE = 4
BK_CVG_UNROLL = 5
BK_CVG_WORKGROUPS = 6
BK_CVG_VARIANTS = 3
BK_CVG_TUNE_MAX = 9

BK_SET_REF = 0
BK_SET_QUERY = 1
BK_PREC_F32 = 0
BK_PREC_F64 = 1

BK_EPANECHNIKOV = 1
BK_GAUSSIAN = 2
BK_EPANECHNIKOV_PRODUCT_2D = 3
BK_GAUSSIAN_PRODUCT_2D = 4
BK_EPANECHNIKOV_CONV_1D = 5
BK_EPANECHNIKOV_PRODUCT_CONV_2D = 6
BK_GAUSSIAN_CONV_1D = 7
BK_GAUSSIAN_PRODUCT_CONV_2D = 8

KIND_NAMES = {
    BK_EPANECHNIKOV: "EpanechnikovKernel",
    BK_GAUSSIAN: "GaussianKernel",
    BK_EPANECHNIKOV_PRODUCT_2D: "EpanechnikovProductKernel2d",
    BK_GAUSSIAN_PRODUCT_2D: "GaussianProductKernel2d",
    BK_EPANECHNIKOV_CONV_1D: "EpanechnikovConvKernel1d",
    BK_EPANECHNIKOV_PRODUCT_CONV_2D: "EpanechnikovProductConvKernel2d",
    BK_GAUSSIAN_CONV_1D: "GaussianConvKernel1d",
    BK_GAUSSIAN_PRODUCT_CONV_2D: "GaussianProductConvKernel2d",
}


class KernelDesc(ctypes.Structure):
    """bk_kernel_desc: kernel kind, dimension and bandwidth(s)."""

    _fields_ = [("kind", ctypes.c_int), ("dim", ctypes.c_int), ("hx", ctypes.c_double), ("hy", ctypes.c_double)]

    def __init__(self, kind: int, dim: int, hx: float, hy: Optional[float] = None):
        super().__init__(int(kind), int(dim), float(hx), float(hx if hy is None else hy))

    def __repr__(self) -> str:
        return f"KernelDesc({KIND_NAMES.get(self.kind, self.kind)}, dim={self.dim}, hx={self.hx}, hy={self.hy})"


class RangePair(ctypes.Structure):
    """bk_range_pair: closed reference and query index ranges."""

    _fields_ = [("r_i", ctypes.c_size_t), ("r_j", ctypes.c_size_t), ("q_i", ctypes.c_size_t), ("q_j", ctypes.c_size_t)]


class BkError(RuntimeError):
    """A non-zero status from the C ABI."""

    def __init__(self, code: int, message: str):
        super().__init__(f"[{code}] {message}")
        self.code = code


_LIB: Optional[ctypes.CDLL] = None

_DBL_P = ctypes.POINTER(ctypes.c_double)
_FLT_P = ctypes.POINTER(ctypes.c_float)


class ServeStats(ctypes.Structure):
    _fields_ = [("n", ctypes.c_size_t), ("n_samples", ctypes.c_size_t), ("n_fallback", ctypes.c_size_t),
                ("cells", ctypes.c_size_t), ("cells_failed", ctypes.c_size_t), ("max_sample_err", ctypes.c_double),
                ("ms_infer", ctypes.c_double), ("ms_check", ctypes.c_double), ("ms_fallback", ctypes.c_double),
                ("ms_total", ctypes.c_double)]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


class BoundStats(ctypes.Structure):
    _fields_ = [("excluded", ctypes.c_ulonglong), ("approximated", ctypes.c_ulonglong),
                ("evaluated", ctypes.c_ulonglong)]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


def library_path() -> Path:
    return LIB_PATH


def load_library() -> ctypes.CDLL:
    """Load libbbrcit_hip.so (raises if it has not been built)."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not LIB_PATH.exists():
        raise RuntimeError(f"{LIB_PATH} is missing: run `python -m bbrcit_kde_amd.build` first")
    lib = ctypes.CDLL(str(LIB_PATH))
    c_ctx_p = ctypes.c_void_p
    sig = {
        "bk_create": (ctypes.c_int, [ctypes.POINTER(c_ctx_p), ctypes.c_int, ctypes.c_int]),
        "bk_destroy": (ctypes.c_int, [c_ctx_p]),
        "bk_last_error": (ctypes.c_char_p, []),
        "bk_set_stream": (ctypes.c_int, [c_ctx_p, ctypes.c_void_p]),
        "bk_set_option": (ctypes.c_int, [c_ctx_p, ctypes.c_int, ctypes.c_int]),
        "bk_upload": (ctypes.c_int, [c_ctx_p, ctypes.c_int, ctypes.c_int, ctypes.c_size_t, _DBL_P, _DBL_P, _DBL_P]),
        "bk_alias_queries_to_refs": (ctypes.c_int, [c_ctx_p]),
        "bk_size": (ctypes.c_int, [c_ctx_p, ctypes.c_int, ctypes.POINTER(ctypes.c_size_t)]),
        "bk_update_refs": (ctypes.c_int, [c_ctx_p, _DBL_P, _DBL_P]),
        "bk_lexico_order": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_size_t, _DBL_P,
                                           ctypes.POINTER(ctypes.c_longlong)]),
        "bk_eval_range": (ctypes.c_int, [c_ctx_p, ctypes.POINTER(KernelDesc), ctypes.c_size_t, ctypes.c_size_t,
                                         ctypes.c_size_t, ctypes.c_size_t, _DBL_P, ctypes.c_int]),
        "bk_eval_range_device": (ctypes.c_int, [c_ctx_p, ctypes.POINTER(KernelDesc), ctypes.c_size_t,
                                                ctypes.c_size_t, ctypes.c_size_t, ctypes.c_size_t, ctypes.c_void_p,
                                                ctypes.c_int]),
        "bk_term_count": (ctypes.c_int, [c_ctx_p, ctypes.POINTER(ctypes.c_ulonglong),
                                         ctypes.POINTER(ctypes.c_ulonglong)]),
        "bk_eval_range_device_noalloc": (ctypes.c_int, [c_ctx_p, ctypes.POINTER(KernelDesc), ctypes.c_size_t,
                                                        ctypes.c_size_t, ctypes.c_size_t, ctypes.c_size_t,
                                                        ctypes.c_void_p, ctypes.c_int]),
        "bk_eval_leaf_pairs": (ctypes.c_int, [c_ctx_p, ctypes.POINTER(KernelDesc), ctypes.POINTER(RangePair),
                                              ctypes.c_size_t, _DBL_P, ctypes.c_int]),
        "bk_eval_bounded": (ctypes.c_int, [c_ctx_p, ctypes.POINTER(KernelDesc), ctypes.c_size_t, ctypes.c_size_t,
                                           ctypes.c_double, ctypes.c_double, _DBL_P, _DBL_P, ctypes.c_int,
                                           ctypes.POINTER(BoundStats)]),
        "bk_eval_bounded_device": (ctypes.c_int, [c_ctx_p, ctypes.POINTER(KernelDesc), ctypes.c_size_t,
                                                  ctypes.c_size_t, ctypes.c_double, ctypes.c_double, ctypes.c_void_p,
                                                  ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(BoundStats)]),
        "bk_self_eval": (ctypes.c_int, [c_ctx_p, ctypes.POINTER(KernelDesc), ctypes.c_int, _DBL_P, ctypes.c_int]),
        "bk_likelihood_cv": (ctypes.c_int, [c_ctx_p, ctypes.POINTER(KernelDesc), _DBL_P, _DBL_P]),
        "bk_lsq_conv_cv": (ctypes.c_int, [c_ctx_p, ctypes.POINTER(KernelDesc), _DBL_P, _DBL_P]),
        "bk_adapt": (ctypes.c_int, [c_ctx_p, ctypes.POINTER(KernelDesc), ctypes.c_double, _DBL_P, _DBL_P, _DBL_P]),
        "bk_cv_grid_create": (ctypes.c_int, [c_ctx_p, ctypes.c_int, ctypes.c_size_t,
                                             ctypes.POINTER(ctypes.c_size_t), _DBL_P, _DBL_P, _DBL_P,
                                             ctypes.POINTER(ctypes.c_void_p)]),
        "bk_cv_grid_destroy": (ctypes.c_int, [ctypes.c_void_p]),
        "bk_cv_grid_run": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_size_t, _DBL_P, ctypes.c_size_t, _DBL_P,
                                          _DBL_P, _DBL_P]),
        "bk_cv_grid_loo": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_size_t, ctypes.c_size_t, ctypes.c_size_t,
                                          _DBL_P]),
        "bk_cv_grid_set_option": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]),
        "bk_cv_grid_get_option": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_int)]),
        "bk_cv_grid_last_timing": (ctypes.c_int, [ctypes.c_void_p, _DBL_P, _DBL_P,
                                                  ctypes.POINTER(ctypes.c_longlong)]),
        "bk_cv_grid_unit_count": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_ulonglong)]),
        "bk_tune_set_cache": (ctypes.c_int, [ctypes.c_char_p]),
        "bk_cv_grid_tune": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_size_t, _DBL_P, ctypes.c_size_t, _DBL_P,
                                           _DBL_P, ctypes.POINTER(ctypes.c_int)]),
        "bk_tune_entry_count": (ctypes.c_int, [ctypes.POINTER(ctypes.c_size_t)]),
        "bk_mlp_create": (ctypes.c_int, [c_ctx_p, ctypes.c_ulonglong, ctypes.POINTER(ctypes.c_void_p)]),
        "bk_mlp_destroy": (ctypes.c_int, [ctypes.c_void_p]),
        "bk_mlp_set_params": (ctypes.c_int, [ctypes.c_void_p, _FLT_P, ctypes.c_double, ctypes.c_double,
                                             ctypes.c_double, ctypes.c_double, ctypes.c_double]),
        "bk_mlp_get_params": (ctypes.c_int, [ctypes.c_void_p, _FLT_P, _DBL_P, _DBL_P, _DBL_P, _DBL_P, _DBL_P]),
        "bk_mlp_save": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_char_p]),
        "bk_mlp_load": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_char_p]),
        "bk_mlp_set_training_data": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_size_t, _DBL_P, _DBL_P,
                                                    ctypes.c_double, ctypes.c_double, ctypes.c_double]),
        "bk_mlp_train": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_size_t, ctypes.c_size_t, ctypes.c_double,
                                        ctypes.c_double, ctypes.c_ulonglong, _DBL_P, _DBL_P]),
        "bk_mlp_loss_grad": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_size_t, _DBL_P, _DBL_P, _DBL_P, _FLT_P]),
        "bk_mlp_forward": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_size_t, _DBL_P, _DBL_P]),
        "bk_mlp_infer_device": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p, ctypes.c_void_p,
                                               ctypes.c_void_p]),
        "bk_mlp_serve_device": (ctypes.c_int, [ctypes.c_void_p, c_ctx_p, ctypes.POINTER(KernelDesc),
                                               ctypes.c_size_t, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                               ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                               ctypes.c_ulonglong, ctypes.POINTER(ServeStats)]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _LIB = lib
    return lib


def _check(status: int) -> None:
    if status == BK_OK:
        return
    msg = load_library().bk_last_error().decode(errors="replace")
    if status == BK_ERR_OUT_OF_RANGE:
        raise IndexError(msg)
    if status == BK_ERR_INVALID_ARG:
        raise ValueError(msg)
    raise BkError(status, msg)


def tune_set_cache(path) -> None:
    """Name the autotuner cache file (loaded now, saved after every new
    choice); None keeps the cache in memory only.  Drops the entries in memory."""
    lib = load_library()
    _check(lib.bk_tune_set_cache(None if path is None else str(path).encode()))


def tune_entry_count() -> int:
    """Entries in the in-memory autotuner cache."""
    n = ctypes.c_size_t()
    _check(load_library().bk_tune_entry_count(ctypes.byref(n)))
    return int(n.value)


def _dptr(a: Optional[np.ndarray]):
    if a is None:
        return None
    return a.ctypes.data_as(_DBL_P)


def _as_points(coords) -> np.ndarray:
    c = np.ascontiguousarray(np.asarray(coords, dtype=np.float64))
    if c.ndim == 1:
        c = c[:, None]
    if c.ndim != 2:
        raise ValueError("coordinates must be an (n, D) array")
    return c


def lexico_order(coords, device: int = 0) -> np.ndarray:
    """Indices of the points in lexicographic order (ExactLexicoLess, ties in
    index order), sorted on the device: bk_lexico_order."""
    c = _as_points(coords)
    order = np.empty(len(c), dtype=np.int64)
    if len(c):
        _check(load_library().bk_lexico_order(int(device), c.shape[1], len(c), _dptr(c),
                                              order.ctypes.data_as(ctypes.POINTER(ctypes.c_longlong))))
    return order


def _opt_vec(v, n: int) -> Optional[np.ndarray]:
    if v is None:
        return None
    a = np.ascontiguousarray(np.asarray(v, dtype=np.float64))
    if a.shape != (n,):
        raise ValueError(f"expected {n} values, got shape {a.shape}")
    return a


class DirectKde:
    """Direct kernel-density evaluator on one GPU (mirror of CudaDirectKde).

    ``refs`` and ``queries`` are (n, D) coordinate arrays; ``ref_mass`` and
    ``ref_abw`` give each reference its mass and adaptive bandwidth factor
    (default 1).  ``queries=None`` makes the query set the reference set
    itself (needed by the leave-one-out calls).
    """

    def __init__(self, refs, queries=None, ref_mass=None, ref_abw=None, *, precision: int = BK_PREC_F32,
                 device: int = 0):
        lib = load_library()
        self._lib = lib
        self._ctx = ctypes.c_void_p()
        _check(lib.bk_create(ctypes.byref(self._ctx), int(device), int(precision)))
        self.precision = precision
        r = _as_points(refs)
        self.dim = r.shape[1]
        self._upload(BK_SET_REF, r, _opt_vec(ref_mass, len(r)), _opt_vec(ref_abw, len(r)))
        if queries is None:
            _check(lib.bk_alias_queries_to_refs(self._ctx))
        else:
            q = _as_points(queries)
            self._upload(BK_SET_QUERY, q, None, None)

    def _upload(self, which: int, coords: np.ndarray, mass, abw) -> None:
        _check(self._lib.bk_upload(self._ctx, which, coords.shape[1], coords.shape[0], _dptr(coords), _dptr(mass),
                                   _dptr(abw)))

    # -- lifetime -----------------------------------------------------------
    def close(self) -> None:
        if getattr(self, "_ctx", None) is not None and self._ctx.value:
            self._lib.bk_destroy(self._ctx)
            self._ctx = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- sizes ----------------------------------------------------------------
    def reference_size(self) -> int:
        n = ctypes.c_size_t()
        _check(self._lib.bk_size(self._ctx, BK_SET_REF, ctypes.byref(n)))
        return n.value

    def query_size(self) -> int:
        n = ctypes.c_size_t()
        _check(self._lib.bk_size(self._ctx, BK_SET_QUERY, ctypes.byref(n)))
        return n.value

    def set_stream(self, stream_handle: int) -> None:
        _check(self._lib.bk_set_stream(self._ctx, ctypes.c_void_p(int(stream_handle))))

    def set_option(self, option: int, value: int) -> None:
        _check(self._lib.bk_set_option(self._ctx, int(option), int(value)))

    def update_refs(self, mass=None, abw=None) -> None:
        n = self.reference_size()
        _check(self._lib.bk_update_refs(self._ctx, _dptr(_opt_vec(mass, n)), _dptr(_opt_vec(abw, n))))

    # -- evaluation -------------------------------------------------------------
    def unnormalized_eval(self, desc: KernelDesc, r_i: int, r_j: int, q_i: int, q_j: int) -> np.ndarray:
        self._check_indices(r_i, r_j, q_i, q_j)
        out = np.empty(max(q_j - q_i + 1, 0), dtype=np.float64)
        _check(self._lib.bk_eval_range(self._ctx, ctypes.byref(desc), r_i, r_j, q_i, q_j, _dptr(out), 0))
        return out

    def eval(self, desc: KernelDesc, r_i: int, r_j: int, q_i: int, q_j: int) -> np.ndarray:
        self._check_indices(r_i, r_j, q_i, q_j)
        out = np.empty(max(q_j - q_i + 1, 0), dtype=np.float64)
        _check(self._lib.bk_eval_range(self._ctx, ctypes.byref(desc), r_i, r_j, q_i, q_j, _dptr(out), 1))
        return out

    def eval_device(self, desc: KernelDesc, r_i: int, r_j: int, q_i: int, q_j: int, out_dev_ptr: int,
                    normalize: bool = True) -> None:
        """Queue an evaluation writing doubles to device memory (no sync)."""
        _check(self._lib.bk_eval_range_device(self._ctx, ctypes.byref(desc), r_i, r_j, q_i, q_j,
                                              ctypes.c_void_p(int(out_dev_ptr)), int(bool(normalize))))

    def term_count(self) -> tuple[int, int]:
        """(lane terms, real-query terms) evaluated since counting was switched on
        (set_option(BK_OPT_COUNT_TERMS, 1)) or last read; resets them."""
        a, b = ctypes.c_ulonglong(), ctypes.c_ulonglong()
        _check(self._lib.bk_term_count(self._ctx, ctypes.byref(a), ctypes.byref(b)))
        return int(a.value), int(b.value)

    def eval_device_noalloc(self, desc: KernelDesc, r_i: int, r_j: int, q_i: int, q_j: int, out_dev_ptr: int,
                            normalize: bool = True) -> None:
        """As eval_device, through the entry point that allocates nothing."""
        _check(self._lib.bk_eval_range_device_noalloc(self._ctx, ctypes.byref(desc), r_i, r_j, q_i, q_j,
                                                      ctypes.c_void_p(int(out_dev_ptr)), int(bool(normalize))))

    def eval_leaf_pairs(self, desc: KernelDesc, pairs: Sequence[Sequence[int]], normalize: bool = False) -> list:
        """Evaluate many (r_i, r_j, q_i, q_j) pairs in one launch; one array per pair."""
        arr = (RangePair * len(pairs))(*[RangePair(*map(int, p)) for p in pairs])
        sizes = [int(p[3]) - int(p[2]) + 1 for p in pairs]
        out = np.empty(max(sum(sizes), 0), dtype=np.float64)
        _check(self._lib.bk_eval_leaf_pairs(self._ctx, ctypes.byref(desc), arr, len(pairs), _dptr(out),
                                            int(bool(normalize))))
        res, off = [], 0
        for s in sizes:
            res.append(out[off:off + s])
            off += s
        return res

    def eval_bounded(self, desc: KernelDesc, q_i: int, q_j: int, rel_err: float, abs_err: float,
                     normalize: bool = True):
        """Dual-tree bounds over every reference: (lower, upper, stats dict)."""
        self._check_indices(q_i, q_j)
        n = max(q_j - q_i + 1, 0)
        lo = np.empty(max(n, 1))
        up = np.empty(max(n, 1))
        st = BoundStats()
        _check(self._lib.bk_eval_bounded(self._ctx, ctypes.byref(desc), q_i, q_j, rel_err, abs_err, _dptr(lo),
                                         _dptr(up), int(normalize), ctypes.byref(st)))
        return lo[:n], up[:n], st.as_dict()

    def eval_bounded_device(self, desc: KernelDesc, q_i: int, q_j: int, rel_err: float, abs_err: float,
                            lower_dev_ptr: int, upper_dev_ptr: int, normalize: bool = True) -> None:
        self._check_indices(q_i, q_j)
        _check(self._lib.bk_eval_bounded_device(self._ctx, ctypes.byref(desc), q_i, q_j, rel_err, abs_err,
                                                ctypes.c_void_p(lower_dev_ptr), ctypes.c_void_p(upper_dev_ptr),
                                                int(normalize), None))

    def self_eval(self, desc: KernelDesc, exclude_self: bool = False, normalize: bool = True) -> np.ndarray:
        out = np.empty(self.reference_size(), dtype=np.float64)
        _check(self._lib.bk_self_eval(self._ctx, ctypes.byref(desc), int(bool(exclude_self)), _dptr(out),
                                      int(bool(normalize))))
        return out

    def likelihood_cv(self, desc: KernelDesc, weights) -> float:
        w = _opt_vec(weights, self.reference_size())
        s = ctypes.c_double()
        _check(self._lib.bk_likelihood_cv(self._ctx, ctypes.byref(desc), _dptr(w), ctypes.byref(s)))
        return s.value

    def lsq_conv_cv(self, desc: KernelDesc, weights) -> float:
        w = _opt_vec(weights, self.reference_size())
        s = ctypes.c_double()
        _check(self._lib.bk_lsq_conv_cv(self._ctx, ctypes.byref(desc), _dptr(w), ctypes.byref(s)))
        return s.value

    def adapt(self, desc: KernelDesc, alpha: float, weights):
        """Adaptive pass; returns (g, lambda) and leaves the refs adapted."""
        n = self.reference_size()
        w = _opt_vec(weights, n)
        lam = np.empty(n, dtype=np.float64)
        g = ctypes.c_double()
        _check(self._lib.bk_adapt(self._ctx, ctypes.byref(desc), float(alpha), _dptr(w), ctypes.byref(g), _dptr(lam)))
        return g.value, lam

    @staticmethod
    def _check_indices(*idx: int) -> None:
        for v in idx:
            if int(v) < 0:
                raise IndexError("indices must be non-negative")


class CvGrid:
    """Batched cross-validation over datasets x a bandwidth grid.

    Mirror of the bandwidth loop of the reference's grid_search tool
    (grid_search/grid_search.cc:366-402) as one device pipeline
    (``bk_cv_grid_*``).  ``datasets`` is a sequence of (n_d, 2) point arrays,
    ``weights`` the matching per-point weights (used as given) and ``abw``
    optional adaptive bandwidth corrections.  :meth:`run` returns
    ``(lcv, lscv)`` arrays of shape (n_sets, n_hx, n_hy).
    """

    def __init__(self, kind: int, datasets, weights, abw=None, *, device: int = 0):
        lib = load_library()
        self._lib = lib
        self._ctx = ctypes.c_void_p()
        self._plan = ctypes.c_void_p()
        _check(lib.bk_create(ctypes.byref(self._ctx), int(device), BK_PREC_F32))
        pts = [_as_points(d) for d in datasets]
        self.sizes = [len(p) for p in pts]
        coords = np.ascontiguousarray(np.concatenate(pts, axis=0)) if pts else np.zeros((0, 2))
        n = len(coords)
        w = np.ascontiguousarray(np.concatenate([np.asarray(x, dtype=np.float64) for x in weights]))
        if w.shape != (n,):
            raise ValueError("weights must match the datasets")
        a = None
        if abw is not None:
            a = np.ascontiguousarray(np.concatenate([np.asarray(x, dtype=np.float64) for x in abw]))
            if a.shape != (n,):
                raise ValueError("abw must match the datasets")
        sizes = (ctypes.c_size_t * max(1, len(self.sizes)))(*self.sizes)
        try:
            _check(lib.bk_cv_grid_create(self._ctx, int(kind), len(self.sizes), sizes, _dptr(coords), _dptr(w),
                                         _dptr(a), ctypes.byref(self._plan)))
        except Exception:
            self.close()
            raise
        self.kind = kind

    def close(self) -> None:
        if getattr(self, "_plan", None) is not None and self._plan.value:
            self._lib.bk_cv_grid_destroy(self._plan)
            self._plan = ctypes.c_void_p()
        if getattr(self, "_ctx", None) is not None and self._ctx.value:
            self._lib.bk_destroy(self._ctx)
            self._ctx = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def set_option(self, option: int, value: int) -> None:
        _check(self._lib.bk_cv_grid_set_option(self._plan, int(option), int(value)))

    def get_option(self, option: int) -> int:
        v = ctypes.c_int()
        _check(self._lib.bk_cv_grid_get_option(self._plan, int(option), ctypes.byref(v)))
        return v.value

    def run(self, hx, hy):
        hx = np.ascontiguousarray(np.asarray(hx, dtype=np.float64).ravel())
        hy = np.ascontiguousarray(np.asarray(hy, dtype=np.float64).ravel())
        shape = (len(self.sizes), len(hx), len(hy))
        lcv = np.empty(shape, dtype=np.float64)
        lscv = np.empty(shape, dtype=np.float64)
        _check(self._lib.bk_cv_grid_run(self._plan, len(hx), _dptr(hx), len(hy), _dptr(hy), _dptr(lcv),
                                        _dptr(lscv)))
        return lcv, lscv

    def loo(self, set_index: int, ix: int, iy: int) -> np.ndarray:
        """Normalised leave-one-out densities of one dataset and bandwidth pair."""
        if min(set_index, ix, iy) < 0:
            raise IndexError("indices must be non-negative")
        n = self.sizes[set_index] if set_index < len(self.sizes) else 0
        out = np.empty(max(n, 1), dtype=np.float64)
        _check(self._lib.bk_cv_grid_loo(self._plan, int(set_index), int(ix), int(iy), _dptr(out)))
        return out[:n]

    def last_timing(self):
        """(pair-kernel ms or nan under graph replay, pipeline ms, rescued entries) of the last run."""
        m, t, r = ctypes.c_double(), ctypes.c_double(), ctypes.c_longlong()
        _check(self._lib.bk_cv_grid_last_timing(self._plan, ctypes.byref(m), ctypes.byref(t), ctypes.byref(r)))
        return m.value, t.value, r.value

    def tune(self, hx, hy) -> list:
        """Time every schedule of the pair kernel on whole runs of this grid and
        record the fastest in the autotuner cache: [(tiling variant, unroll
        index, pair-kernel ms), ...] in candidate order."""
        hx = np.ascontiguousarray(hx, dtype=np.float64)
        hy = np.ascontiguousarray(hy, dtype=np.float64)
        ms = np.zeros(BK_CVG_TUNE_MAX)
        n = ctypes.c_int()
        _check(self._lib.bk_cv_grid_tune(self._plan, len(hx), _dptr(hx), len(hy), _dptr(hy), _dptr(ms),
                                         ctypes.byref(n)))
        per = n.value // BK_CVG_VARIANTS
        return [(i // per, i % per, float(ms[i])) for i in range(n.value)]

    def unit_count(self) -> int:
        """(query, reference, bandwidth pair) units the pair kernel evaluated in
        the last run, which ran with set_option(BK_CVG_COUNT_UNITS, 1)."""
        u = ctypes.c_ulonglong()
        _check(self._lib.bk_cv_grid_unit_count(self._plan, ctypes.byref(u)))
        return int(u.value)


class Surrogate:
    """Learned density surrogate (``bk_mlp_*``): an MLP 2-128-128-128-1 with
    BF16 weights, trained by the library's own HIP kernels on densities of
    the exact direct sum and served with a sampled exact check."""

    def __init__(self, seed: int = 0, *, device: int = 0):
        lib = load_library()
        self._lib = lib
        self._ctx = ctypes.c_void_p()
        self._mlp = ctypes.c_void_p()
        _check(lib.bk_create(ctypes.byref(self._ctx), int(device), BK_PREC_F32))
        try:
            _check(lib.bk_mlp_create(self._ctx, int(seed), ctypes.byref(self._mlp)))
        except Exception:
            self.close()
            raise

    def close(self) -> None:
        if getattr(self, "_mlp", None) is not None and self._mlp.value:
            self._lib.bk_mlp_destroy(self._mlp)
            self._mlp = ctypes.c_void_p()
        if getattr(self, "_ctx", None) is not None and self._ctx.value:
            self._lib.bk_destroy(self._ctx)
            self._ctx = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def params(self):
        """(theta float32[BK_MLP_NUM_PARAMS], lo, hi, mu, sigma, eps)."""
        th = np.empty(BK_MLP_NUM_PARAMS, dtype=np.float32)
        vals = [ctypes.c_double() for _ in range(5)]
        _check(self._lib.bk_mlp_get_params(self._mlp, th.ctypes.data_as(_FLT_P), *[ctypes.byref(v) for v in vals]))
        return (th,) + tuple(v.value for v in vals)

    def save(self, path) -> None:
        """Checkpoint: weights, working copy, Adam state and normalisation (bk_mlp_save)."""
        _check(self._lib.bk_mlp_save(self._mlp, str(path).encode()))

    def load(self, path) -> None:
        """Restore a checkpoint written by save (bk_mlp_load)."""
        _check(self._lib.bk_mlp_load(self._mlp, str(path).encode()))

    def set_params(self, theta, lo, hi, mu, sigma, eps) -> None:
        th = np.ascontiguousarray(np.asarray(theta, dtype=np.float32))
        if th.shape != (BK_MLP_NUM_PARAMS,):
            raise ValueError(f"expected {BK_MLP_NUM_PARAMS} parameters")
        _check(self._lib.bk_mlp_set_params(self._mlp, th.ctypes.data_as(_FLT_P), float(lo), float(hi), float(mu),
                                           float(sigma), float(eps)))

    def set_training_data(self, q, f, lo: float, hi: float, eps: float) -> None:
        q = _as_points(q)
        f = _opt_vec(f, len(q))
        _check(self._lib.bk_mlp_set_training_data(self._mlp, len(q), _dptr(q), _dptr(f), float(lo), float(hi),
                                                  float(eps)))

    def train(self, steps: int, batch: int, lr_max: float, lr_min: float, seed: int = 0):
        """Returns (mean loss of the last batches, device milliseconds)."""
        loss, ms = ctypes.c_double(), ctypes.c_double()
        _check(self._lib.bk_mlp_train(self._mlp, int(steps), int(batch), float(lr_max), float(lr_min), int(seed),
                                      ctypes.byref(loss), ctypes.byref(ms)))
        return loss.value, ms.value

    def loss_grad(self, x, t):
        x = _as_points(x)
        t = _opt_vec(t, len(x))
        g = np.empty(BK_MLP_NUM_PARAMS, dtype=np.float32)
        loss = ctypes.c_double()
        _check(self._lib.bk_mlp_loss_grad(self._mlp, len(x), _dptr(x), _dptr(t), ctypes.byref(loss),
                                          g.ctypes.data_as(_FLT_P)))
        return loss.value, g

    def forward(self, x) -> np.ndarray:
        x = _as_points(x)
        y = np.empty(len(x), dtype=np.float64)
        _check(self._lib.bk_mlp_forward(self._mlp, len(x), _dptr(x), _dptr(y)))
        return y

    def infer_device(self, n: int, qx_ptr: int, qy_ptr: int, out_ptr: int) -> None:
        _check(self._lib.bk_mlp_infer_device(self._mlp, int(n), ctypes.c_void_p(qx_ptr), ctypes.c_void_p(qy_ptr),
                                             ctypes.c_void_p(out_ptr)))

    def serve_device(self, kde: "DirectKde", desc: KernelDesc, n: int, qx_ptr: int, qy_ptr: int, out_ptr: int, *,
                     sample_fraction: float = 0.01, rel_tol: float = 1e-3, floor: float, cells: int = 64,
                     seed: int = 0) -> dict:
        st = ServeStats()
        _check(self._lib.bk_mlp_serve_device(self._mlp, kde._ctx, ctypes.byref(desc), int(n), ctypes.c_void_p(qx_ptr),
                                             ctypes.c_void_p(qy_ptr), ctypes.c_void_p(out_ptr),
                                             float(sample_fraction), float(rel_tol), float(floor), int(cells),
                                             int(seed), ctypes.byref(st)))
        return st.as_dict()
