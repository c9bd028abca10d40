"""ctypes view of the oracle libraries.  TEST INFRASTRUCTURE ONLY.

Only tests/, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline leg
may import this module, and only as the checker (or as the timed CPU
baseline), never as a path of the product.

* :class:`Oracle` wraps ``oracle/liboracle.so``, the FP64 restatement in
  ``kde_oracle.c``.
* :class:`Reference` wraps ``oracle/_ref/libbbrcit_ref.so``, the reference's
  own CPU implementation compiled from its sources (see ``oracle/Makefile``);
  it is present only where the reference sources were available at build
  time.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from pathlib import Path
from typing import Optional

import numpy as np

ORACLE_DIR = Path(__file__).resolve().parent
ORACLE_LIB = ORACLE_DIR / "liboracle.so"
REF_LIB = ORACLE_DIR / "_ref" / "libbbrcit_ref.so"

_D = ctypes.c_double
_DP = ctypes.POINTER(ctypes.c_double)
_SZ = ctypes.c_size_t
_I = ctypes.c_int


def build(reference_dir: Optional[str] = None) -> None:
    """Run oracle/Makefile (liboracle.so, and _ref/ when the sources exist)."""
    env = dict(os.environ)
    if reference_dir:
        env["REF"] = reference_dir
    proc = subprocess.run(["make", "-C", str(ORACLE_DIR), "all"], capture_output=True, text=True, env=env)
    if proc.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + proc.stdout + proc.stderr)


def _p(a: Optional[np.ndarray]):
    return None if a is None else a.ctypes.data_as(_DP)


def _pts(x) -> np.ndarray:
    a = np.ascontiguousarray(np.asarray(x, dtype=np.float64))
    return a[:, None] if a.ndim == 1 else a


def _vec(v, n) -> Optional[np.ndarray]:
    if v is None:
        return None
    a = np.ascontiguousarray(np.asarray(v, dtype=np.float64))
    assert a.shape == (n,)
    return a


def default_threads() -> int:
    return max(1, min(16, os.cpu_count() or 1))


class Oracle:
    """FP64 restatement (kde_oracle.c)."""

    def __init__(self, path: Path = ORACLE_LIB):
        if not Path(path).exists():
            build()
        lib = ctypes.CDLL(str(path))
        lib.bko_normalization.restype = _D
        lib.bko_normalization.argtypes = [_I, _I, _D, _D]
        lib.bko_kernel_u.restype = _D
        lib.bko_kernel_u.argtypes = [_I, _I, _D, _D, _DP, _D]
        lib.bko_direct.restype = _I
        lib.bko_direct.argtypes = [_I, _I, _D, _D, _DP, _DP, _DP, _SZ, _DP, _SZ, _I, _DP, _I]
        lib.bko_direct_loo_at.restype = _I
        lib.bko_direct_loo_at.argtypes = [_I, _I, _D, _D, _DP, _DP, _DP, _SZ, ctypes.POINTER(ctypes.c_longlong),
                                          _SZ, _DP, _I]
        for name in ("bko_likelihood_cv", "bko_lsq_conv_cv"):
            f = getattr(lib, name)
            f.restype = _D
            f.argtypes = [_I, _I, _D, _D, _DP, _DP, _DP, _DP, _SZ, _I]
        lib.bko_adapt.restype = _D
        lib.bko_adapt.argtypes = [_I, _I, _D, _D, _DP, _DP, _SZ, _D, _DP, _I]
        self.lib = lib

    def normalization(self, kind: int, dim: int, hx: float, hy: Optional[float] = None) -> float:
        return self.lib.bko_normalization(kind, dim, hx, hx if hy is None else hy)

    def kernel_u(self, kind, dim, hx, hy, dx, a=1.0) -> float:
        d = np.ascontiguousarray(np.asarray(dx, dtype=np.float64).reshape(-1))
        return self.lib.bko_kernel_u(kind, dim, hx, hx if hy is None else hy, _p(d), a)

    def direct(self, kind, hx, hy, refs, queries, mass=None, abw=None, exclude_self=False, normalize=False,
               threads: Optional[int] = None) -> np.ndarray:
        r = _pts(refs)
        q = _pts(queries)
        assert r.shape[1] == q.shape[1]
        dim = r.shape[1]
        hy = hx if hy is None else hy
        out = np.empty(len(q), dtype=np.float64)
        rc = self.lib.bko_direct(kind, dim, hx, hy, _p(r), _p(_vec(mass, len(r))), _p(_vec(abw, len(r))), len(r),
                                 _p(q), len(q), int(bool(exclude_self)), _p(out), threads or default_threads())
        if rc != 0:
            raise RuntimeError(f"bko_direct failed ({rc})")
        if normalize:
            out *= self.normalization(kind, dim, hx, hy)
        return out

    def direct_loo_at(self, kind, hx, hy, pts, idx, mass=None, abw=None, normalize=False,
                      threads: Optional[int] = None) -> np.ndarray:
        """Leave-one-out sums of the points pts[idx] over all of pts (pair (i, i) left out)."""
        p = _pts(pts)
        ii = np.ascontiguousarray(np.asarray(idx, dtype=np.int64))
        hy = hx if hy is None else hy
        out = np.empty(len(ii), dtype=np.float64)
        rc = self.lib.bko_direct_loo_at(kind, p.shape[1], hx, hy, _p(p), _p(_vec(mass, len(p))),
                                        _p(_vec(abw, len(p))), len(p),
                                        ii.ctypes.data_as(ctypes.POINTER(ctypes.c_longlong)), len(ii), _p(out),
                                        threads or default_threads())
        if rc != 0:
            raise RuntimeError(f"bko_direct_loo_at failed ({rc})")
        if normalize:
            out *= self.normalization(kind, p.shape[1], hx, hy)
        return out

    def likelihood_cv(self, kind, hx, hy, pts, weight, mass=None, abw=None, threads=None) -> float:
        p = _pts(pts)
        n = len(p)
        return self.lib.bko_likelihood_cv(kind, p.shape[1], hx, hx if hy is None else hy, _p(p), _p(_vec(weight, n)),
                                          _p(_vec(mass if mass is not None else weight, n)), _p(_vec(abw, n)), n,
                                          threads or default_threads())

    def lsq_conv_cv(self, kind, hx, hy, pts, weight, mass=None, abw=None, threads=None) -> float:
        p = _pts(pts)
        n = len(p)
        return self.lib.bko_lsq_conv_cv(kind, p.shape[1], hx, hx if hy is None else hy, _p(p), _p(_vec(weight, n)),
                                        _p(_vec(mass if mass is not None else weight, n)), _p(_vec(abw, n)), n,
                                        threads or default_threads())

    def adapt(self, kind, hx, hy, pts, weight, alpha, threads=None):
        p = _pts(pts)
        n = len(p)
        lam = np.empty(n, dtype=np.float64)
        g = self.lib.bko_adapt(kind, p.shape[1], hx, hx if hy is None else hy, _p(p), _p(_vec(weight, n)), n, alpha,
                               _p(lam), threads or default_threads())
        return g, lam


class Reference:
    """The reference's CPU implementation (oracle/_ref), when built."""

    def __init__(self, path: Path = REF_LIB):
        if not Path(path).exists():
            raise FileNotFoundError(f"{path} not built (reference sources absent?)")
        lib = ctypes.CDLL(str(path))
        lib.ref_kernel.restype = _I
        lib.ref_kernel.argtypes = [_I, _I, _I, _D, _D, _DP, _DP, _D, _DP, _DP]
        lib.ref_direct_eval.restype = _I
        lib.ref_direct_eval.argtypes = [_I, _I, _I, _D, _D, _DP, _DP, _SZ, _I, _DP, _SZ, _DP]
        lib.ref_direct_eval_timed.restype = _I
        lib.ref_direct_eval_timed.argtypes = [_I, _I, _I, _D, _D, _DP, _DP, _SZ, _I, _DP, _SZ, _I, _DP,
                                              ctypes.POINTER(_D)]
        lib.ref_simulate.restype = _I
        lib.ref_simulate.argtypes = [_I, _I, _I, _D, _D, _DP, _DP, _SZ, _I, ctypes.c_ulonglong, _SZ, _DP]
        lib.ref_tree_points.restype = ctypes.c_long
        lib.ref_tree_points.argtypes = [_I, _I, _I, _D, _D, _DP, _DP, _SZ, _I, _DP, _DP, _DP, _DP]
        lib.ref_tree_eval.restype = ctypes.c_long
        lib.ref_tree_eval.argtypes = [_I, _I, _I, _D, _D, _DP, _DP, _SZ, _I, _DP, _SZ, _D, _D, _I, _DP, _DP, _DP]
        for name in ("ref_likelihood_cv", "ref_lsq_conv_cv"):
            f = getattr(lib, name)
            f.restype = _I
            f.argtypes = [_I, _I, _I, _D, _D, _DP, _DP, _SZ, _I, _D, _D, _DP]
        lib.ref_lsq_numint_cv.restype = _I
        lib.ref_lsq_numint_cv.argtypes = [_I, _I, _D, _D, _DP, _DP, _SZ, _I, _D, _D, _I, _D, _D, _I, _D, _D, _I, _DP]
        lib.ref_adapt.restype = ctypes.c_long
        lib.ref_adapt.argtypes = [_I, _I, _I, _D, _D, _DP, _DP, _SZ, _I, _D, _D, _D, _DP, _DP, _DP, _DP]
        self.lib = lib

    def kernel(self, kind, dim, kernel_float, hx, hy, p, q, a=1.0):
        pp = np.ascontiguousarray(np.asarray(p, dtype=np.float64).reshape(-1))
        qq = np.ascontiguousarray(np.asarray(q, dtype=np.float64).reshape(-1))
        norm, u = _D(), _D()
        rc = self.lib.ref_kernel(kind, dim, kernel_float, hx, hx if hy is None else hy, _p(pp), _p(qq), a,
                                 ctypes.byref(norm), ctypes.byref(u))
        if rc != 0:
            raise ValueError("unsupported kind/dim")
        return norm.value, u.value

    def direct_eval(self, kind, kernel_float, hx, hy, pts, w, queries, leaf_nmax=2) -> np.ndarray:
        p, q = _pts(pts), _pts(queries)
        out = np.empty(len(q))
        rc = self.lib.ref_direct_eval(kind, p.shape[1], kernel_float, hx, hx if hy is None else hy, _p(p),
                                      _p(_vec(w, len(p))), len(p), leaf_nmax, _p(q), len(q), _p(out))
        if rc != 0:
            raise RuntimeError(f"ref_direct_eval failed ({rc})")
        return out

    def direct_eval_timed(self, kind, kernel_float, hx, hy, pts, w, queries, threads=1, leaf_nmax=2):
        """direct_eval with the KernelDensity built outside the timed region and
        the queries split over `threads` threads: (values, seconds)."""
        p, q = _pts(pts), _pts(queries)
        out = np.empty(len(q))
        sec = _D()
        rc = self.lib.ref_direct_eval_timed(kind, p.shape[1], kernel_float, hx, hx if hy is None else hy, _p(p),
                                            _p(_vec(w, len(p))), len(p), leaf_nmax, _p(q), len(q), int(threads),
                                            _p(out), ctypes.byref(sec))
        if rc != 0:
            raise RuntimeError(f"ref_direct_eval_timed failed ({rc})")
        return out, sec.value

    def simulate(self, kind, kernel_float, hx, hy, pts, w, seed, n_samples, leaf_nmax=2):
        """n_samples calls of KernelDensity::simulate(e, p), e = std::mt19937_64(seed)."""
        p = _pts(pts)
        out = np.empty((n_samples, p.shape[1]))
        rc = self.lib.ref_simulate(kind, p.shape[1], kernel_float, hx, hx if hy is None else hy, _p(p),
                                   _p(_vec(w, len(p))), len(p), leaf_nmax, int(seed), int(n_samples), _p(out))
        if rc != 0:
            raise RuntimeError(f"ref_simulate failed ({rc})")
        return out

    def tree_points(self, kind, kernel_float, hx, hy, pts, w, leaf_nmax=2):
        p = _pts(pts)
        n, dim = p.shape
        c, ww, m, a = np.empty((n, dim)), np.empty(n), np.empty(n), np.empty(n)
        k = self.lib.ref_tree_points(kind, dim, kernel_float, hx, hx if hy is None else hy, _p(p), _p(_vec(w, n)), n,
                                     leaf_nmax, _p(c), _p(ww), _p(m), _p(a))
        if k < 0:
            raise RuntimeError("ref_tree_points failed")
        return c[:k], ww[:k], m[:k], a[:k]

    def tree_eval(self, kind, kernel_float, hx, hy, pts, w, queries, rel, abs_err, leaf_nmax=2, q_leaf_nmax=2):
        p, q = _pts(pts), _pts(queries)
        dim = p.shape[1]
        qo, lo, up = np.empty((len(q), dim)), np.empty(len(q)), np.empty(len(q))
        k = self.lib.ref_tree_eval(kind, dim, kernel_float, hx, hx if hy is None else hy, _p(p), _p(_vec(w, len(p))),
                                   len(p), leaf_nmax, _p(q), len(q), rel, abs_err, q_leaf_nmax, _p(qo), _p(lo), _p(up))
        if k < 0:
            raise RuntimeError("ref_tree_eval failed")
        return qo[:k], lo[:k], up[:k]

    def _cv(self, fn, kind, kernel_float, hx, hy, pts, w, rel, abs_err, leaf_nmax):
        p = _pts(pts)
        out = _D()
        rc = fn(kind, p.shape[1], kernel_float, hx, hx if hy is None else hy, _p(p), _p(_vec(w, len(p))), len(p),
                leaf_nmax, rel, abs_err, ctypes.byref(out))
        if rc != 0:
            raise RuntimeError(f"reference CV call failed ({rc})")
        return out.value

    def likelihood_cv(self, kind, kernel_float, hx, hy, pts, w, rel=1e-6, abs_err=1e-6, leaf_nmax=2):
        return self._cv(self.lib.ref_likelihood_cv, kind, kernel_float, hx, hy, pts, w, rel, abs_err, leaf_nmax)

    def lsq_conv_cv(self, kind, kernel_float, hx, hy, pts, w, rel=1e-6, abs_err=1e-6, leaf_nmax=2):
        return self._cv(self.lib.ref_lsq_conv_cv, kind, kernel_float, hx, hy, pts, w, rel, abs_err, leaf_nmax)

    def lsq_numint_cv(self, kind, kernel_float, hx, hy, pts, w, start_x, end_x, steps_x, start_y, end_y, steps_y,
                      rel=1e-6, abs_err=1e-8, leaf_nmax=2, q_leaf_nmax=2):
        p = _pts(pts)
        out = _D()
        rc = self.lib.ref_lsq_numint_cv(kind, kernel_float, hx, hx if hy is None else hy, _p(p), _p(_vec(w, len(p))),
                                        len(p), leaf_nmax, start_x, end_x, steps_x, start_y, end_y, steps_y, rel,
                                        abs_err, q_leaf_nmax, ctypes.byref(out))
        if rc != 0:
            raise RuntimeError("ref_lsq_numint_cv failed")
        return out.value

    def adapt(self, kind, kernel_float, hx, hy, pts, w, alpha, rel=1e-6, abs_err=1e-6, leaf_nmax=2):
        p = _pts(pts)
        n, dim = p.shape
        c, ww, a, m = np.empty((n, dim)), np.empty(n), np.empty(n), np.empty(n)
        k = self.lib.ref_adapt(kind, dim, kernel_float, hx, hx if hy is None else hy, _p(p), _p(_vec(w, n)), n,
                               leaf_nmax, alpha, rel, abs_err, _p(c), _p(ww), _p(a), _p(m))
        if k < 0:
            raise RuntimeError("ref_adapt failed")
        return c[:k], ww[:k], a[:k], m[:k]
