"""ctypes binding of the CPU oracle (liboracle.so, built from sg_oracle.c).

TEST INFRASTRUCTURE ONLY.  The tests, ``__graft_entry__.smoke()`` and the CPU-baseline
leg of ``bench.py`` use this module as the checker.  The product library never imports
it, and nothing here is ever measured as the GPU path.

Every function follows the reference C benchmark (adaptive_sparse_grid_bench.c, "asg.c"):
tables asg.c:101-140, gp2idx asg.c:187-207, idx2gp asg.c:210-243, idx2l/l2idx
asg.c:245-279, hierarchize0 asg.c:557-571, level-group hierarchize asg.c:660-708,
evaluate0 asg.c:1237-1274.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

_i64p = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")


def build() -> None:
    """Compile liboracle.so (and nothing else) with the Makefile beside this file."""
    subprocess.run(["make", "-s", "-C", _HERE, "all"], check=True)


def _load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB_PATH):
        build()
    lib = C.CDLL(_LIB_PATH)
    lib.sgo_tables.argtypes = [C.c_int, C.c_int, _i32p, _i64p, _i64p,
                               C.POINTER(C.c_int64), C.POINTER(C.c_int64)]
    lib.sgo_tables.restype = C.c_int
    lib.sgo_l2idx.argtypes = [C.c_int, C.c_int, _i64p, _i32p]
    lib.sgo_l2idx.restype = C.c_int64
    lib.sgo_idx2l.argtypes = [C.c_int, C.c_int, _i64p, C.c_int64, C.c_int, _i32p]
    lib.sgo_idx2l.restype = None
    lib.sgo_gp2idx.argtypes = [C.c_int, C.c_int, _i64p, _i64p, _i32p, _i32p]
    lib.sgo_gp2idx.restype = C.c_int64
    lib.sgo_idx2gp.argtypes = [C.c_int, C.c_int, _i64p, _i64p, C.c_int64, _i32p, _i32p]
    lib.sgo_idx2gp.restype = None
    lib.sgo_idx2gp_all.argtypes = [C.c_int, C.c_int, _i64p, _i64p, C.c_int64, _i32p, _i32p]
    lib.sgo_idx2gp_all.restype = None
    lib.sgo_gp2idx_many.argtypes = [C.c_int, C.c_int, _i64p, _i64p, C.c_int64, _i32p,
                                    _i32p, _i64p]
    lib.sgo_gp2idx_many.restype = None
    for sfx, fp in (("f64", _f64p), ("f32", _f32p)):
        for name in ("sgo_hierarchize0", "sgo_hierarchize"):
            fn = getattr(lib, f"{name}_{sfx}")
            fn.argtypes = [C.c_int, C.c_int, _i64p, _i64p, C.c_int64, fp]
            fn.restype = None
        fn = getattr(lib, f"sgo_evaluate_{sfx}")
        fn.argtypes = [C.c_int, C.c_int, _i64p, _i64p, fp, fp, C.c_int64, fp]
        fn.restype = None
        fn = getattr(lib, f"sgo_evaluate_brute_{sfx}")
        fn.argtypes = [C.c_int, C.c_int, _i64p, _i64p, C.c_int64, fp, fp, C.c_int64, fp]
        fn.restype = None
    _u8p = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")
    lib.sgo_x_offsets.argtypes = [C.c_int, C.c_int64, _u8p, _i64p]
    lib.sgo_x_offsets.restype = C.c_int64
    lib.sgo_x_find.argtypes = [C.c_int, C.c_int64, _u8p, _u8p]
    lib.sgo_x_find.restype = C.c_int64
    for sfx, fp in (("f64", _f64p), ("f32", _f32p)):
        fn = getattr(lib, f"sgo_x_hierarchize_{sfx}")
        fn.argtypes = [C.c_int, C.c_int64, _u8p, _i64p, fp]
        fn.restype = None
        fn = getattr(lib, f"sgo_x_evaluate_{sfx}")
        fn.argtypes = [C.c_int, C.c_int64, _u8p, _i64p, fp, fp, C.c_int64, fp]
        fn.restype = None
    _lib = lib
    return lib


@dataclass
class Shape:
    """Shape tables of one grid: dp[d][n], lidx[n+1] (lidx[n] = N), N, S."""
    d: int
    n: int
    limits: np.ndarray
    dp: np.ndarray
    lidx: np.ndarray
    N: int
    S: int


def tables(d: int, n: int, limits=None) -> Shape:
    lib = _load()
    lim = np.full(d, n - 1, np.int32) if limits is None else np.ascontiguousarray(limits, np.int32)
    dp = np.zeros(d * n, np.int64)
    lidx = np.zeros(n + 1, np.int64)
    N = C.c_int64()
    S = C.c_int64()
    if lib.sgo_tables(d, n, lim, dp, lidx, C.byref(N), C.byref(S)) != 0:
        raise ValueError("bad shape")
    return Shape(d, n, lim, dp, lidx, N.value, S.value)


def l2idx(sh: Shape, levels) -> int:
    return int(_load().sgo_l2idx(sh.d, sh.n, sh.dp, np.ascontiguousarray(levels, np.int32)))


def idx2l(sh: Shape, rank: int, level_sum: int) -> np.ndarray:
    out = np.zeros(sh.d, np.int32)
    _load().sgo_idx2l(sh.d, sh.n, sh.dp, rank, level_sum, out)
    return out


def gp2idx(sh: Shape, levels, indices) -> int:
    return int(_load().sgo_gp2idx(sh.d, sh.n, sh.dp, sh.lidx,
                                  np.ascontiguousarray(levels, np.int32),
                                  np.ascontiguousarray(indices, np.int32)))


def idx2gp(sh: Shape, offset: int):
    lv = np.zeros(sh.d, np.int32)
    iv = np.zeros(sh.d, np.int32)
    _load().sgo_idx2gp(sh.d, sh.n, sh.dp, sh.lidx, offset, lv, iv)
    return lv, iv


def idx2gp_all(sh: Shape):
    lv = np.zeros((sh.N, sh.d), np.int32)
    iv = np.zeros((sh.N, sh.d), np.int32)
    _load().sgo_idx2gp_all(sh.d, sh.n, sh.dp, sh.lidx, sh.N, lv, iv)
    return lv, iv


def gp2idx_many(sh: Shape, levels, indices) -> np.ndarray:
    lv = np.ascontiguousarray(levels, np.int32)
    iv = np.ascontiguousarray(indices, np.int32)
    out = np.zeros(lv.shape[0], np.int64)
    _load().sgo_gp2idx_many(sh.d, sh.n, sh.dp, sh.lidx, lv.shape[0], lv, iv, out)
    return out


def _sfx(dtype) -> str:
    return {np.dtype(np.float64): "f64", np.dtype(np.float32): "f32"}[np.dtype(dtype)]


def hierarchize0(sh: Shape, values: np.ndarray) -> np.ndarray:
    """hierarchize0 (asg.c:557-571) on a copy of ``values``; O(d*N*(d+n))."""
    v = np.array(values, copy=True, order="C")
    getattr(_load(), f"sgo_hierarchize0_{_sfx(v.dtype)}")(sh.d, sh.n, sh.dp, sh.lidx, sh.N, v)
    return v


def hierarchize(sh: Shape, values: np.ndarray) -> np.ndarray:
    """Level-group hierarchization (asg.c:660-708) on a copy of ``values``."""
    v = np.array(values, copy=True, order="C")
    getattr(_load(), f"sgo_hierarchize_{_sfx(v.dtype)}")(sh.d, sh.n, sh.dp, sh.lidx, sh.N, v)
    return v


def evaluate(sh: Shape, alpha: np.ndarray, coords: np.ndarray) -> np.ndarray:
    """evaluate0 (asg.c:1237-1274): reference summation order, row-major [P][d] points."""
    a = np.ascontiguousarray(alpha)
    x = np.ascontiguousarray(coords, a.dtype)
    P = x.shape[0]
    out = np.zeros(P, a.dtype)
    getattr(_load(), f"sgo_evaluate_{_sfx(a.dtype)}")(sh.d, sh.n, sh.dp, sh.lidx, a, x, P, out)
    return out


def evaluate_brute(sh: Shape, alpha: np.ndarray, coords: np.ndarray) -> np.ndarray:
    """Sum over all N grid points of alpha * prod of hats; tiny grids only."""
    a = np.ascontiguousarray(alpha)
    x = np.ascontiguousarray(coords, a.dtype)
    P = x.shape[0]
    out = np.zeros(P, a.dtype)
    getattr(_load(), f"sgo_evaluate_brute_{_sfx(a.dtype)}")(sh.d, sh.n, sh.dp, sh.lidx, sh.N,
                                                            a, x, P, out)
    return out


def grid_coords(sh: Shape, dtype=np.float64) -> np.ndarray:
    """Coordinates (2 i + 1) / 2^(l+1) of every grid point in compact order (asg.c:150)."""
    lv, iv = idx2gp_all(sh)
    return ((2 * iv.astype(np.float64) + 1) / np.exp2(lv.astype(np.float64) + 1)).astype(dtype)


# ------------------------------------------------ explicit (refined) level sets --

@dataclass
class XShape:
    """An explicit downward-closed level set in compact order with its offsets."""
    d: int
    levels: np.ndarray      # [S][d] uint8, compact order
    base: np.ndarray        # [S] int64
    N: int
    S: int


def compact_sort(levels) -> np.ndarray:
    """Sort level vectors into compact order: level sum, then l_{d-1}, l_{d-2}, ..."""
    lv = np.unique(np.ascontiguousarray(levels, np.uint8), axis=0)
    keys = [lv[:, j] for j in range(lv.shape[1])] + [lv.sum(1)]
    return np.ascontiguousarray(lv[np.lexsort(keys)])


def xshape(levels) -> XShape:
    lv = compact_sort(levels)
    S, d = lv.shape
    base = np.zeros(S, np.int64)
    N = _load().sgo_x_offsets(d, S, lv, base)
    if N < 0:
        raise ValueError("level list not in compact order")
    return XShape(d, lv, base, int(N), S)


def x_grid_coords(xs: XShape, dtype=np.float64) -> np.ndarray:
    """Grid-point coordinates of an explicit set in compact order."""
    out = np.zeros((xs.N, xs.d))
    for q in range(xs.S):
        l = xs.levels[q].astype(np.int64)
        npts = 1 << int(l.sum())
        r = np.arange(npts)
        for k in range(xs.d - 1, -1, -1):
            i = r & ((1 << int(l[k])) - 1)
            r = r >> int(l[k])
            out[xs.base[q]:xs.base[q] + npts, k] = (2 * i + 1) / 2.0 ** (l[k] + 1)
    return out.astype(dtype)


def x_hierarchize(xs: XShape, values: np.ndarray) -> np.ndarray:
    v = np.array(values, copy=True, order="C")
    getattr(_load(), f"sgo_x_hierarchize_{_sfx(v.dtype)}")(xs.d, xs.S, xs.levels, xs.base, v)
    return v


def x_evaluate(xs: XShape, alpha: np.ndarray, coords: np.ndarray) -> np.ndarray:
    a = np.ascontiguousarray(alpha)
    x = np.ascontiguousarray(coords, a.dtype)
    out = np.zeros(x.shape[0], a.dtype)
    getattr(_load(), f"sgo_x_evaluate_{_sfx(a.dtype)}")(xs.d, xs.S, xs.levels, xs.base, a, x,
                                                        x.shape[0], out)
    return out


def x_evaluate_brute(xs: XShape, alpha: np.ndarray, coords: np.ndarray) -> np.ndarray:
    """Sum over every grid point of alpha * prod of hats (numpy; tiny sets only)."""
    gx = x_grid_coords(xs)
    lv = np.zeros((xs.N, xs.d))
    for q in range(xs.S):
        npts = 1 << int(xs.levels[q].astype(np.int64).sum())
        lv[xs.base[q]:xs.base[q] + npts] = xs.levels[q]
    out = np.zeros(len(coords))
    for p, x in enumerate(np.asarray(coords, np.float64)):
        h = np.clip(1.0 - np.abs(x[None, :] * 2.0 ** (lv + 1) - gx * 2.0 ** (lv + 1)), 0, None)
        out[p] = (np.prod(h, axis=1) * alpha).sum()
    return out
