"""ctypes binding of the reference C benchmark compiled from its own source
(oracle/_ref/libasgref.so, built by ``make -C oracle ref`` from
adaptive_sparse_grid_bench.c with main() renamed).

TEST INFRASTRUCTURE ONLY: used to write golden fixtures in the container that holds the
reference sources, to pin the oracle against the reference itself, and as the timed CPU
baseline (the reference's own omp_hierarchize3 / omp_evaluate5) in ``bench.py``.

The reference keeps its grid in process globals (asg.c:77-85) and ``init()`` allocates
without freeing (asg.c:101-184), so one process can hold one shape at a time; calling
``setup`` again re-initialises (leaking the previous tables, as the reference does).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_ref", "libasgref.so")

_f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")

HIERARCHIZE = ["hierarchize0", "hierarchize1", "hierarchize2", "hierarchize3",
               "hierarchize4", "hierarchize5", "hierarchize6", "hierarchize7",
               "hierarchize8", "omp_hierarchize3"]
EVALUATE = ["evaluate0", "evaluate1", "evaluate2", "evaluate3", "evaluate4",
            "evaluate5", "omp_evaluate2", "omp_evaluate5"]


def available() -> bool:
    return os.path.exists(LIB_PATH)


class RefGrid:
    """The reference's global grid state for one (d, n, limits)."""

    def __init__(self, d: int, n: int, limits=None):
        self.lib = C.CDLL(LIB_PATH)
        lib = self.lib
        for name in HIERARCHIZE:
            getattr(lib, name).restype = C.c_int
            getattr(lib, name).argtypes = []
        for name in EVALUATE:
            fn = getattr(lib, name)
            fn.restype = C.c_int
            fn.argtypes = [C.c_void_p, _f32p, _f32p, C.c_int]
        lib.gp2idx.argtypes = [_i32p, _i32p]
        lib.gp2idx.restype = C.c_int
        lib.idx2gp.argtypes = [C.c_int, _i32p, _i32p]
        lib.idx2gp.restype = C.c_int
        lib.idx2l.argtypes = [C.c_int, C.c_int, _i32p]
        lib.idx2l.restype = C.c_int
        lib.l2idx.argtypes = [_i32p]
        lib.l2idx.restype = C.c_int
        lib.init.argtypes = []
        lib.init.restype = None
        self.d, self.n = d, n
        self._limits = np.full(d, n - 1, np.int32) if limits is None else np.array(limits, np.int32)
        C.c_int.in_dll(lib, "num_dims").value = d
        C.c_int.in_dll(lib, "num_levels").value = n
        C.c_void_p.in_dll(lib, "limits").value = self._limits.ctypes.data
        lib.init()
        self.N = C.c_int.in_dll(lib, "num_grid_points").value
        self.S = C.c_int.in_dll(lib, "num_subspaces").value
        self._sg1d_ptr = C.POINTER(C.c_float).in_dll(lib, "sg1d")
        dpp = C.POINTER(C.POINTER(C.c_int)).in_dll(lib, "dp_mat")
        lip = C.POINTER(C.c_int).in_dll(lib, "lidx_vect")
        self.dp = np.array([[dpp[i][s] for s in range(n)] for i in range(d)], np.int64)
        self.lidx = np.array([lip[s] for s in range(n)], np.int64)

    def omp_threads(self) -> int:
        """OpenMP threads the reference's omp_* functions run with in this process."""
        fn = self.lib.omp_get_max_threads
        fn.restype = C.c_int
        fn.argtypes = []
        return int(fn())

    # the reference grid values (float *sg1d, asg.c:79)
    def get_values(self) -> np.ndarray:
        return np.ctypeslib.as_array(self._sg1d_ptr, shape=(self.N,)).copy()

    def set_values(self, v) -> None:
        arr = np.ctypeslib.as_array(self._sg1d_ptr, shape=(self.N,))
        arr[:] = np.asarray(v, np.float32)

    def sg1d_address(self) -> int:
        return C.cast(self._sg1d_ptr, C.c_void_p).value

    def hierarchize(self, variant: str = "hierarchize0") -> None:
        getattr(self.lib, variant)()

    def evaluate(self, coords: np.ndarray, variant: str = "evaluate0") -> np.ndarray:
        x = np.ascontiguousarray(coords, np.float32)
        P = x.shape[0]
        out = np.zeros(P, np.float32)
        getattr(self.lib, variant)(self.sg1d_address(), x, out, P)
        return out

    def idx2gp(self, offset: int):
        lv = np.zeros(self.d, np.int32)
        iv = np.zeros(self.d, np.int32)
        self.lib.idx2gp(offset, lv, iv)
        return lv, iv

    def gp2idx(self, levels, indices) -> int:
        return int(self.lib.gp2idx(np.ascontiguousarray(levels, np.int32),
                                   np.ascontiguousarray(indices, np.int32)))

    def idx2l(self, rank: int, level_sum: int):
        lv = np.zeros(self.d, np.int32)
        self.lib.idx2l(rank, level_sum, lv)
        return lv

    def l2idx(self, levels) -> int:
        return int(self.lib.l2idx(np.ascontiguousarray(levels, np.int32)))
