"""Python mirror of the reference benchmark's sparse grid operations over libsgrid.so.

``Grid`` carries the reference's operator interface (adaptive_sparse_grid_bench.c,
"asg.c"): ``gp2idx`` (asg.c:187), ``idx2gp`` (asg.c:210), ``idx2l`` (asg.c:245),
``l2idx`` (asg.c:266), ``hierarchize`` (asg.c:557, in place) and ``evaluate``
(asg.c:1237, ``out`` overwritten).  Device buffers are torch tensors on the GPU (torch is
only the allocator and the stream source here); every computation runs in the HIP
kernels of libsgrid.so.  The module raises at import if the library is missing: there
is no fallback path.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# SGRID_LIB overrides the library path (for comparing builds of the same sources)
LIB_PATH = os.environ.get("SGRID_LIB") or os.path.join(_HERE, "libsgrid.so")

SG_OK = 0
SG_E_INVALID = -1
SG_E_NOMEM = -2
SG_E_HIP = -3
SG_E_RANGE = -4
SG_E_UNSUPPORTED = -5

# test functions of sg_sample_* (include/sgrid.h)
SAMPLE_FUNCTIONS = {"reference": 0, "parabola": 1, "gauss": 2, "gaussian_sum": 3,
                    "sin_product": 4}

# every symbol include/sgrid.h declares
EXPORTS = [
    "sg_last_error", "sg_status_string", "sg_version",
    "sg_grid_create", "sg_grid_create_explicit", "sg_grid_destroy", "sg_grid_info",
    "sg_grid_tables", "sg_grid_subspaces", "sg_grid_get_tuning", "sg_grid_set_tuning",
    "sg_grid_walk_variant", "sg_gp2idx", "sg_idx2gp", "sg_idx2l", "sg_l2idx",
    "sg_idx2gp_dev", "sg_gp2idx_dev", "sg_grid_coords_f64",
    "sg_hierarchize_f64", "sg_hierarchize_f32",
    "sg_evaluate_f64", "sg_evaluate_f32c", "sg_evaluate_ref_f64", "sg_evaluate_ref_f32",
    "sg_evaluate_f64_noalloc", "sg_uniform_points_f64", "sg_uniform_points_f32", "sg_convert_f64_to_f32",
    "sg_gather_probe_f64", "sg_gather_probe_f32", "sg_stage_probe_f64", "sg_stage_probe_f32",
    "sg_sample_f64", "sg_sample_f32",
    "sg_file_write", "sg_file_info", "sg_file_read",
    "sg_points_to_row_major_f64", "sg_points_to_row_major_f32",
    "sg_batch_create", "sg_batch_destroy", "sg_batch_values", "sg_batch_set_points",
    "sg_batch_hierarchize_f64", "sg_batch_hierarchize_f32", "sg_batch_evaluate_f64",
    "sg_batch_evaluate_f32c", "sg_batch_run_graph_f64", "sg_batch_run_graph_f32c",
    "sg_server_create", "sg_server_destroy", "sg_server_add_grid", "sg_server_submit",
    "sg_server_poll", "sg_server_wait", "sg_server_stats", "sg_server_inject_launch_failure",
]


class Tuning(C.Structure):
    """sg_tuning (include/sgrid.h)."""
    _fields_ = [("eval_block", C.c_int), ("eval_variant_f64", C.c_int),
                ("eval_variant_f32", C.c_int), ("bin_min_points", C.c_int),
                ("hier_tile_values", C.c_int)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


EVAL_VARIANTS = 6
STAGED_VARIANT = 3          # surplus blocks staged in LDS by the workgroup
BLK_VARIANT = 4             # block-staged walk (restricted-regular grids, compile-time n)
BLK2_VARIANT = 5            # block-staged walk with two points per lane (FP64, depth-first)
STAGED_VARIANTS = (STAGED_VARIANT, BLK_VARIANT, BLK2_VARIANT)
# the block-staged walk visits its blocks in the sparse order from these dimensions on
# (sg_grid.hpp blk_sparse; tests/test_stage_plan.py checks the two agree): each leaf's
# weight is the depth-first walk's, but the blocks are summed in another order, so the
# outputs agree with the other variants within the FP64 bound, not bit for bit
BLK_SPARSE_DMIN = {"f64": 12, "f32": 15}


def blk_sparse(d: int, prec: str) -> bool:
    """Whether the block-staged walk of a d-dimensional grid in precision prec ("f64" or
    "f32") runs in the sparse order."""
    return d >= BLK_SPARSE_DMIN[prec]


class ServerConfig(C.Structure):
    """sg_server_config (include/sgrid.h)."""
    _fields_ = [("precision", C.c_int), ("slots", C.c_int), ("max_batch_points", C.c_longlong),
                ("max_requests", C.c_int), ("budget_bytes", C.c_longlong)]


class SgError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"{msg} (status {status})")
        self.status = status


def _load() -> C.CDLL:
    # torch ships its own HIP runtime under the same soname as /opt/rocm's; whichever
    # is loaded first serves the whole process.  Load torch's first so that the tensors
    # and streams it hands over belong to the runtime the library launches on.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                          "(make -C adaptive_sparse_grid_bench_amd/csrc)")
    lib = C.CDLL(LIB_PATH)
    vp, ll, i = C.c_void_p, C.c_longlong, C.c_int
    ip = C.POINTER(C.c_int)
    llp = C.POINTER(C.c_longlong)
    sig = {
        "sg_last_error": (C.c_char_p, []),
        "sg_status_string": (C.c_char_p, [i]),
        "sg_version": (i, []),
        "sg_grid_create": (i, [i, i, ip, C.POINTER(vp)]),
        "sg_grid_create_explicit": (i, [i, C.POINTER(C.c_ubyte), ll, C.POINTER(vp)]),
        "sg_grid_destroy": (i, [vp]),
        "sg_grid_info": (i, [vp, ip, ip, llp, llp]),
        "sg_grid_tables": (i, [vp, llp, llp]),
        "sg_grid_subspaces": (i, [vp, C.POINTER(C.c_ubyte), llp]),
        "sg_grid_get_tuning": (i, [vp, C.POINTER(Tuning)]),
        "sg_grid_set_tuning": (i, [vp, C.POINTER(Tuning)]),
        "sg_grid_walk_variant": (i, [vp, i, i, ip]),
        "sg_gp2idx": (i, [vp, ip, ip, llp]),
        "sg_idx2gp": (i, [vp, ll, ip, ip]),
        "sg_idx2l": (i, [vp, ll, i, ip]),
        "sg_l2idx": (i, [vp, ip, llp]),
        "sg_idx2gp_dev": (i, [vp, vp, ll, vp, vp, vp]),
        "sg_gp2idx_dev": (i, [vp, vp, vp, ll, vp, vp]),
        "sg_grid_coords_f64": (i, [vp, vp, vp]),
        "sg_hierarchize_f64": (i, [vp, vp, vp]),
        "sg_hierarchize_f32": (i, [vp, vp, vp]),
        "sg_evaluate_f64": (i, [vp, vp, vp, ll, vp, vp]),
        "sg_evaluate_f32c": (i, [vp, vp, vp, ll, vp, vp]),
        "sg_evaluate_ref_f64": (i, [vp, vp, vp, ll, vp, vp]),
        "sg_evaluate_ref_f32": (i, [vp, vp, vp, ll, vp, vp]),
        "sg_evaluate_f64_noalloc": (i, [vp, vp, vp, ll, vp, vp]),
        "sg_uniform_points_f64": (i, [ll, C.c_ulonglong, vp, vp]),
        "sg_uniform_points_f32": (i, [ll, C.c_ulonglong, vp, vp]),
        "sg_convert_f64_to_f32": (i, [vp, vp, ll, vp]),
        "sg_points_to_row_major_f64": (i, [vp, i, i, ll, vp, vp]),
        "sg_points_to_row_major_f32": (i, [vp, i, i, ll, vp, vp]),
        "sg_file_write": (i, [C.c_char_p, vp, vp, i]),
        "sg_file_info": (i, [C.c_char_p, ip, ip, llp, llp, ip]),
        "sg_file_read": (i, [C.c_char_p, C.POINTER(vp), vp, ll]),
        "sg_sample_f64": (i, [vp, i, C.POINTER(C.c_double), ll, vp, vp]),
        "sg_sample_f32": (i, [vp, i, C.POINTER(C.c_double), ll, vp, vp]),
        "sg_gather_probe_f64": (i, [vp, vp, ll, C.c_ulonglong, vp, vp]),
        "sg_gather_probe_f32": (i, [vp, vp, ll, C.c_ulonglong, vp, vp]),
        "sg_stage_probe_f64": (i, [vp, vp, ll, C.c_ulonglong, vp, vp]),
        "sg_stage_probe_f32": (i, [vp, vp, ll, C.c_ulonglong, vp, vp]),
        "sg_batch_create": (i, [C.POINTER(vp), i, C.POINTER(vp)]),
        "sg_batch_destroy": (i, [vp]),
        "sg_batch_values": (i, [vp, llp, llp]),
        "sg_batch_set_points": (i, [vp, llp, llp, llp]),
        "sg_batch_hierarchize_f64": (i, [vp, vp, vp]),
        "sg_batch_hierarchize_f32": (i, [vp, vp, vp]),
        "sg_batch_evaluate_f64": (i, [vp, vp, vp, vp, vp]),
        "sg_batch_evaluate_f32c": (i, [vp, vp, vp, vp, vp]),
        "sg_batch_run_graph_f64": (i, [vp, vp, vp, vp, i, vp]),
        "sg_batch_run_graph_f32c": (i, [vp, vp, vp, vp, i, vp]),
        "sg_server_create": (i, [C.POINTER(ServerConfig), C.POINTER(vp)]),
        "sg_server_destroy": (i, [vp]),
        "sg_server_add_grid": (i, [vp, vp, vp, ip]),
        "sg_server_submit": (i, [vp, i, vp, ll, vp, llp]),
        "sg_server_poll": (i, [vp, ll, ip]),
        "sg_server_wait": (i, [vp, ll, C.POINTER(C.c_double)]),
        "sg_server_stats": (i, [vp, llp, llp, llp, llp]),
        "sg_server_inject_launch_failure": (i, [vp, i, i]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


lib = _load()


def check(status: int) -> None:
    if status != SG_OK:
        raise SgError(status, lib.sg_last_error().decode())


def _ip(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_int))


def _ptr(t) -> int:
    """Device address of a torch tensor (or an int address, or None)."""
    if t is None:
        return None
    if isinstance(t, int):
        return t
    return t.data_ptr()


def _fp_dtype(t, name: str):
    """float64 or float32 dtype of tensor t (ValueError for anything else)."""
    import torch
    if not isinstance(t, torch.Tensor):
        raise ValueError(f"{name}: expected a torch tensor, got {type(t).__name__}")
    if t.dtype not in (torch.float64, torch.float32):
        raise ValueError(f"{name}: dtype {t.dtype}; float64 or float32 expected")
    return t.dtype


def _check(t, name: str, dtype, numel=None, shape=None, min_numel=None) -> None:
    """Validate a tensor handed to the library: dtype, C-contiguity, size.  Raises
    ValueError.  _on_device checks the devices once every tensor has passed this."""
    import torch
    if not isinstance(t, torch.Tensor):
        raise ValueError(f"{name}: expected a torch tensor, got {type(t).__name__}")
    if t.dtype != dtype:
        raise ValueError(f"{name}: dtype {t.dtype}, expected {dtype}")
    if not t.is_contiguous():
        raise ValueError(f"{name}: must be contiguous (row-major)")
    if shape is not None and tuple(t.shape) != tuple(shape):
        raise ValueError(f"{name}: shape {tuple(t.shape)}, expected {tuple(shape)}")
    if numel is not None and t.numel() != numel:
        raise ValueError(f"{name}: {t.numel()} elements, expected {numel}")
    if min_numel is not None and t.numel() < min_numel:
        raise ValueError(f"{name}: {t.numel()} elements, at least {min_numel} expected")


def _on_device(**tensors) -> None:
    for name, t in tensors.items():
        if not t.is_cuda:
            raise ValueError(f"{name}: must be a CUDA (device) tensor")


def _stream(stream) -> int:
    if stream is None:
        import torch
        return torch.cuda.current_stream().cuda_stream
    if isinstance(stream, int):
        return stream
    return stream.cuda_stream


class Grid:
    """A compact sparse grid resident on the current GPU.

    ``Grid(d, n, limits=None)`` is the restricted-regular grid
    { l : |l|_1 <= n-1, l_k <= limits[k] } (asg.c:82-83, 101-140);
    ``Grid.explicit(d, levels)`` takes a downward-closed set of level vectors.
    """

    def __init__(self, d: int, n: int, limits=None, _handle=None):
        h = C.c_void_p()
        if _handle is None:
            lim = None if limits is None else np.ascontiguousarray(limits, np.int32)
            check(lib.sg_grid_create(d, n, None if lim is None else _ip(lim), C.byref(h)))
            self._h = h
        else:
            self._h = _handle
        dd, nn = C.c_int(), C.c_int()
        NN, SS = C.c_longlong(), C.c_longlong()
        check(lib.sg_grid_info(self._h, C.byref(dd), C.byref(nn), C.byref(NN), C.byref(SS)))
        self.d, self.n, self.N, self.S = dd.value, nn.value, NN.value, SS.value

    @classmethod
    def explicit(cls, d: int, levels) -> "Grid":
        lv = np.ascontiguousarray(levels, np.uint8).reshape(-1, d)
        h = C.c_void_p()
        check(lib.sg_grid_create_explicit(d, lv.ctypes.data_as(C.POINTER(C.c_ubyte)),
                                          lv.shape[0], C.byref(h)))
        return cls(d, 0, _handle=h)

    def close(self) -> None:
        if getattr(self, "_h", None):
            lib.sg_grid_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- shape tables --------------------------------------------------------
    def tables(self):
        dp = np.zeros(self.d * self.n, np.int64)
        lidx = np.zeros(self.n + 1, np.int64)
        check(lib.sg_grid_tables(self._h, dp.ctypes.data_as(C.POINTER(C.c_longlong)),
                                 lidx.ctypes.data_as(C.POINTER(C.c_longlong))))
        return dp.reshape(self.d, self.n), lidx

    @property
    def is_explicit(self) -> bool:
        """True for an explicit (refined-subspace) level set: it has no dp matrix."""
        dp = np.zeros(max(self.d * self.n, 1), np.int64)
        st = lib.sg_grid_tables(self._h, dp.ctypes.data_as(C.POINTER(C.c_longlong)), None)
        if st == SG_E_UNSUPPORTED:
            return True
        check(st)
        return False

    @property
    def limits(self) -> np.ndarray:
        """Largest level of each dimension in the level set (the restriction vector of a
        restricted-regular grid, asg.c:82-83)."""
        lv, _ = self.subspaces()
        return lv.max(axis=0).astype(np.int32)

    def subspaces(self):
        lv = np.zeros((self.S, self.d), np.uint8)
        base = np.zeros(self.S, np.int64)
        check(lib.sg_grid_subspaces(self._h, lv.ctypes.data_as(C.POINTER(C.c_ubyte)),
                                    base.ctypes.data_as(C.POINTER(C.c_longlong))))
        return lv, base

    def get_tuning(self) -> dict:
        tu = Tuning()
        check(lib.sg_grid_get_tuning(self._h, C.byref(tu)))
        return tu.as_dict()

    def set_tuning(self, **kw) -> None:
        """Update launch tuning fields (sg_tuning names); others keep their value."""
        cur = self.get_tuning()
        unknown = set(kw) - set(cur)
        if unknown:
            raise TypeError(f"unknown tuning fields {sorted(unknown)}")
        cur.update(kw)
        check(lib.sg_grid_set_tuning(self._h, C.byref(Tuning(**cur))))

    def walk_variant(self, variant: int, prec: str = "f64") -> int:
        """The walk variant an evaluate in precision prec ("f64" or "f32") runs on this
        grid when its tuning names `variant` (a variant the grid does not cover falls
        back to a more general one)."""
        out = C.c_int()
        check(lib.sg_grid_walk_variant(self._h, 8 if prec == "f64" else 4, variant,
                                       C.byref(out)))
        return out.value

    # ---- host bijection (asg.c:187-279) --------------------------------------
    def gp2idx(self, levels, indices) -> int:
        lv = np.ascontiguousarray(levels, np.int32)
        iv = np.ascontiguousarray(indices, np.int32)
        out = C.c_longlong()
        check(lib.sg_gp2idx(self._h, _ip(lv), _ip(iv), C.byref(out)))
        return out.value

    def idx2gp(self, index: int):
        lv = np.zeros(self.d, np.int32)
        iv = np.zeros(self.d, np.int32)
        check(lib.sg_idx2gp(self._h, index, _ip(lv), _ip(iv)))
        return lv, iv

    def idx2l(self, rank: int, level_sum: int):
        lv = np.zeros(self.d, np.int32)
        check(lib.sg_idx2l(self._h, rank, level_sum, _ip(lv)))
        return lv

    def l2idx(self, levels) -> int:
        lv = np.ascontiguousarray(levels, np.int32)
        out = C.c_longlong()
        check(lib.sg_l2idx(self._h, _ip(lv), C.byref(out)))
        return out.value

    # ---- device operations ---------------------------------------------------
    def idx2gp_dev(self, levels, indices, index=None, count=None, stream=None) -> None:
        cnt = self.N if count is None else count
        check(lib.sg_idx2gp_dev(self._h, _ptr(index), cnt, _ptr(levels), _ptr(indices),
                                _stream(stream)))

    def gp2idx_dev(self, levels, indices, index, count: int, stream=None) -> None:
        check(lib.sg_gp2idx_dev(self._h, _ptr(levels), _ptr(indices), count, _ptr(index),
                                _stream(stream)))

    def grid_coords(self, coords, stream=None) -> None:
        import torch
        _check(coords, "coords", torch.float64, numel=self.N * self.d)
        _on_device(coords=coords)
        check(lib.sg_grid_coords_f64(self._h, _ptr(coords), _stream(stream)))

    def hierarchize(self, values, stream=None) -> None:
        """In place; FP64 or FP32 by the tensor's dtype (asg.c:557-571)."""
        import torch
        dt = _fp_dtype(values, "values")
        _check(values, "values", dt, numel=self.N)
        _on_device(values=values)
        fn = lib.sg_hierarchize_f64 if dt == torch.float64 else lib.sg_hierarchize_f32
        check(fn(self._h, _ptr(values), _stream(stream)))

    def sample(self, function: str, out, params=None, stream=None) -> None:
        """out[o] = function at grid point o, on the device (asg.c:143-152).  function is
        a SAMPLE_FUNCTIONS name; params as include/sgrid.h states (a flat float sequence)."""
        import torch
        dt = _fp_dtype(out, "out")
        _check(out, "out", dt, numel=self.N)
        _on_device(out=out)
        fn = lib.sg_sample_f64 if dt == torch.float64 else lib.sg_sample_f32
        p = np.ascontiguousarray(params if params is not None else [], np.float64)
        check(fn(self._h, SAMPLE_FUNCTIONS[function],
                 p.ctypes.data_as(C.POINTER(C.c_double)) if p.size else None, p.size,
                 _ptr(out), _stream(stream)))

    def evaluate(self, alpha, coords, out, stream=None, mode: str = "fast") -> None:
        """out[p] = interpolant at coords[p] (asg.c:1237-1274).

        mode "fast": FP64 walk (float64 tensors) or FP32 Kahan (float32 tensors);
        mode "ref": the reference's operation order, bit-identical to evaluate0.
        """
        import torch
        dt = _fp_dtype(alpha, "alpha")
        _check(alpha, "alpha", dt, numel=self.N)
        if not hasattr(coords, "shape") or coords.dim() != 2:
            raise ValueError("coords: expected a [P][d] tensor")
        P = coords.shape[0]
        _check(coords, "coords", dt, shape=(P, self.d))
        _check(out, "out", dt, min_numel=P)
        _on_device(alpha=alpha, coords=coords, out=out)
        f64 = dt == torch.float64
        if mode == "fast":
            fn = lib.sg_evaluate_f64 if f64 else lib.sg_evaluate_f32c
        elif mode == "ref":
            fn = lib.sg_evaluate_ref_f64 if f64 else lib.sg_evaluate_ref_f32
        else:
            raise ValueError(mode)
        check(fn(self._h, _ptr(alpha), _ptr(coords), P, _ptr(out), _stream(stream)))


def save(path: str, grid: "Grid", values=None) -> None:
    """Write a compact-grid file: the level set and, if given, the values (a float64 or
    float32 array or CPU tensor in compact order)."""
    if values is None:
        check(lib.sg_file_write(os.fsencode(path), grid._h, None, 0))
        return
    v = np.ascontiguousarray(values.cpu().numpy() if hasattr(values, "cpu") else values)
    if v.dtype not in (np.float64, np.float32) or v.shape != (grid.N,):
        raise ValueError("values must be float64 or float32 of shape (N,)")
    check(lib.sg_file_write(os.fsencode(path), grid._h, v.ctypes.data, v.itemsize))


def file_info(path: str) -> dict:
    d, n, es = C.c_int(), C.c_int(), C.c_int()
    N, S = C.c_longlong(), C.c_longlong()
    check(lib.sg_file_info(os.fsencode(path), C.byref(d), C.byref(n), C.byref(N), C.byref(S),
                           C.byref(es)))
    return {"d": d.value, "n": n.value, "N": N.value, "S": S.value, "elem_bytes": es.value}


def load(path: str):
    """Read a compact-grid file: (Grid, values as a numpy array or None)."""
    info = file_info(path)
    es = info["elem_bytes"]
    vals = np.empty(info["N"], np.float64 if es == 8 else np.float32) if es else None
    h = C.c_void_p()
    check(lib.sg_file_read(os.fsencode(path), C.byref(h),
                           vals.ctypes.data if vals is not None else None,
                           vals.nbytes if vals is not None else 0))
    return Grid(0, 0, _handle=h), vals


def gather_probe(grid: "Grid", alpha, out, seed: int = 1, stream=None) -> None:
    """The achievable-gather pattern of grid's shape for out.numel() points (one random
    in-block gather per subspace per point, no hats): the evaluate kernels' gather
    roofline.  out[p] = sum of point p's gathered values."""
    import torch
    dt = _fp_dtype(alpha, "alpha")
    _check(alpha, "alpha", dt, numel=grid.N)
    _check(out, "out", dt)
    _on_device(alpha=alpha, out=out)
    fn = lib.sg_gather_probe_f64 if dt == torch.float64 else lib.sg_gather_probe_f32
    check(fn(grid._h, _ptr(alpha), out.numel(), seed, _ptr(out), _stream(stream)))


def stage_probe(grid: "Grid", alpha, out, seed: int = 1, stream=None) -> None:
    """The staged-pattern probe of the staged walk grid runs in alpha's precision, for
    out.numel() points: the walk's stage copies and one read (LDS or global) plus one FMA
    per subspace per point, no hats and no walk; a floor for that walk's time."""
    import torch
    dt = _fp_dtype(alpha, "alpha")
    _check(alpha, "alpha", dt, numel=grid.N)
    _check(out, "out", dt)
    _on_device(alpha=alpha, out=out)
    fn = lib.sg_stage_probe_f64 if dt == torch.float64 else lib.sg_stage_probe_f32
    check(fn(grid._h, _ptr(alpha), out.numel(), seed, _ptr(out), _stream(stream)))


POINT_LAYOUTS = {"row_major": 0, "column_major": 1, "tiled32": 2}


def points_to_row_major(src, layout: str, d: int, num_points: int, out, stream=None) -> None:
    """Device copy of query points held column-major ([d][P]) or in the CUDA reference's
    32-point tiles ([ceil(P/32)][d][32], asg.cu:167-180) into row-major out [P][d]."""
    import torch
    fn = lib.sg_points_to_row_major_f64 if out.dtype == torch.float64 else \
        lib.sg_points_to_row_major_f32
    check(fn(_ptr(src), POINT_LAYOUTS[layout], d, num_points, _ptr(out), _stream(stream)))


def uniform_points(out, seed: int = 42, stream=None) -> None:
    """Fill a float64/float32 tensor with k * 2^-24 values (splitmix64 stream `seed`)."""
    import torch
    dt = _fp_dtype(out, "out")
    _check(out, "out", dt)
    _on_device(out=out)
    fn = lib.sg_uniform_points_f64 if dt == torch.float64 else lib.sg_uniform_points_f32
    check(fn(out.numel(), seed, _ptr(out), _stream(stream)))


def to_f32(src, dst, stream=None) -> None:
    import torch
    _check(src, "src", torch.float64)
    _check(dst, "dst", torch.float32, min_numel=src.numel())
    _on_device(src=src, dst=dst)
    check(lib.sg_convert_f64_to_f32(_ptr(src), _ptr(dst), src.numel(), _stream(stream)))


class Batch:
    """Independent grids resident together (sg_batch in include/sgrid.h).

    ``values`` buffers hold grid g's surpluses at ``value_offsets[g]``; after
    ``set_points(counts)`` coordinates of grid g start at element ``coord_offsets[g]``
    and its outputs at ``point_offsets[g]``.
    """

    def __init__(self, grids):
        self.grids = list(grids)                 # keep the grids alive
        arr = (C.c_void_p * len(self.grids))(*[g._h.value if isinstance(g._h, C.c_void_p)
                                                 else g._h for g in self.grids])
        h = C.c_void_p()
        check(lib.sg_batch_create(arr, len(self.grids), C.byref(h)))
        self._h = h
        total = C.c_longlong()
        offs = np.zeros(len(self.grids) + 1, np.int64)
        check(lib.sg_batch_values(self._h, C.byref(total),
                                  offs.ctypes.data_as(C.POINTER(C.c_longlong))))
        self.total_values = total.value
        self.value_offsets = offs
        self.point_offsets = None
        self.coord_offsets = None

    def close(self):
        if getattr(self, "_h", None):
            lib.sg_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_points(self, counts):
        c = np.ascontiguousarray(counts, np.int64)
        po = np.zeros(len(self.grids) + 1, np.int64)
        co = np.zeros(len(self.grids) + 1, np.int64)
        llp = C.POINTER(C.c_longlong)
        check(lib.sg_batch_set_points(self._h, c.ctypes.data_as(llp), po.ctypes.data_as(llp),
                                      co.ctypes.data_as(llp)))
        self.point_offsets, self.coord_offsets = po, co

    def _check_eval(self, alpha, coords, out):
        if self.point_offsets is None:
            raise ValueError("set_points() first")
        dt = _fp_dtype(alpha, "alpha")
        _check(alpha, "alpha", dt, numel=self.total_values)
        _check(coords, "coords", dt, numel=int(self.coord_offsets[-1]))
        _check(out, "out", dt, min_numel=int(self.point_offsets[-1]))
        _on_device(alpha=alpha, coords=coords, out=out)
        return dt

    def hierarchize(self, values, stream=None):
        import torch
        dt = _fp_dtype(values, "values")
        _check(values, "values", dt, numel=self.total_values)
        _on_device(values=values)
        fn = lib.sg_batch_hierarchize_f64 if dt == torch.float64 else \
            lib.sg_batch_hierarchize_f32
        check(fn(self._h, _ptr(values), _stream(stream)))

    def evaluate(self, alpha, coords, out, stream=None):
        import torch
        dt = self._check_eval(alpha, coords, out)
        fn = lib.sg_batch_evaluate_f64 if dt == torch.float64 else \
            lib.sg_batch_evaluate_f32c
        check(fn(self._h, _ptr(alpha), _ptr(coords), _ptr(out), _stream(stream)))

    def run_graph(self, alpha, coords, out, hierarchize=True, stream=None):
        """Hierarchize (in place) then evaluate as one hipGraph launch."""
        import torch
        dt = self._check_eval(alpha, coords, out)
        fn = lib.sg_batch_run_graph_f64 if dt == torch.float64 else \
            lib.sg_batch_run_graph_f32c
        check(fn(self._h, _ptr(alpha), _ptr(coords), _ptr(out), int(hierarchize),
                 _stream(stream)))


class Server:
    """Evaluation server with continuous batching (sg_server in include/sgrid.h).

    ``precision`` "f64" or "f32c" fixes the dtype of surpluses, points and outputs.
    Requests carry host numpy arrays (row-major [P][d] points, [P] outputs) that are
    kept alive here until ``wait`` returns.
    """

    def __init__(self, precision: str = "f64", slots: int = 3, max_batch_points: int = 1 << 20,
                 max_requests: int = 1024, budget_bytes: int = 0):
        self.dtype = np.float64 if precision == "f64" else np.float32
        cfg = ServerConfig(0 if precision == "f64" else 1, slots, max_batch_points, max_requests,
                           budget_bytes)
        h = C.c_void_p()
        check(lib.sg_server_create(C.byref(cfg), C.byref(h)))
        self._h = h
        self._keep = {}
        self._grids = []

    def close(self):
        if getattr(self, "_h", None):
            lib.sg_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add_grid(self, grid: "Grid", alpha) -> int:
        gid = C.c_int()
        check(lib.sg_server_add_grid(self._h, grid._h, _ptr(alpha), C.byref(gid)))
        self._grids.append((grid, alpha))
        return gid.value

    def submit(self, grid_id: int, points: np.ndarray, out: np.ndarray) -> int:
        assert points.dtype == self.dtype and out.dtype == self.dtype
        assert points.flags["C_CONTIGUOUS"] and out.flags["C_CONTIGUOUS"]
        rid = C.c_longlong()
        check(lib.sg_server_submit(self._h, grid_id, points.ctypes.data, out.shape[0],
                                   out.ctypes.data, C.byref(rid)))
        self._keep[rid.value] = (points, out)
        return rid.value

    def poll(self, rid: int) -> bool:
        done = C.c_int()
        check(lib.sg_server_poll(self._h, rid, C.byref(done)))
        return bool(done.value)

    def wait(self, rid: int) -> float:
        """Block until request ``rid`` is done; returns its latency in seconds."""
        lat = C.c_double()
        try:
            check(lib.sg_server_wait(self._h, rid, C.byref(lat)))
        finally:
            self._keep.pop(rid, None)
        return lat.value

    def inject_launch_failure(self, grid_id: int, launches: int = 1) -> None:
        """Test hook: the next `launches` launch groups holding requests of grid_id fail to
        launch (their requests raise SG_E_HIP at wait); other requests are unaffected."""
        check(lib.sg_server_inject_launch_failure(self._h, grid_id, launches))

    def stats(self) -> dict:
        v = [C.c_longlong() for _ in range(4)]
        check(lib.sg_server_stats(self._h, *[C.byref(x) for x in v]))
        return dict(zip(("launches", "points_done", "waiting", "budget_used"), [x.value for x in v]))
