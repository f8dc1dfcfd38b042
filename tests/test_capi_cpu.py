"""CPU-side checks of the C ABI library (no kernel launches): it loads, exports every
symbol include/*.h declares, and its host-side grid tables and bijection agree with the
oracle and the reference's golden vectors bit for bit."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from oracle import oracle as O

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))


def declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(sg_\w+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    from adaptive_sparse_grid_bench_amd import sgrid
    names = declared("sgrid.h")
    assert names, "no declarations parsed"
    lib = C.CDLL(sgrid.LIB_PATH)
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert set(names) == set(sgrid.EXPORTS)


@pytest.mark.parametrize("d,n,lim", [(3, 5, None), (4, 5, [4, 1, 4, 2]), (10, 8, None), (20, 6, None)])
def test_host_tables_and_bijection(d, n, lim):
    from adaptive_sparse_grid_bench_amd import sgrid
    g = sgrid.Grid(d, n, lim)
    sh = O.tables(d, n, lim)
    assert (g.N, g.S) == (sh.N, sh.S)
    dp, lidx = g.tables()
    assert (dp.ravel() == sh.dp).all() and (lidx == sh.lidx).all()
    rng = np.random.default_rng(d * 100 + n)
    for o in list(rng.integers(0, sh.N, 300)) + [0, sh.N - 1]:
        lv, iv = g.idx2gp(int(o))
        rl, ri = O.idx2gp(sh, int(o))
        assert (lv == rl).all() and (iv == ri).all()
        assert g.gp2idx(lv, iv) == o
        s = int(lv.sum())
        r = g.l2idx(lv)
        assert r == O.l2idx(sh, lv)
        assert (g.idx2l(r, s) == lv).all()


def test_host_errors():
    from adaptive_sparse_grid_bench_amd import sgrid
    with pytest.raises(sgrid.SgError):
        sgrid.Grid(0, 3)
    with pytest.raises(sgrid.SgError):
        sgrid.Grid(21, 3)
    g = sgrid.Grid(3, 4, [3, 1, 2])
    with pytest.raises(sgrid.SgError) as e:
        g.idx2gp(g.N)
    assert e.value.status == sgrid.SG_E_RANGE
    with pytest.raises(sgrid.SgError):
        g.gp2idx([0, 2, 0], [0, 0, 0])       # l_1 above its limit
    with pytest.raises(sgrid.SgError):
        g.idx2l(0, 4)                        # level sum >= n
    with pytest.raises(sgrid.SgError):
        g.set_tuning(eval_block=96)
    with pytest.raises(sgrid.SgError):
        g.set_tuning(eval_variant_f64=sgrid.EVAL_VARIANTS)
    with pytest.raises(TypeError):
        g.set_tuning(points_per_lane=2)
    g.set_tuning(eval_block=128, bin_min_points=-1)
    assert g.get_tuning()["eval_block"] == 128 and g.get_tuning()["bin_min_points"] == -1


def test_noalloc_evaluate_argument_checks():
    """sg_evaluate_f64_noalloc refuses bad arguments and point counts beyond its fixed
    scratch before it touches a device."""
    from adaptive_sparse_grid_bench_amd import sgrid
    g = sgrid.Grid(3, 4)
    fn = sgrid.lib.sg_evaluate_f64_noalloc
    assert fn(None, 8, 8, 4, 8, None) == sgrid.SG_E_INVALID
    assert fn(g._h, None, 8, 4, 8, None) == sgrid.SG_E_INVALID
    assert fn(g._h, 8, 8, -1, 8, None) == sgrid.SG_E_INVALID
    assert fn(g._h, 8, 8, 2097152 + 1, 8, None) == sgrid.SG_E_RANGE
    assert "SG_NOALLOC_MAX_POINTS" in sgrid.lib.sg_last_error().decode()


def test_walk_variant_fallbacks():
    """sg_grid_walk_variant reports the walk a tuned variant runs: the two-point walk
    only in FP64 at d = 9..11 (depth-first order), the block-staged walks only on regular
    grids with n - 1 in 2..7 (FP64 from d = 9, FP32 from d = 7), the staged walk on
    explicit sets and where its stages hold the blocks, gathers elsewhere; the gathering
    walks 1 and 2 only where their tabled dimensions stay below level 8 (not on explicit
    sets)."""
    from adaptive_sparse_grid_bench_amd import sgrid
    g = sgrid.Grid(10, 8)
    assert [g.walk_variant(v, "f64") for v in range(6)] == [0, 1, 2, 3, 4, 5]
    assert [g.walk_variant(v, "f32") for v in range(6)] == [0, 1, 2, 3, 4, 4]
    g20 = sgrid.Grid(20, 6)                  # sparse block order: one point per lane
    assert g20.walk_variant(5, "f64") == 4 and g20.walk_variant(5, "f32") == 4
    assert sgrid.Grid(12, 7).walk_variant(5, "f64") == 4
    g8 = sgrid.Grid(8, 5)
    assert g8.walk_variant(5, "f64") == 3 and g8.walk_variant(5, "f32") == 4
    assert sgrid.Grid(6, 5).walk_variant(4, "f32") == 3
    # n - 1 = 8: no block-staged kernel, level-8 blocks exceed the staged walk's tables,
    # and the gathering walks run without tabled dimensions, as one walk
    assert [sgrid.Grid(10, 9).walk_variant(v, "f64") for v in range(6)] == [0] * 6
    assert [sgrid.Grid(3, 9, [2, 2, 8]).walk_variant(v, "f64") for v in (1, 2)] == [1, 0]
    lv, _ = sgrid.Grid(4, 5, [4, 1, 4, 2]).subspaces()
    ge = sgrid.Grid.explicit(4, lv)
    assert [ge.walk_variant(v, "f64") for v in range(6)] == [0, 0, 0, 3, 3, 3]
    for bad in ((g._h, 2, 0), (g._h, 8, 6), (g._h, 8, -1), (None, 8, 0)):
        out = C.c_int()
        assert sgrid.lib.sg_grid_walk_variant(*bad, C.byref(out)) == sgrid.SG_E_INVALID
    assert sgrid.lib.sg_grid_walk_variant(g._h, 8, 0, None) == sgrid.SG_E_INVALID


def test_explicit_set_matches_restricted_regular():
    """A restricted-regular level set given explicitly gets the identical compact order."""
    from adaptive_sparse_grid_bench_amd import sgrid
    lim = [4, 1, 4, 2]
    g = sgrid.Grid(4, 5, lim)
    lv, base = g.subspaces()
    perm = np.random.default_rng(3).permutation(len(lv))
    ge = sgrid.Grid.explicit(4, lv[perm])
    lv2, base2 = ge.subspaces()
    assert (lv2 == lv).all() and (base2 == base).all() and ge.N == g.N
    for o in (0, 7, g.N // 2, g.N - 1):
        a, b = g.idx2gp(o)
        c, e = ge.idx2gp(o)
        assert (a == c).all() and (b == e).all()
        assert ge.gp2idx(a, b) == o
    with pytest.raises(sgrid.SgError):      # not downward closed
        sgrid.Grid.explicit(2, np.array([[0, 0], [0, 2]], np.uint8))


def test_sample_argument_errors():
    """Parameter checks of sg_sample_* run before any device work."""
    from adaptive_sparse_grid_bench_amd import sgrid
    g = sgrid.Grid(3, 4)
    fake = C.c_void_p(1)
    bad = [
        (9, None, 0),                                   # unknown function
        (sgrid.SAMPLE_FUNCTIONS["parabola"], [1.0], 1),  # takes no parameters
        (sgrid.SAMPLE_FUNCTIONS["gauss"], [0.5, 0.5], 2),  # 0 or d centres
        (sgrid.SAMPLE_FUNCTIONS["gaussian_sum"], [9.0], 1),  # more than 8 terms
        (sgrid.SAMPLE_FUNCTIONS["gaussian_sum"], [1.0] + [0.5] * 3 + [0.0] * 3, 7),  # width 0
        (sgrid.SAMPLE_FUNCTIONS["sin_product"], [1.0, 2.0], 2),
    ]
    for fn, params, count in bad:
        arr = (C.c_double * max(count, 1))(*(params or [0.0]))
        st = sgrid.lib.sg_sample_f64(g._h, fn, arr if params else None, count, fake, None)
        assert st == sgrid.SG_E_INVALID, (fn, params)


def test_wrappers_reject_bad_tensors():
    """The Python mirror checks dtype, layout and sizes before handing a tensor to the
    library (a mismatch would otherwise be read out of bounds or misread on the device);
    host tensors are refused last."""
    import torch
    from adaptive_sparse_grid_bench_amd import sgrid
    g = sgrid.Grid(3, 4)
    P = 10
    a64 = torch.zeros(g.N, dtype=torch.float64)
    x64 = torch.zeros((P, 3), dtype=torch.float64)
    o64 = torch.zeros(P, dtype=torch.float64)
    cases = [
        (lambda: g.evaluate(a64.float(), x64, o64), "dtype"),              # alpha FP32, coords FP64
        (lambda: g.evaluate(a64, x64.float(), o64), "dtype"),
        (lambda: g.evaluate(a64, x64, o64.float()), "dtype"),
        (lambda: g.evaluate(a64.half(), x64.half(), o64.half()), "float64 or float32"),
        (lambda: g.evaluate(a64[:-1], x64, o64), "elements"),               # alpha too short
        (lambda: g.evaluate(a64, x64.reshape(-1), o64), "[P][d]"),          # flattened points
        (lambda: g.evaluate(a64, torch.zeros((P, 4), dtype=torch.float64), o64), "shape"),
        (lambda: g.evaluate(a64, torch.zeros((3, P), dtype=torch.float64).t(), o64),
         "contiguous"),                                                     # column-major view
        (lambda: g.evaluate(a64, x64, o64[:-1]), "at least"),               # out too short
        (lambda: g.evaluate(a64, x64, o64), "CUDA"),                        # all fine but host
        (lambda: g.hierarchize(a64[:-1]), "elements"),
        (lambda: g.hierarchize(a64.int()), "float64 or float32"),
        (lambda: g.hierarchize(a64), "CUDA"),
        (lambda: g.grid_coords(torch.zeros(g.N * 3, dtype=torch.float32)), "dtype"),
        (lambda: g.sample("gauss", a64[:5]), "elements"),
        (lambda: sgrid.uniform_points(torch.zeros(4, dtype=torch.int64)), "float64 or float32"),
        (lambda: sgrid.to_f32(a64, torch.zeros(3, dtype=torch.float32)), "at least"),
    ]
    for fn, what in cases:
        with pytest.raises(ValueError) as e:
            fn()
        assert what in str(e.value), (what, str(e.value))


def test_batch_rejects_oversized_members():
    """A batch member whose surpluses would not fit the evaluate kernels' 2 GiB buffer
    range is refused before any device work (a single-grid evaluate refuses it too)."""
    from adaptive_sparse_grid_bench_amd import sgrid
    big = sgrid.Grid(4, 22, [7, 7, 7, 7])        # N = 671,480,833: FP32 bytes > 2^31
    assert big.N * 4 > 0x7fffffff
    with pytest.raises(sgrid.SgError) as e:
        sgrid.Batch([sgrid.Grid(3, 4), big])
    assert e.value.status == sgrid.SG_E_RANGE
    assert "2 GiB" in str(e.value)


def test_batch_wrappers_check_sizes():
    import torch
    from adaptive_sparse_grid_bench_amd import sgrid
    b = sgrid.Batch.__new__(sgrid.Batch)      # sizes only: no library batch behind it
    b.total_values, b.point_offsets, b.coord_offsets = 100, None, None
    with pytest.raises(ValueError):
        b.evaluate(torch.zeros(100, dtype=torch.float64), torch.zeros(3, dtype=torch.float64),
                   torch.zeros(1, dtype=torch.float64))
    b.point_offsets = np.array([0, 4, 10])
    b.coord_offsets = np.array([0, 8, 26])
    a, x, o = (torch.zeros(100, dtype=torch.float64), torch.zeros(26, dtype=torch.float64),
               torch.zeros(10, dtype=torch.float64))
    for args, what in [((a[:-1], x, o), "elements"), ((a, x[:-1], o), "elements"),
                       ((a, x, o[:-1]), "at least"), ((a, x.float(), o), "dtype"),
                       ((a, x, o), "CUDA")]:
        with pytest.raises(ValueError) as e:
            b.evaluate(*args)
        assert what in str(e.value)
    with pytest.raises(ValueError):
        b.hierarchize(a[:50])
    b._h = None
