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
