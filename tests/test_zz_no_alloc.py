"""The allocation-free evaluate entry point (sg_evaluate_f64_noalloc) through the C ABI.

It has no counterpart in the reference.  Bar: its outputs are bit-identical to
sg_evaluate_f64's on the same inputs (the binning permutation lives in the library's
fixed device scratch instead of stream-ordered allocations; every point's arithmetic is
the same).  Each case makes two calls in a row, the second with fewer points, so the
scratch holds the first call's contents when the second starts.  One sample of each
output is also checked against the CPU oracle (FP64 bound 1e-12 * sum |terms|).
"""
import ctypes as C

import numpy as np
import pytest

from oracle import oracle as O
from adaptive_sparse_grid_bench_amd import workloads as W

pytestmark = pytest.mark.gpu

MAX_POINTS = 2097152          # SG_NOALLOC_MAX_POINTS (include/sgrid.h)


@pytest.fixture(scope="module")
def sg(torch_gpu):
    from adaptive_sparse_grid_bench_amd import sgrid
    return sgrid


def _noalloc(sg, g, alpha, coords, P, out, stream=None):
    return sg.lib.sg_evaluate_f64_noalloc(g._h, alpha.data_ptr(), coords.data_ptr(), P,
                                          out.data_ptr(), sg._stream(stream))


def _setup(sg, torch, d, n, variant=None):
    sh = O.tables(d, n)
    g = sg.Grid(d, n)
    if variant is not None:
        g.set_tuning(eval_variant_f64=variant, bin_min_points=4096)
    a = torch.from_numpy(W.f_gauss(O.grid_coords(sh))).cuda()
    g.hierarchize(a)                       # uploads the grid's device tables
    return sh, g, a


# (d, n, variant): a gathering walk that bins its points (S >= 1024, P >= 4096) and the
# default staged walk, which does not bin
CASES = [(8, 6, 0, (MAX_POINTS, 5000)), (8, 6, 2, (70000, 4096)), (5, 5, None, (40000, 777))]


@pytest.mark.parametrize("d,n,variant,sizes", CASES,
                         ids=[f"d{c[0]}n{c[1]}v{c[2]}" for c in CASES])
def test_noalloc_equals_evaluate(sg, torch_gpu, d, n, variant, sizes):
    torch = torch_gpu
    print(f"noalloc case d={d} n={n} variant={variant} sizes={sizes}", flush=True)
    sh, g, a = _setup(sg, torch, d, n, variant)
    alpha_h = a.cpu().numpy()
    for k, P in enumerate(sizes):
        x = torch.empty((P, d), dtype=torch.float64, device="cuda")
        sg.uniform_points(x, seed=100 + k)
        ref = torch.empty(P, dtype=torch.float64, device="cuda")
        out = torch.full((P,), float("nan"), dtype=torch.float64, device="cuda")
        g.evaluate(a, x, ref)
        st = _noalloc(sg, g, a, x, P, out)
        assert st == sg.SG_OK, sg.lib.sg_last_error().decode()
        torch.cuda.synchronize()
        o, r = out.cpu().numpy(), ref.cpu().numpy()
        assert np.array_equal(o, r), f"call {k}: differs from sg_evaluate_f64"
        idx = np.linspace(0, P - 1, 256).astype(np.int64)
        pts = x.cpu().numpy()[idx]
        want = O.evaluate(sh, alpha_h, pts)
        bound = 1e-12 * O.evaluate(sh, np.abs(alpha_h), pts) + 1e-300
        assert (np.abs(o[idx] - want) <= bound).all()


def test_noalloc_rejects(sg, torch_gpu):
    torch = torch_gpu
    d, n = 4, 4
    g = sg.Grid(d, n)
    a = torch.zeros(g.N, dtype=torch.float64, device="cuda")
    x = torch.zeros((8, d), dtype=torch.float64, device="cuda")
    o = torch.zeros(8, dtype=torch.float64, device="cuda")
    # tables not uploaded yet: uploading would allocate
    assert _noalloc(sg, g, a, x, 8, o) == sg.SG_E_INVALID
    assert "not uploaded" in sg.lib.sg_last_error().decode()
    g.hierarchize(a)
    assert _noalloc(sg, g, a, x, 8, o) == sg.SG_OK
    assert _noalloc(sg, g, a, x, MAX_POINTS + 1, o) == sg.SG_E_RANGE
    assert "SG_NOALLOC_MAX_POINTS" in sg.lib.sg_last_error().decode()
    assert _noalloc(sg, g, a, x, 0, o) == sg.SG_OK
    torch.cuda.synchronize()
