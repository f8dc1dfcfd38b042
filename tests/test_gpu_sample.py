"""Device sampling of the test functions at the grid points (the step before
hierarchization, asg.c:143-152).

Bars: the reference's own function in FP32 is bit-identical to the reference's init()
values (golden fixtures written by the reference itself); the FP64 functions agree with
numpy on the oracle's grid coordinates within a few ulp (the device exp / sin are not
glibc's), stated per function below."""
import glob
import os

import numpy as np
import pytest

from oracle import oracle as O
from adaptive_sparse_grid_bench_amd import workloads as W

pytestmark = pytest.mark.gpu

GOLDEN = sorted(glob.glob(os.path.join(os.path.dirname(__file__), "golden", "ref_*.npz")))


@pytest.fixture(scope="module")
def sg(torch_gpu):
    from adaptive_sparse_grid_bench_amd import sgrid
    return sgrid


def with_init(path):
    with np.load(path) as z:
        return "init_values" in z.files and "levels" in z.files


@pytest.mark.parametrize("path", [p for p in GOLDEN if with_init(p)],
                         ids=[os.path.basename(p)[4:-4] for p in GOLDEN if with_init(p)])
def test_reference_function_f32_equals_reference_init(sg, torch_gpu, path):
    torch = torch_gpu
    z = np.load(path)
    d, n = int(z["d"]), int(z["n"])
    lim = [int(v) for v in z["limits"]]
    g = sg.Grid(d, n, lim)
    out = torch.empty(g.N, dtype=torch.float32, device="cuda")
    g.sample("reference", out)
    torch.cuda.synchronize()
    assert np.array_equal(out.cpu().numpy(), z["init_values"])


SHAPES = [(3, 5, None), (10, 5, None), (4, 5, [4, 1, 4, 2]), (20, 3, None)]


@pytest.mark.parametrize("d,n,lim", SHAPES, ids=[f"d{d}n{n}" for d, n, _ in SHAPES])
def test_functions_f64_match_numpy(sg, torch_gpu, d, n, lim):
    torch = torch_gpu
    sh = O.tables(d, n, lim)
    x = O.grid_coords(sh)
    g = sg.Grid(d, n, lim)
    out = torch.empty(g.N, dtype=torch.float64, device="cuda")

    def run(name, params=None):
        g.sample(name, out, params)
        torch.cuda.synchronize()
        return out.cpu().numpy()

    # only +, -, * in the same order: exact up to numpy's product order (1 ulp per factor)
    assert np.allclose(run("reference"), W.f_reference(x), rtol=4 * d * 2.0 ** -53, atol=0)
    assert np.allclose(run("parabola"), W.f_prod_parabola(x), rtol=4 * d * 2.0 ** -53, atol=0)
    # exp and sin: device library within a few ulp
    assert np.allclose(run("gauss"), W.f_gauss(x), rtol=1e-14, atol=0)
    c = np.linspace(0.3, 0.7, d)
    assert np.allclose(run("gauss", c), np.exp(-np.sum((x - c) ** 2, axis=1)), rtol=1e-14, atol=0)
    cs, ws = W.gaussian_sum_params(d, 8, seed=7)
    got = run("gaussian_sum", np.concatenate([[8.0], cs.ravel(), ws.ravel()]))
    assert np.allclose(got, W.f_gaussian_sum(x, cs, ws), rtol=1e-13, atol=1e-300)
    assert np.allclose(run("sin_product", [1.05]), W.f_sin_prod(x, 0.5), rtol=1e-14, atol=1e-300)


def test_explicit_set_sampling(sg, torch_gpu):
    torch = torch_gpu
    lv, _ = W.adaptive_level_set(d=8, base_n=3, extra=4, seed=11)
    xs = O.xshape(lv)
    g = sg.Grid.explicit(8, lv)
    out = torch.empty(g.N, dtype=torch.float64, device="cuda")
    g.sample("gauss", out)
    torch.cuda.synchronize()
    assert np.allclose(out.cpu().numpy(), W.f_gauss(O.x_grid_coords(xs)), rtol=1e-14, atol=0)
