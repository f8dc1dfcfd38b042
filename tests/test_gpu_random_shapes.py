"""Seeded random grid shapes through every walk variant.

For each shape: regular or restricted grids with d = 1..20 and up to 8 level sums, with
random per-dimension limits, and dimension-adaptive explicit level sets.  The checks:
  * the staged walk (variant 3) equals variant 0 bit for bit, in FP64 and FP32-Kahan:
    same leaves, same order, same arithmetic per point;
  * FP64 variants 0..3 meet |gpu - oracle| <= 1e-12 * sum_j |term_j| against the oracle's
    asg.c-order evaluate;
  * FP32-Kahan stays within 2e-6 * sum_j |term_j| + 1e-7 of the oracle.
Shapes are drawn with a fixed seed; sizes stay small enough for the oracle (N <= 300,000,
S <= 20,000, a few hundred points)."""
import numpy as np
import pytest

from oracle import oracle as O
from adaptive_sparse_grid_bench_amd import workloads as W

pytestmark = pytest.mark.gpu

REL_F64 = 1e-12
STAGED = 3


def draw_shapes(count: int, seed: int):
    rng = np.random.default_rng(seed)
    out = []
    while len(out) < count:
        d = int(rng.integers(1, 21))
        n = int(rng.integers(1, 9))
        lim = None
        if rng.random() < 0.4:
            lim = [int(x) for x in rng.integers(0, n, d)]
        sh = O.tables(d, n, lim)
        if sh.N > 300000 or sh.S > 20000 or (d, n, lim) in out:
            continue
        out.append((d, n, lim))
    return out


SHAPES = draw_shapes(36, seed=20261204)


@pytest.fixture(scope="module")
def sg(torch_gpu):
    from adaptive_sparse_grid_bench_amd import sgrid
    return sgrid


def run(torch, g, alpha, pts, dtype, variant):
    g.set_tuning(eval_variant_f64=variant, eval_variant_f32=variant, bin_min_points=-1)
    o = torch.full((len(pts),), float("nan"), dtype=dtype, device="cuda")
    g.evaluate(torch.from_numpy(alpha).cuda().to(dtype), torch.from_numpy(pts).cuda().to(dtype), o)
    torch.cuda.synchronize()
    return o.cpu().numpy()


def check_all_variants(sg, torch, g, alpha, pts, ref, absref, tag):
    out64 = {v: run(torch, g, alpha, pts, torch.float64, v) for v in range(sg.EVAL_VARIANTS)}
    for v, o in out64.items():
        assert np.isfinite(o).all(), (tag, v)
        err = np.abs(o - ref)
        assert (err <= REL_F64 * absref + 1e-300).all(), (tag, v, float(err.max()))
    assert np.array_equal(out64[STAGED], out64[0]), tag
    o32 = run(torch, g, alpha, pts, torch.float32, STAGED)
    assert np.array_equal(o32, run(torch, g, alpha, pts, torch.float32, 0)), tag
    err = np.abs(o32.astype(np.float64) - ref)
    assert (err <= 2e-6 * absref + 1e-7).all(), (tag, float(err.max()))


@pytest.mark.parametrize("d,n,lim", SHAPES,
                         ids=[f"d{d}n{n}" + ("" if l is None else "lim" + "".join(map(str, l)))
                              for d, n, l in SHAPES])
def test_random_regular_shapes(sg, torch_gpu, d, n, lim):
    sh = O.tables(d, n, lim)
    x = O.grid_coords(sh)
    alpha = O.hierarchize(sh, W.f_gauss(x) - 0.25 * W.f_prod_parabola(x))
    pts = W.uniform_points(300, d, seed=d * 31 + n)
    ref = O.evaluate(sh, alpha, pts)
    absref = O.evaluate(sh, np.abs(alpha), pts)
    check_all_variants(sg, torch_gpu, sg.Grid(d, n, lim), alpha, pts, ref, absref, (d, n, lim))


@pytest.mark.parametrize("seed", [101, 202, 303, 404, 505, 606])
def test_random_adaptive_sets(sg, torch_gpu, seed):
    rng = np.random.default_rng(seed)
    d = int(rng.integers(2, 11))
    lv = W.adaptive_level_set(d=d, base_n=int(rng.integers(2, 4)), extra=int(rng.integers(1, 5)),
                              seed=seed, max_subspaces=int(rng.integers(50, 1500)))[0]
    xs = O.xshape(lv)
    alpha = O.x_hierarchize(xs, W.f_gauss(O.x_grid_coords(xs)))
    pts = W.uniform_points(300, d, seed=seed)
    ref = O.x_evaluate(xs, alpha, pts)
    absref = O.x_evaluate(xs, np.abs(alpha), pts)
    check_all_variants(sg, torch_gpu, sg.Grid.explicit(d, lv), alpha, pts, ref, absref,
                       ("explicit", seed, d))
