"""The staged walks (evaluate variants 3 and 4): the workgroup copies each coming subspace
block into LDS with buffer-to-LDS loads and every lane reads its coefficient there;
variant 4 lays the blocks out in walk order with compile-time offsets (regular grids with
n - 1 in 2..7 and the dimensions it is built for; other grids run variant 3).

Bars: the staged walks visit the same leaves in the same order and do the same
arithmetic per point as variant 0 (three tabled dimensions, batches of 8), so their
outputs are bit-identical to variant 0's in FP64 and FP32-Kahan; and they satisfy the
FP64 bound against the oracle, |gpu - oracle| <= 1e-12 * sum_j |term_j|.  Where variant
4 walks the sparse order (sgrid.blk_sparse: FP64 from d = 12, FP32 from d = 15) it sums
the blocks in another order, so there it must agree with variant 0 within the bound
(2e-12 sum |terms| FP64, 4e-6 sum |terms| + 2e-7 FP32) instead of bit for bit.  Shapes cover
one and many stages, blocks too large for a stage (gathered inside the staged walk),
blocks cut by restrictions (single-leaf batches), grids that cannot be staged (variant 0
runs), explicit level sets, workgroups with idle lanes, and every dimension-dependent
form of the walk."""
import numpy as np
import pytest

from oracle import oracle as O
from adaptive_sparse_grid_bench_amd import workloads as W

pytestmark = pytest.mark.gpu

REL_F64 = 1e-12
STAGED = 3
BLK = 4


@pytest.fixture(scope="module")
def sg(torch_gpu):
    from adaptive_sparse_grid_bench_amd import sgrid
    return sgrid


def dev(torch, a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def host(t):
    import torch
    torch.cuda.synchronize()
    return t.cpu().numpy()


def check_f64(out, ref, absref):
    assert np.isfinite(out).all()
    err = np.abs(out - ref)
    assert (err <= REL_F64 * absref + 1e-300).all(), float((err / (absref + 1e-300)).max())


def blk_agrees(o_blk, o0, d, prec, absref):
    """The block-staged walk against variant 0: bit-identical in the depth-first order,
    within the bound in the sparse order (sgrid.blk_sparse)."""
    from adaptive_sparse_grid_bench_amd import sgrid
    if not sgrid.blk_sparse(d, prec):
        return np.array_equal(o_blk, o0)
    e = np.abs(o_blk.astype(np.float64) - o0.astype(np.float64))
    bound = 2e-12 * absref + 1e-300 if prec == "f64" else 4e-6 * absref + 2e-7
    return bool((e <= bound).all())


def run_variants(torch, g, alpha, pts, dtype, binning):
    outs = {}
    for v in (0, STAGED, BLK):
        g.set_tuning(eval_variant_f64=v, eval_variant_f32=v, bin_min_points=binning)
        o = torch.full((len(pts),), float("nan"), dtype=dtype, device="cuda")
        g.evaluate(dev(torch, alpha).to(dtype), dev(torch, pts).to(dtype), o)
        outs[v] = host(o)
    return outs


SHAPES = [(10, 8, None, 3000), (10, 8, None, 65), (4, 10, None, 2000), (1, 8, None, 700),
          (2, 8, None, 1), (1, 1, None, 300), (20, 4, None, 1500), (3, 5, None, 257),
          (6, 7, [3, 1, 3, 2, 6, 6], 2500), (3, 6, [1, 5, 5], 999), (3, 11, None, 600),
          # every dimension-dependent form of the walk: batches of 8 (d <= 8) or 4, the
          # unroll caps of dimension 3 (d = 9..16) and of dimensions 4 and 3 (d = 10..13)
          (8, 6, None, 700), (9, 6, None, 700), (11, 5, None, 600), (12, 5, None, 600),
          (13, 5, None, 500), (14, 4, None, 500), (15, 4, None, 500), (16, 4, None, 500),
          (12, 6, [5, 5, 5, 5, 2, 5, 5, 5, 5, 5, 5, 5], 400)]


@pytest.mark.parametrize("d,n,lim,P", SHAPES,
                         ids=[f"d{d}n{n}" + ("" if l is None else "lim" + "".join(map(str, l)))
                              + f"P{P}" for d, n, l, P in SHAPES])
@pytest.mark.parametrize("binning", [-1, 1])
def test_staged_equals_variant0_and_oracle(sg, torch_gpu, d, n, lim, P, binning):
    torch = torch_gpu
    sh = O.tables(d, n, lim)
    x = O.grid_coords(sh)
    alpha = O.hierarchize(sh, W.f_gauss(x) - 0.5 * W.f_reference(x))
    pts = W.uniform_points(P, d, seed=50 + d + n)
    ref = O.evaluate(sh, alpha, pts)
    absref = O.evaluate(sh, np.abs(alpha), pts)
    g = sg.Grid(d, n, lim)
    o64 = run_variants(torch, g, alpha, pts, torch.float64, binning)
    assert np.array_equal(o64[STAGED], o64[0])
    assert blk_agrees(o64[BLK], o64[0], d, "f64", absref)
    check_f64(o64[STAGED], ref, absref)
    check_f64(o64[BLK], ref, absref)
    o32 = run_variants(torch, g, alpha, pts, torch.float32, binning)
    assert np.array_equal(o32[STAGED], o32[0])
    assert blk_agrees(o32[BLK], o32[0], d, "f32", absref)
    err = np.abs(o32[STAGED].astype(np.float64) - ref)
    assert (err <= 2e-6 * absref + 1e-7).all(), err.max()


def test_staged_interpolates_at_every_grid_point(sg, torch_gpu):
    """Size-independent check at the full d=10, n=8 grid: the interpolant equals the
    nodal values at all 1,862,145 grid points (FP64 bound relative to sum |terms|)."""
    torch = torch_gpu
    g = sg.Grid(10, 8)
    g.set_tuning(eval_variant_f64=STAGED)
    x = torch.empty((g.N, 10), dtype=torch.float64, device="cuda")
    g.grid_coords(x)
    f = torch.exp(-((x - 0.5) ** 2).sum(1))
    a = f.clone()
    g.hierarchize(a)
    out = torch.empty(g.N, dtype=torch.float64, device="cuda")
    g.evaluate(a, x, out)
    absu = torch.empty_like(out)
    g.evaluate(a.abs(), x, absu)
    assert bool(((out - f).abs() <= 1e-12 * absu + 1e-300).all())


def test_staged_explicit_sets(sg, torch_gpu):
    torch = torch_gpu
    for seed, d, cap in ((11, 8, 2000), (3, 3, 60), (5, 6, 400)):
        lv = W.adaptive_level_set(d=d, base_n=3, extra=4, seed=seed, max_subspaces=cap)[0]
        xs = O.xshape(lv)
        g = sg.Grid.explicit(d, lv)
        alpha = O.x_hierarchize(xs, W.f_gauss(O.x_grid_coords(xs)))
        pts = W.uniform_points(1500, d, seed=seed)
        ref = O.x_evaluate(xs, alpha, pts)
        absref = O.x_evaluate(xs, np.abs(alpha), pts)
        for binning in (-1, 1):
            o64 = run_variants(torch, g, alpha, pts, torch.float64, binning)
            assert np.array_equal(o64[STAGED], o64[0]), (seed, binning)
            assert np.array_equal(o64[BLK], o64[0]), (seed, binning)
            check_f64(o64[STAGED], ref, absref)
            o32 = run_variants(torch, g, alpha, pts, torch.float32, binning)
            assert np.array_equal(o32[STAGED], o32[0]), (seed, binning)
            assert np.array_equal(o32[BLK], o32[0]), (seed, binning)


def test_staged_batch_and_graph(sg, torch_gpu):
    """Batch launches with every member on the staged walk equal the per-grid calls,
    eagerly and through the captured graph."""
    torch = torch_gpu
    rng = np.random.default_rng(8)
    shapes = [(2, 3, None), (12, 4, None), (5, 6, [5, 1, 5, 2, 3]), (3, 8, None),
              (10, 5, None), (4, 10, None)]
    grids, vals = [], []
    for d, n, lim in shapes:
        sh = O.tables(d, n, lim)
        g = sg.Grid(d, n, lim)
        g.set_tuning(eval_variant_f64=STAGED, eval_variant_f32=STAGED)
        grids.append(g)
        vals.append(np.prod(np.sin(np.pi * O.grid_coords(sh)), axis=1))
    counts = [int(c) for c in rng.integers(1, 3000, len(grids))]
    b = sg.Batch(grids)
    b.set_points(counts)
    nodal = torch.from_numpy(np.concatenate(vals)).cuda()
    coords = torch.empty(int(b.coord_offsets[-1]), dtype=torch.float64, device="cuda")
    sg.uniform_points(coords, seed=12)
    for dt in (torch.float64, torch.float32):
        a = nodal.to(dt).clone()
        b.hierarchize(a)
        out = torch.full((int(b.point_offsets[-1]),), float("nan"), dtype=dt, device="cuda")
        b.evaluate(a, coords.to(dt), out)
        torch.cuda.synchronize()
        for i, g in enumerate(grids):
            lo, hi = int(b.value_offsets[i]), int(b.value_offsets[i + 1])
            c0, c1 = int(b.coord_offsets[i]), int(b.coord_offsets[i + 1])
            p0, p1 = int(b.point_offsets[i]), int(b.point_offsets[i + 1])
            oi = torch.empty(p1 - p0, dtype=dt, device="cuda")
            g.evaluate(a[lo:hi], coords.to(dt)[c0:c1].view(p1 - p0, g.d), oi)
            assert torch.equal(oi, out[p0:p1]), (i, dt)
        ref = out.clone()
        for _ in range(3):
            a.copy_(nodal.to(dt))
            out.fill_(float("nan"))
            b.run_graph(a, coords.to(dt), out, hierarchize=True)
            torch.cuda.synchronize()
            assert torch.equal(out, ref)
