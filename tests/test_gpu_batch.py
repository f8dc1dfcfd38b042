"""The batched form: many grids of uneven dimension and level resident together,
hierarchized and evaluated by batch launches and by one hipGraph launch.  Bar: every
grid's surpluses and outputs bit-identical to the per-grid calls (same kernels, same
operation order per point); the graph's replays identical to the eager launches."""
import numpy as np
import pytest

from oracle import oracle as O
from adaptive_sparse_grid_bench_amd import workloads as W

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def sg(torch_gpu):
    from adaptive_sparse_grid_bench_amd import sgrid
    return sgrid


def make_members(sg, rng):
    grids, vals = [], []
    shapes = [(2, 3, None), (12, 4, None), (5, 6, [5, 1, 5, 2, 3]), (3, 8, None),
              (10, 5, None), (7, 3, None)]
    for d, n, lim in shapes:
        sh = O.tables(d, n, lim)
        g = sg.Grid(d, n, lim)
        x = O.grid_coords(sh)
        grids.append(g)
        vals.append(np.prod(np.sin(np.pi * x), axis=1) * (1 + 0.1 * rng.random()))
    lv, _ = W.adaptive_level_set(d=8, base_n=3, extra=4, seed=2, max_subspaces=300)
    xs = O.xshape(lv)
    g = sg.Grid.explicit(8, lv)
    grids.append(g)
    vals.append(W.f_gauss(O.x_grid_coords(xs)))
    return grids, vals


def test_batch_bit_identical_to_per_grid(sg, torch_gpu):
    torch = torch_gpu
    rng = np.random.default_rng(3)
    grids, vals = make_members(sg, rng)
    counts = [int(c) for c in rng.integers(1, 3000, len(grids))]
    counts[1] = 256                          # exactly one workgroup
    b = sg.Batch(grids)
    b.set_points(counts)
    alpha_all = torch.from_numpy(np.concatenate(vals)).cuda()
    assert alpha_all.numel() == b.total_values
    pts = [W.uniform_points(c, g.d, seed=40 + i) for i, (c, g) in enumerate(zip(counts, grids))]
    coords = torch.from_numpy(np.concatenate([p.ravel() for p in pts])).cuda()
    assert coords.numel() == b.coord_offsets[-1]
    for dt in (torch.float64, torch.float32):
        a = alpha_all.to(dt).clone()
        b.hierarchize(a)
        out = torch.full((b.point_offsets[-1],), float("nan"), dtype=dt, device="cuda")
        b.evaluate(a, coords.to(dt), out)
        torch.cuda.synchronize()
        for i, g in enumerate(grids):
            lo, hi = b.value_offsets[i], b.value_offsets[i + 1]
            ai = alpha_all[lo:hi].to(dt).clone()
            g.hierarchize(ai)
            assert torch.equal(ai, a[lo:hi]), (i, "hierarchize")
            oi = torch.empty(counts[i], dtype=dt, device="cuda")
            g.evaluate(ai, torch.from_numpy(pts[i]).cuda().to(dt), oi)
            assert torch.equal(oi, out[b.point_offsets[i]:b.point_offsets[i + 1]]), (i, "eval")


def test_batch_graph_replays(sg, torch_gpu):
    """hipGraph capture of hierarchize + evaluate, replayed several times with eager
    work and synchronisation between replays; every replay equals the eager result."""
    torch = torch_gpu
    rng = np.random.default_rng(4)
    grids, vals = make_members(sg, rng)
    counts = [int(c) for c in rng.integers(100, 2000, len(grids))]
    b = sg.Batch(grids)
    b.set_points(counts)
    nodal = torch.from_numpy(np.concatenate(vals)).cuda()
    coords = torch.empty(int(b.coord_offsets[-1]), dtype=torch.float64, device="cuda")
    sg.uniform_points(coords, seed=9)
    ref_a = nodal.clone()
    b.hierarchize(ref_a)
    ref_o = torch.empty(int(b.point_offsets[-1]), dtype=torch.float64, device="cuda")
    b.evaluate(ref_a, coords, ref_o)
    a = torch.empty_like(nodal)
    out = torch.empty_like(ref_o)
    for rep in range(4):
        a.copy_(nodal)
        out.fill_(float("nan"))
        b.run_graph(a, coords, out, hierarchize=True)
        torch.cuda.synchronize()
        assert torch.equal(a, ref_a), rep
        assert torch.equal(out, ref_o), rep
        junk = torch.randn(1 << 16, device="cuda").sum()     # eager work between replays
        torch.cuda.synchronize()
        assert torch.isfinite(junk)
    # evaluate-only graph on the FP32 copy
    a32 = ref_a.float()
    c32 = coords.float()
    o32 = torch.empty(int(b.point_offsets[-1]), dtype=torch.float32, device="cuda")
    ref32 = torch.empty_like(o32)
    b.evaluate(a32, c32, ref32)
    for rep in range(3):
        o32.zero_()
        b.run_graph(a32, c32, o32, hierarchize=False)
        torch.cuda.synchronize()
        assert torch.equal(o32, ref32), rep
