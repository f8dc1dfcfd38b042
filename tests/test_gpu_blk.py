"""The block-staged walk (evaluate variant 4) and the staged-pattern probe.

Variant 4 runs on regular grids with n - 1 in 2..7, FP64 for d = 9..20 and FP32 for
d = 7..20 (other grids run variant 3).  Its static blocks sit in LDS in walk order at
compile-time offsets (n - 1 is a template constant).  From d = 12 (FP64) and d = 15
(FP32) on it walks the sparse order (sgrid.blk_sparse).  Grids with n = 8 at d = 9..11
and n = 7 at d = 10, 11 run it in 512-thread workgroups at 4 waves per SIMD with packed
lane tables (d9n8, d10n7, d10n8, d11n8 below).  Bars:
  * depth-first order: outputs bit-identical to variant 0 (same leaves, same order, same
    arithmetic per point; the FP64 walk computes each outer hat from x at its level,
    which on the 2^-24 lattice equals the incremental form);
  * sparse order: the blocks are summed in another order and one-leaf blocks take a
    tabled product of level-0 hats, so outputs are within the bound of variant 0's and
    of the oracle's, not bit-identical;
  * every output within the FP64 bound of the oracle,
    |gpu - oracle| <= 1e-12 * sum_j |term_j|; FP32 within 2e-6 sum_j |term_j| + 1e-7;
  * interpolation at every grid point of the d=10, n=8 (FP32) and d=20, n=6 (FP64) grids;
  * ragged point counts (idle lanes in the last workgroup);
  * variant 5, the same walk with two points per lane (FP64, depth-first order),
    bit-identical to variants 4 and 0, and running as variant 4 where it is not built;
  * the staged-pattern probe reads once per subspace per point, each read landing on
    copied surplus data (all-ones surpluses give S times the point's weight).
"""
import numpy as np
import pytest

from oracle import oracle as O
from adaptive_sparse_grid_bench_amd import workloads as W

pytestmark = pytest.mark.gpu

BLK, STAGED = 4, 3


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


def evaluate(torch, g, alpha, pts, dt, variant):
    g.set_tuning(eval_variant_f64=variant, eval_variant_f32=variant)
    o = torch.full((len(pts),), float("nan"), dtype=dt, device="cuda")
    g.evaluate(dev(torch, alpha).to(dt), dev(torch, pts).to(dt), o)
    return host(o)


SHAPES64 = [(9, 6, 700), (9, 8, 700), (10, 7, 600), (10, 8, 1100), (11, 3, 900), (11, 8, 300),
            (12, 5, 700),
            (14, 4, 800),
            (16, 6, 200), (17, 4, 500), (20, 3, 1000), (20, 5, 300), (20, 8, 40)]
SHAPES32 = [(7, 8, 900), (8, 3, 1000), (9, 8, 1100), (10, 8, 500), (11, 7, 600), (11, 8, 300),
            (13, 6, 300),
            (15, 5, 400), (16, 4, 700),
            (18, 5, 300), (20, 6, 100)]


def same_walk_result(o4, o0, d, prec, absref):
    """Variant 4's outputs against variant 0's: bit-identical in the depth-first order,
    within the precision's bound (relative to sum |terms|) in the sparse order."""
    from adaptive_sparse_grid_bench_amd import sgrid
    if not sgrid.blk_sparse(d, prec):
        return np.array_equal(o4, o0)
    e = np.abs(o4.astype(np.float64) - o0.astype(np.float64))
    bound = 2e-12 * absref + 1e-300 if prec == "f64" else 4e-6 * absref + 2e-7
    return bool((e <= bound).all())


@pytest.mark.parametrize("prec,d,n,P", [("f64",) + s for s in SHAPES64] +
                         [("f32",) + s for s in SHAPES32],
                         ids=[f"f64-d{d}n{n}" for d, n, _ in SHAPES64] +
                             [f"f32-d{d}n{n}" for d, n, _ in SHAPES32])
def test_blk_equals_variant0_and_oracle(sg, torch_gpu, prec, d, n, P):
    torch = torch_gpu
    sh = O.tables(d, n)
    x = O.grid_coords(sh)
    alpha = O.hierarchize(sh, W.f_gauss(x) - 0.5 * W.f_reference(x))
    pts = W.uniform_points(P, d, seed=70 + d + n)
    pts[0] = 0.0
    ref = O.evaluate(sh, alpha, pts)
    absref = O.evaluate(sh, np.abs(alpha), pts)
    g = sg.Grid(d, n)
    dt = torch.float64 if prec == "f64" else torch.float32
    o4 = evaluate(torch, g, alpha, pts, dt, BLK)
    o0 = evaluate(torch, g, alpha, pts, dt, 0)
    assert same_walk_result(o4, o0, d, prec, absref)
    err = np.abs(o4.astype(np.float64) - ref)
    if prec == "f64":
        assert (err <= 1e-12 * absref + 1e-300).all(), float((err / (absref + 1e-300)).max())
    else:
        assert (err <= 2e-6 * absref + 1e-7).all(), float(err.max())


@pytest.mark.parametrize("d,n", [(9, 6), (10, 8), (11, 5), (9, 3)])
def test_two_points_per_lane(sg, torch_gpu, d, n):
    """Variant 5 (two points per lane, FP64, depth-first order) equals variants 4 and 0
    bit for bit (every point's arithmetic is the lean walk's) on ragged counts around the
    512 points of a workgroup, and is within the FP64 bound of the oracle."""
    torch = torch_gpu
    sh = O.tables(d, n)
    alpha = O.hierarchize(sh, W.f_gauss(O.grid_coords(sh)) - 0.5 * W.f_reference(O.grid_coords(sh)))
    g = sg.Grid(d, n)
    for P in (1, 255, 256, 257, 511, 512, 513, 1500):
        pts = W.uniform_points(P, d, seed=900 + P)
        o5 = evaluate(torch, g, alpha, pts, torch.float64, 5)
        o4 = evaluate(torch, g, alpha, pts, torch.float64, BLK)
        o0 = evaluate(torch, g, alpha, pts, torch.float64, 0)
        assert np.array_equal(o5, o4) and np.array_equal(o5, o0), P
        ref = O.evaluate(sh, alpha, pts)
        absref = O.evaluate(sh, np.abs(alpha), pts)
        assert (np.abs(o5 - ref) <= 1e-12 * absref + 1e-300).all()


@pytest.mark.parametrize("d,n,prec", [(12, 5, "f64"), (10, 6, "f32"), (6, 4, "f64")])
def test_two_points_per_lane_falls_back(sg, torch_gpu, d, n, prec):
    """Where variant 5 is not built (the sparse order, FP32, dimensions without the
    block-staged walk) it runs as the variant below it: outputs equal variant 4's."""
    torch = torch_gpu
    sh = O.tables(d, n)
    alpha = O.hierarchize(sh, W.f_gauss(O.grid_coords(sh)))
    g = sg.Grid(d, n)
    pts = W.uniform_points(700, d, seed=d)
    dt = torch.float64 if prec == "f64" else torch.float32
    assert np.array_equal(evaluate(torch, g, alpha, pts, dt, 5),
                          evaluate(torch, g, alpha, pts, dt, BLK))


def test_blk_ragged_counts(sg, torch_gpu):
    torch = torch_gpu
    d, n = 12, 5
    sh = O.tables(d, n)
    alpha = O.hierarchize(sh, W.f_gauss(O.grid_coords(sh)))
    g = sg.Grid(d, n)
    for P in (0, 1, 63, 255, 256, 257, 511, 513):
        pts = W.uniform_points(max(P, 1), d, seed=P)[:P]
        absref = O.evaluate(sh, np.abs(alpha), pts) if P else np.zeros(0)
        for dt, prec in ((torch.float64, "f64"), (torch.float32, "f32")):
            o4 = evaluate(torch, g, alpha, pts, dt, BLK)
            o0 = evaluate(torch, g, alpha, pts, dt, 0)
            assert same_walk_result(o4, o0, d, prec, absref), (P, dt)


@pytest.mark.parametrize("d,n,prec", [(10, 8, "f32"), (20, 6, "f64")])
def test_blk_interpolates_at_every_grid_point(sg, torch_gpu, d, n, prec):
    """Size-independent check at full size: the interpolant equals the nodal values at
    every grid point (FP64: 1e-12 sum|terms|; FP32 Kahan on FP32 surpluses: 2e-6 sum|terms|
    + 1e-7 against the FP64 nodal values)."""
    torch = torch_gpu
    g = sg.Grid(d, n)
    dt = torch.float64 if prec == "f64" else torch.float32
    g.set_tuning(eval_variant_f64=BLK, eval_variant_f32=BLK)
    x = torch.empty((g.N, d), dtype=torch.float64, device="cuda")
    g.grid_coords(x)
    f = torch.exp(-((x - 0.5) ** 2).sum(1))
    a = f.clone()
    g.hierarchize(a)
    out = torch.empty(g.N, dtype=dt, device="cuda")
    g.evaluate(a.to(dt), x.to(dt), out)
    absu = torch.empty(g.N, dtype=torch.float64, device="cuda")
    g.evaluate(a.abs(), x, absu)
    err = (out.double() - f).abs()
    bound = 1e-12 * absu + 1e-300 if prec == "f64" else 2e-6 * absu + 1e-7
    assert bool((err <= bound).all()), float(err.max())


# ----------------------------------------------------------- staged-pattern probe --

def probe_weight(P, seed):
    M = 0xFFFFFFFF
    w = np.zeros(P)
    for p in range(P):
        h = ((p * 0x9E3779B9) & M) ^ seed
        h ^= h >> 15
        h = (h * 0x85EBCA6B) & M
        h ^= h >> 13
        w[p] = 1.0 + (h & 1023) / 4096.0
    return w


@pytest.mark.parametrize("d,n,prec,variant", [(12, 4, "f64", BLK), (8, 5, "f32", BLK),
                                              (6, 5, "f64", STAGED), (10, 8, "f64", STAGED),
                                              (10, 8, "f32", BLK), (20, 6, "f64", BLK),
                                              (10, 8, "f64", 5), (9, 4, "f64", 5),
                                              (10, 8, "f64", BLK), (9, 8, "f32", BLK)])
def test_stage_probe_reads_once_per_subspace(sg, torch_gpu, d, n, prec, variant):
    """On all-ones surpluses every read of the probe returns 1 exactly when it reads a
    copied surplus (LDS cells of a stage that its copies filled, or a global cell of a
    gathered leaf), so each point's sum is S times its weight: the probe reads once per
    subspace per point, all from real surplus data.  (Sums stay exact: weights have 13
    significant bits; FP32 only on grids with S < 2^11.)"""
    torch = torch_gpu
    g = sg.Grid(d, n)
    g.set_tuning(eval_variant_f64=variant, eval_variant_f32=variant)
    dt = torch.float64 if prec == "f64" else torch.float32
    P = 777
    out = torch.empty(P, dtype=dt, device="cuda")
    sg.stage_probe(g, torch.ones(g.N, dtype=dt, device="cuda"), out, seed=5)
    got = host(out).astype(np.float64)
    want = probe_weight(P, 5) * g.S
    if prec == "f32" and g.S >= 2048:
        # FP32 sums of ~S terms near S w: each add rounds by up to ulp(S w) / 2
        assert np.allclose(got, want, rtol=g.S * 2.0 ** -24, atol=0)
    else:
        assert np.array_equal(got, want), float(np.max(np.abs(got - want)))


def test_blk_in_batches_and_the_server(sg, torch_gpu):
    """Work-list launches (batches, their graph, the server) run the block-staged walk
    for the grids tuned to it, grouped by n; outputs equal the per-grid calls.  The
    n = 8 grids at d = 9..11 run it in 512-thread workgroups with work items of 512
    points (ragged counts around that size); the d = 11 one is tuned to a gathering walk
    in FP32, so in a batch its FP64 walk falls back to the staged one."""
    torch = torch_gpu
    shapes = [(12, 5), (11, 4), (9, 6), (12, 6), (8, 5), (3, 4), (10, 8), (9, 8), (11, 8)]
    grids, vals = [], []
    for d, n in shapes:
        sh = O.tables(d, n)
        g = sg.Grid(d, n)
        g.set_tuning(eval_variant_f64=BLK, eval_variant_f32=2 if (d, n) == (11, 8) else BLK)
        grids.append(g)
        vals.append(W.f_gauss(O.grid_coords(sh)))
    counts = [300, 1000, 257, 64, 700, 5, 1100, 513, 900]
    b = sg.Batch(grids)
    b.set_points(counts)
    nodal = torch.from_numpy(np.concatenate(vals)).cuda()
    coords = torch.empty(int(b.coord_offsets[-1]), dtype=torch.float64, device="cuda")
    sg.uniform_points(coords, seed=21)
    for dt in (torch.float64, torch.float32):
        a = nodal.to(dt).clone()
        b.hierarchize(a)
        out = torch.full((int(b.point_offsets[-1]),), float("nan"), dtype=dt, device="cuda")
        b.run_graph(a.clone(), coords.to(dt), out, hierarchize=False)
        torch.cuda.synchronize()
        for i, g in enumerate(grids):
            lo, hi = int(b.value_offsets[i]), int(b.value_offsets[i + 1])
            c0, c1 = int(b.coord_offsets[i]), int(b.coord_offsets[i + 1])
            p0, p1 = int(b.point_offsets[i]), int(b.point_offsets[i + 1])
            oi = torch.empty(p1 - p0, dtype=dt, device="cuda")
            g.evaluate(a[lo:hi], coords.to(dt)[c0:c1].view(p1 - p0, g.d), oi)
            assert torch.equal(oi, out[p0:p1]), (i, dt)
        prec = "f64" if dt == torch.float64 else "f32c"
        srv = sg.Server(prec, slots=2, max_batch_points=1 << 13, max_requests=32)
        ids = [srv.add_grid(g, a[int(b.value_offsets[i]):int(b.value_offsets[i + 1])])
               for i, g in enumerate(grids)]
        npdt = np.float64 if dt == torch.float64 else np.float32
        reqs = []
        for k in range(18):
            i = k % len(grids)
            pts = W.uniform_points(50 + 37 * k, grids[i].d, seed=300 + k).astype(npdt)
            o = np.full(len(pts), np.nan, npdt)
            reqs.append((i, pts, o, srv.submit(ids[i], pts, o)))
        for i, pts, o, rid in reqs:
            srv.wait(rid)
            g = grids[i]
            ref = torch.empty(len(pts), dtype=dt, device="cuda")
            g.evaluate(a[int(b.value_offsets[i]):int(b.value_offsets[i + 1])],
                       torch.from_numpy(pts).cuda(), ref)
            assert np.array_equal(o, ref.cpu().numpy()), (i, prec)
        srv.close()
