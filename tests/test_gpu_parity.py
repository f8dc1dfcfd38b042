"""GPU parity of libsgrid.so against the CPU oracle and the reference's golden vectors.

Bars (each stated where it is tested):
  * index bijection, grid coordinates, hierarchization (FP64 and FP32), the
    reference-order evaluate and the point generator: bit-exact;
  * FP64 evaluate (depth-first walk, different summation order):
        |gpu - oracle| <= 1e-12 * sum_j |term_j|   (+1e-300),
    where sum_j |term_j| = interpolant of |alpha| (every hat is >= 0);
  * FP32 Kahan evaluate: error vs the FP64 oracle stated per test.
"""
import numpy as np
import pytest

from oracle import oracle as O
from adaptive_sparse_grid_bench_amd import workloads as W

pytestmark = pytest.mark.gpu

REL_F64 = 1e-12


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


SHAPES = [(1, 1, None), (1, 7, None), (2, 3, None), (3, 5, None), (2, 3, [2, 1]),
          (4, 5, [4, 1, 4, 2]), (5, 4, [3, 3, 1, 2, 3]), (10, 8, None), (20, 6, None)]


def ids(shapes):
    return [f"d{d}n{n}" + ("" if l is None else "lim" + "".join(map(str, l))) for d, n, l in shapes]


# ------------------------------------------------------------------- bijection --

@pytest.mark.parametrize("d,n,lim", SHAPES, ids=ids(SHAPES))
def test_idx2gp_gp2idx_exhaustive(sg, torch_gpu, d, n, lim):
    torch = torch_gpu
    sh = O.tables(d, n, lim)
    g = sg.Grid(d, n, lim)
    assert (g.N, g.S) == (sh.N, sh.S)
    lv = torch.empty((sh.N, d), dtype=torch.int32, device="cuda")
    iv = torch.empty_like(lv)
    g.idx2gp_dev(lv, iv)
    rl, ri = O.idx2gp_all(sh)
    assert (host(lv) == rl).all() and (host(iv) == ri).all()
    idx = torch.empty(sh.N, dtype=torch.int32, device="cuda")
    g.gp2idx_dev(lv, iv, idx, sh.N)
    assert (host(idx) == np.arange(sh.N)).all()


def test_bijection_invalid_inputs(sg, torch_gpu):
    torch = torch_gpu
    g = sg.Grid(3, 4, [3, 1, 2])
    offs = torch.tensor([-1, 0, g.N - 1, g.N, 2 ** 30], dtype=torch.int32, device="cuda")
    lv = torch.empty((5, 3), dtype=torch.int32, device="cuda")
    iv = torch.empty_like(lv)
    g.idx2gp_dev(lv, iv, index=offs, count=5)
    L = host(lv)
    assert (L[0] == -1).all() and (L[3] == -1).all() and (L[4] == -1).all()
    assert (L[1] == 0).all()
    bad_l = dev(torch, np.array([[0, 2, 0], [1, 0, 0], [4, 0, 0], [1, 1, 2]], np.int32))
    bad_i = dev(torch, np.array([[0, 0, 0], [2, 0, 0], [0, 0, 0], [0, 0, 0]], np.int32))
    out = torch.empty(4, dtype=torch.int32, device="cuda")
    g.gp2idx_dev(bad_l, bad_i, out, 4)
    # l_1 = 2 > limit 1; i_0 = 2 >= 2^1; l_0 = 4 > limit 3; level sum 4 >= n
    assert (host(out) == -1).all()


@pytest.mark.parametrize("tag", ["d10n8_sampled", "d20n6_sampled"])
def test_bijection_vs_golden_sampled(sg, torch_gpu, golden, tag):
    torch = torch_gpu
    gd = golden(tag)
    g = sg.Grid(int(gd["d"]), int(gd["n"]))
    offs = dev(torch, gd["offsets"].astype(np.int32))
    m = len(gd["offsets"])
    lv = torch.empty((m, g.d), dtype=torch.int32, device="cuda")
    iv = torch.empty_like(lv)
    g.idx2gp_dev(lv, iv, index=offs, count=m)
    assert (host(lv) == gd["levels"]).all() and (host(iv) == gd["indices"]).all()


@pytest.mark.parametrize("d,n,lim", SHAPES[3:8], ids=ids(SHAPES[3:8]))
def test_grid_coords(sg, torch_gpu, d, n, lim):
    torch = torch_gpu
    sh = O.tables(d, n, lim)
    g = sg.Grid(d, n, lim)
    c = torch.empty((g.N, d), dtype=torch.float64, device="cuda")
    g.grid_coords(c)
    assert (host(c) == O.grid_coords(sh)).all()


def test_uniform_points_match_host_generator(sg, torch_gpu):
    torch = torch_gpu
    for dt, npdt in ((torch.float64, np.float64), (torch.float32, np.float32)):
        t = torch.empty((4099, 7), dtype=dt, device="cuda")
        sg.uniform_points(t, seed=42)
        assert (host(t) == W.uniform_points(4099, 7, seed=42, dtype=npdt)).all()


# ----------------------------------------------------------------- hierarchize --

def sample(sh, fname="gauss"):
    x = O.grid_coords(sh)
    if fname == "gauss":
        return W.f_gauss(x)
    if fname == "parabola":
        return W.f_prod_parabola(x)
    return W.f_gauss(x) - 3.0 * W.f_reference(x)


@pytest.mark.parametrize("d,n,lim", SHAPES, ids=ids(SHAPES))
@pytest.mark.parametrize("fname", ["gauss", "mixed"])
def test_hierarchize_f64_bit_exact(sg, torch_gpu, d, n, lim, fname):
    torch = torch_gpu
    sh = O.tables(d, n, lim)
    v = sample(sh, fname)
    g = sg.Grid(d, n, lim)
    t = dev(torch, v)
    g.hierarchize(t)
    assert (host(t) == O.hierarchize(sh, v)).all()


def test_hierarchize_parabola_closed_form(sg, torch_gpu):
    torch = torch_gpu
    sh = O.tables(6, 6)
    lv, _ = O.idx2gp_all(sh)
    g = sg.Grid(6, 6)
    t = dev(torch, sample(sh, "parabola"))
    g.hierarchize(t)
    assert (host(t) == 4.0 ** -lv.sum(1)).all()


@pytest.mark.parametrize("tag", ["d2n3", "d3n5", "d4n4", "d5n3", "d2n6", "d2n3_lim21",
                                 "d3n5_lim423", "d4n5_lim4142"])
def test_hierarchize_f32_and_evaluate_ref_f32_vs_golden(sg, torch_gpu, golden, tag):
    """FP32 bit-exact against the reference's own hierarchize0 and evaluate0."""
    torch = torch_gpu
    gd = golden(tag)
    g = sg.Grid(int(gd["d"]), int(gd["n"]), gd["limits"])
    for fname in ("reference", "parabola", "gauss"):
        t = dev(torch, gd[f"values_{fname}"])
        g.hierarchize(t)
        assert (host(t) == gd[f"alpha_{fname}"]).all()
        pts = dev(torch, gd["points"])
        out = torch.empty(len(gd["points"]), dtype=torch.float32, device="cuda")
        g.evaluate(t, pts, out, mode="ref")
        assert (host(out) == gd[f"eval_{fname}"]).all()


@pytest.mark.parametrize("tile", [64, 256, 1024, 4096])
def test_hierarchize_tile_sizes(sg, torch_gpu, tile):
    torch = torch_gpu
    sh = O.tables(7, 7, [6, 2, 6, 3, 6, 1, 6])
    v = sample(sh, "mixed")
    g = sg.Grid(7, 7, [6, 2, 6, 3, 6, 1, 6])
    g.set_tuning(hier_tile_values=tile)
    t = dev(torch, v)
    g.hierarchize(t)
    assert (host(t) == O.hierarchize(sh, v)).all()


LARGE_POLES = [(1, 15, None), (2, 14, None), (3, 14, [13, 1, 2])]


@pytest.mark.parametrize("d,n,lim", LARGE_POLES, ids=ids(LARGE_POLES))
def test_hierarchize_pole_families_beyond_lds(sg, torch_gpu, d, n, lim):
    """A dimension reaching level 13 or 14 with the other levels low makes pole families
    of 16,383-32,767 values, more than a sweep tile holds in LDS: those sweeps run in
    global memory.  Bit-exact to the oracle in FP64 and FP32, alone and in a batch with
    an ordinary grid."""
    torch = torch_gpu
    sh = O.tables(d, n, lim)
    v = sample(sh, "mixed")
    g = sg.Grid(d, n, lim)
    for dt in (np.float64, np.float32):
        vv = v.astype(dt)
        t = dev(torch, vv)
        g.hierarchize(t)
        assert (host(t) == O.hierarchize(sh, vv)).all(), dt
    small = sg.Grid(3, 4)
    sh_s = O.tables(3, 4)
    v_s = sample(sh_s, "mixed")
    b = sg.Batch([small, g])
    b.set_points([1, 1])
    t = dev(torch, np.concatenate([v_s, v]))
    b.hierarchize(t)
    out = host(t)
    assert (out[:small.N] == O.hierarchize(sh_s, v_s)).all()
    assert (out[small.N:] == O.hierarchize(sh, v)).all()


# -------------------------------------------------------------------- evaluate --

def check_f64(out, ref, absref):
    err = np.abs(out - ref)
    bound = REL_F64 * absref + 1e-300
    assert (err <= bound).all(), f"max err {err.max()} rel {np.max(err / (absref + 1e-300))}"


EVAL_SHAPES = [(1, 1, None, 100), (1, 7, None, 1000), (2, 3, None, 777), (3, 5, None, 5000),
               (2, 3, [2, 1], 300), (4, 5, [4, 1, 4, 2], 2000), (5, 4, [3, 3, 1, 2, 3], 999),
               (10, 8, None, 2048), (20, 6, None, 300), (12, 8, None, 200),
               (16, 4, None, 500), (5, 11, [10, 1, 2, 1, 10], 600)]


@pytest.mark.parametrize("d,n,lim,P", EVAL_SHAPES, ids=[f"d{a}n{b}P{p}" for a, b, _, p in EVAL_SHAPES])
def test_evaluate_f64_vs_oracle(sg, torch_gpu, d, n, lim, P):
    torch = torch_gpu
    sh = O.tables(d, n, lim)
    alpha = O.hierarchize(sh, sample(sh, "mixed"))
    pts = W.uniform_points(P, d, seed=1000 + d)
    pts[0] = 0.0                       # x = 0 exactly in every dimension
    ref = O.evaluate(sh, alpha, pts)
    absref = O.evaluate(sh, np.abs(alpha), pts)
    g = sg.Grid(d, n, lim)
    a, x = dev(torch, alpha), dev(torch, pts)
    out = torch.full((P,), np.nan, dtype=torch.float64, device="cuda")
    g.evaluate(a, x, out)
    check_f64(host(out), ref, absref)
    # the reference-order kernel is bit-identical to evaluate0 in FP64
    g.evaluate(a, x, out, mode="ref")
    assert (host(out) == ref).all()


@pytest.mark.parametrize("variant", range(5))
def test_headline_gaussian_relative_to_u(sg, torch_gpu, variant):
    """The headline workload (d=10, n=8, f = exp(-sum (x - 1/2)^2), hierarchized on the
    GPU) meets the FP64 bound as stated for it, relative to the value itself:
    |gpu - oracle| <= 1e-12 |u|, on 4,096 of its seeded points, for every walk variant.
    (Every surplus of this product of 1-D Gaussians is positive, so |u| = sum |terms|.)"""
    torch = torch_gpu
    d, n, P = 10, 8, 4096
    sh = O.tables(d, n)
    g = sg.Grid(d, n)
    x = torch.empty((g.N, d), dtype=torch.float64, device="cuda")
    g.grid_coords(x)
    a = torch.exp(-((x - 0.5) ** 2).sum(1))
    nodal = host(a)
    g.hierarchize(a)
    alpha = host(a)
    assert np.array_equal(alpha, O.hierarchize(sh, nodal))
    assert (alpha > 0).all()
    pts = torch.empty((P, d), dtype=torch.float64, device="cuda")
    sg.uniform_points(pts, seed=42)
    ref = O.evaluate(sh, alpha, host(pts))
    g.set_tuning(eval_variant_f64=variant)
    out = torch.full((P,), np.nan, dtype=torch.float64, device="cuda")
    g.evaluate(a, pts, out)
    err = np.abs(host(out) - ref)
    assert (err <= 1e-12 * np.abs(ref)).all(), float(np.max(err / np.abs(ref)))


@pytest.mark.parametrize("block", [64, 128, 256])
def test_evaluate_block_sizes_and_ragged_counts(sg, torch_gpu, block):
    torch = torch_gpu
    sh = O.tables(4, 6)
    alpha = O.hierarchize(sh, sample(sh, "mixed"))
    g = sg.Grid(4, 6)
    g.set_tuning(eval_block=block)
    a = dev(torch, alpha)
    for P in (0, 1, 63, 64, 65, 257, 1000):
        pts = W.uniform_points(max(P, 1), 4, seed=P)[:P]
        out = torch.full((P,), np.nan, dtype=torch.float64, device="cuda")
        g.evaluate(a, dev(torch, pts), out)
        if P:
            check_f64(host(out), O.evaluate(sh, alpha, pts), O.evaluate(sh, np.abs(alpha), pts))


def test_evaluate_interpolates_full_d10n8(sg, torch_gpu):
    """At every one of the 1,862,145 grid points the interpolant equals f."""
    torch = torch_gpu
    g = sg.Grid(10, 8)
    x = torch.empty((g.N, 10), dtype=torch.float64, device="cuda")
    g.grid_coords(x)
    f = torch.exp(-((x - 0.5) ** 2).sum(1))
    a = f.clone()
    g.hierarchize(a)
    out = torch.empty_like(f)
    g.evaluate(a, x, out)
    err = (out - f).abs().max().item()
    assert err <= 1e-13, err


def test_evaluate_linearity_large(sg, torch_gpu):
    """u[a1 + a2] = u[a1] + u[a2] on 2^18 points at d=10 n=8 (size-independent check)."""
    torch = torch_gpu
    g = sg.Grid(10, 8)
    x = torch.empty((g.N, 10), dtype=torch.float64, device="cuda")
    g.grid_coords(x)
    a1 = torch.exp(-((x - 0.5) ** 2).sum(1))
    a2 = torch.sin(3.0 * x).prod(1)
    g.hierarchize(a1)
    g.hierarchize(a2)
    P = 1 << 18
    pts = torch.empty((P, 10), dtype=torch.float64, device="cuda")
    sg.uniform_points(pts, seed=9)
    o1, o2, o12, oabs = (torch.empty(P, dtype=torch.float64, device="cuda") for _ in range(4))
    g.evaluate(a1, pts, o1)
    g.evaluate(a2, pts, o2)
    g.evaluate(a1 + a2, pts, o12)
    g.evaluate(a1.abs() + a2.abs(), pts, oabs)
    assert ((o12 - o1 - o2).abs() <= 3 * REL_F64 * oabs + 1e-300).all()


def test_evaluate_f32c_error_bound(sg, torch_gpu):
    """FP32 surpluses and points with Kahan accumulation vs the FP64 oracle at d=10 n=8."""
    torch = torch_gpu
    sh = O.tables(10, 8)
    alpha = O.hierarchize(sh, sample(sh, "gauss"))
    pts = W.uniform_points(2048, 10, seed=77)
    ref = O.evaluate(sh, alpha, pts)
    g = sg.Grid(10, 8)
    out = torch.empty(2048, dtype=torch.float32, device="cuda")
    g.evaluate(dev(torch, alpha.astype(np.float32)), dev(torch, pts.astype(np.float32)), out)
    err = np.abs(host(out).astype(np.float64) - ref)
    print("f32c max err", err.max(), "rms", np.sqrt((err ** 2).mean()))
    assert err.max() <= 1e-6


# ------------------------------------------------ variants, binning, explicit sets --

@pytest.mark.parametrize("variant", range(4))
@pytest.mark.parametrize("binning", [-1, 1])
@pytest.mark.parametrize("d,n,P", [(10, 8, 3000), (5, 6, 5000), (2, 8, 4000)])
def test_evaluate_variants_and_binning(sg, torch_gpu, variant, binning, d, n, P):
    """Every compiled walk variant, with and without spatial binning, in FP64 and FP32."""
    torch = torch_gpu
    sh = O.tables(d, n)
    alpha = O.hierarchize(sh, sample(sh, "mixed"))
    pts = W.uniform_points(P, d, seed=31 + d)
    ref = O.evaluate(sh, alpha, pts)
    absref = O.evaluate(sh, np.abs(alpha), pts)
    g = sg.Grid(d, n)
    g.set_tuning(eval_variant_f64=variant, eval_variant_f32=variant, bin_min_points=binning)
    out = torch.full((P,), np.nan, dtype=torch.float64, device="cuda")
    g.evaluate(dev(torch, alpha), dev(torch, pts), out)
    check_f64(host(out), ref, absref)
    o32 = torch.full((P,), np.nan, dtype=torch.float32, device="cuda")
    g.evaluate(dev(torch, alpha.astype(np.float32)), dev(torch, pts.astype(np.float32)), o32)
    err = np.abs(host(o32).astype(np.float64) - ref)
    assert (err <= 2e-6 * absref + 1e-7).all(), err.max()


def explicit_sets():
    out = []
    lv = np.array([O.idx2l(O.tables(4, 5, [4, 1, 4, 2]), r, s) for s in range(5)
                   for r in range(int(O.tables(4, 5, [4, 1, 4, 2]).dp.reshape(4, 5)[3, s]))],
                  np.uint8)
    out.append(("regular_as_explicit", lv))
    out.append(("adaptive_d8", W.adaptive_level_set(d=8, base_n=3, extra=4, seed=11)[0]))
    out.append(("adaptive_d3", W.adaptive_level_set(d=3, base_n=3, extra=4, seed=3,
                                                     max_subspaces=60)[0]))
    # dimension 0 refined past level 7: the table-free walk
    deep = [(a, b) for a in range(11) for b in range(3) if a + b <= 10]
    out.append(("deep_dim0", np.array(deep, np.uint8)))
    return out


@pytest.mark.parametrize("name,levels", explicit_sets(), ids=[n for n, _ in explicit_sets()])
def test_explicit_sets(sg, torch_gpu, name, levels):
    torch = torch_gpu
    xs = O.xshape(levels)
    g = sg.Grid.explicit(xs.d, levels)
    assert (g.N, g.S) == (xs.N, xs.S)
    glv, gbase = g.subspaces()
    assert (glv == xs.levels).all() and (gbase == xs.base).all()
    x = O.x_grid_coords(xs)
    # device bijection and coordinates
    c = torch.empty((g.N, xs.d), dtype=torch.float64, device="cuda")
    g.grid_coords(c)
    assert (host(c) == x).all()
    lv = torch.empty((g.N, xs.d), dtype=torch.int32, device="cuda")
    iv = torch.empty_like(lv)
    g.idx2gp_dev(lv, iv)
    idx = torch.empty(g.N, dtype=torch.int32, device="cuda")
    g.gp2idx_dev(lv, iv, idx, g.N)
    assert (host(idx) == np.arange(g.N)).all()
    # hierarchize bit-exact, evaluate within the FP64 bound
    v = W.f_gauss(x) - 0.5 * W.f_reference(x)
    t = dev(torch, v)
    g.hierarchize(t)
    alpha = O.x_hierarchize(xs, v)
    assert (host(t) == alpha).all()
    pts = W.uniform_points(3000, xs.d, seed=17)
    ref = O.x_evaluate(xs, alpha, pts)
    absref = O.x_evaluate(xs, np.abs(alpha), pts)
    out = torch.empty(3000, dtype=torch.float64, device="cuda")
    for binning in (-1, 1):
        g.set_tuning(bin_min_points=binning)
        g.evaluate(t, dev(torch, pts), out)
        check_f64(host(out), ref, absref)
    o32 = torch.empty(3000, dtype=torch.float32, device="cuda")
    g.evaluate(dev(torch, alpha.astype(np.float32)), dev(torch, pts.astype(np.float32)), o32)
    assert (np.abs(host(o32) - ref) <= 2e-6 * absref + 1e-7).all()
