"""Query-point layouts: column-major and the CUDA reference's 32-point tiles
(asg.cu:167-180) converted on the device to the row-major layout the kernels read.
Bar: bit-identical to the numpy transpose of the same points, including a tail tile."""
import numpy as np
import pytest

from adaptive_sparse_grid_bench_amd import workloads as W

pytestmark = pytest.mark.gpu


def tiled32(pts):
    """The reference's init_coord_mat_trans layout: [ceil(P/32)][d][32], padded."""
    P, d = pts.shape
    T = (P + 31) // 32
    out = np.zeros((T, d, 32), pts.dtype)
    for p in range(P):
        out[p // 32, :, p % 32] = pts[p]
    return out


@pytest.mark.parametrize("P,d", [(1, 3), (31, 10), (32, 1), (1000, 20), (4097, 7)])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_layouts_to_row_major(torch_gpu, P, d, dtype):
    torch = torch_gpu
    from adaptive_sparse_grid_bench_amd import sgrid
    pts = W.uniform_points(P, d, seed=P + d).astype(dtype)
    tdt = torch.float64 if dtype == np.float64 else torch.float32
    for name, src in (("row_major", pts), ("column_major", np.ascontiguousarray(pts.T)),
                      ("tiled32", tiled32(pts))):
        out = torch.full((P, d), float("nan"), dtype=tdt, device="cuda")
        sgrid.points_to_row_major(torch.from_numpy(np.ascontiguousarray(src)).cuda(), name, d,
                                  P, out)
        torch.cuda.synchronize()
        assert np.array_equal(out.cpu().numpy(), pts), name


def test_evaluate_from_tiled_points(torch_gpu):
    """Points handed over in the CUDA reference's layout evaluate to the same values."""
    torch = torch_gpu
    from adaptive_sparse_grid_bench_amd import sgrid
    g = sgrid.Grid(5, 5)
    a = torch.empty(g.N, dtype=torch.float64, device="cuda")
    g.sample("gauss", a)
    g.hierarchize(a)
    P = 777
    pts = W.uniform_points(P, 5, seed=3)
    x = torch.empty((P, 5), dtype=torch.float64, device="cuda")
    sgrid.points_to_row_major(torch.from_numpy(tiled32(pts)).cuda(), "tiled32", 5, P, x)
    o1 = torch.empty(P, dtype=torch.float64, device="cuda")
    o2 = torch.empty_like(o1)
    g.evaluate(a, x, o1)
    g.evaluate(a, torch.from_numpy(pts).cuda(), o2)
    torch.cuda.synchronize()
    assert torch.equal(o1, o2)
