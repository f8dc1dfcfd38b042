"""The precision form: FP32 evaluation with Kahan-compensated accumulation across
subspaces, reading FP32 copies of surpluses hierarchized in FP64.

Workload: d=20, n=6 (N=1,515,009, S=53,130); grid values are a seeded sum of 8
anisotropic Gaussians, centres uniform in [0.2, 0.8]^20, widths log-uniform in
[0.05, 0.5]; query points on the 2^-24 lattice (exact in FP32 and FP64).

Stated bound: max |u_f32c - u_f64| <= 1e-6 (absolute; the interpolant's scale here is
O(1)).  Errors are measured against the CPU FP64 oracle directly on a sample of points,
and against the GPU FP64 walk (which agrees with the oracle within 1e-12 relative to
sum |terms|, tested elsewhere) on a large point set; max and RMS are printed."""
import numpy as np
import pytest

from oracle import oracle as O
from adaptive_sparse_grid_bench_amd import workloads as W

pytestmark = pytest.mark.gpu

BOUND = 1e-6


@pytest.fixture(scope="module")
def grid_d20(torch_gpu):
    from adaptive_sparse_grid_bench_amd import sgrid
    torch = torch_gpu
    g = sgrid.Grid(20, 6)
    x = torch.empty((g.N, 20), dtype=torch.float64, device="cuda")
    g.grid_coords(x)
    c, w = W.gaussian_sum_params(20, 8, seed=7)
    f = torch.zeros(g.N, dtype=torch.float64, device="cuda")
    for cc, ww in zip(c, w):
        cc = torch.from_numpy(cc).cuda()
        ww = torch.from_numpy(ww).cuda()
        f += torch.exp(-(((x - cc) ** 2) / (2 * ww * ww)).sum(1))
    del x
    a = f.clone()
    g.hierarchize(a)
    torch.cuda.synchronize()
    return g, f, a


def test_precision_vs_cpu_oracle_sample(grid_d20, torch_gpu):
    torch = torch_gpu
    g, f, a = grid_d20
    sh = O.tables(20, 6)
    alpha = O.hierarchize(sh, f.cpu().numpy())
    assert (alpha == a.cpu().numpy()).all()          # FP64 hierarchization bit-exact
    P = 1024
    pts = W.uniform_points(P, 20, seed=3)
    ref = O.evaluate(sh, alpha, pts)
    o32 = torch.empty(P, dtype=torch.float32, device="cuda")
    g.evaluate(a.float(), torch.from_numpy(pts.astype(np.float32)).cuda(), o32)
    err = np.abs(o32.cpu().numpy().astype(np.float64) - ref)
    print(f"\nprecision d=20 n=6 vs CPU FP64 oracle, {P} points: max {err.max():.3e} "
          f"rms {np.sqrt(np.mean(err ** 2)):.3e} (|u| max {np.abs(ref).max():.3e})")
    assert err.max() <= BOUND


def test_precision_vs_gpu_f64_large(grid_d20, torch_gpu):
    torch = torch_gpu
    g, f, a = grid_d20
    P = 1 << 22
    x = torch.empty((P, 20), dtype=torch.float64, device="cuda")
    from adaptive_sparse_grid_bench_amd import sgrid
    sgrid.uniform_points(x, seed=11)
    o64 = torch.empty(P, dtype=torch.float64, device="cuda")
    g.evaluate(a, x, o64)
    o32 = torch.empty(P, dtype=torch.float32, device="cuda")
    g.evaluate(a.float(), x.float(), o32)
    err = (o32.double() - o64).abs()
    mx, rms = err.max().item(), err.pow(2).mean().sqrt().item()
    print(f"\nprecision d=20 n=6 vs GPU FP64, {P} points: max {mx:.3e} rms {rms:.3e}")
    assert mx <= BOUND
