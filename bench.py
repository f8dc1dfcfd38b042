"""Benchmark of the compact sparse grid hot path on one MI355X.

Workload (the drop-in FP64 configuration): regular sparse grid d=10, n=8 (N=1,862,145
grid points, S=19,448 subspaces), nodal values f(x) = exp(-sum_i (x_i - 1/2)^2) sampled
at the grid points, P = 2^24 seeded uniform query points (k * 2^-24 lattice), FP64.

One step = restore the nodal values into the grid buffer, hierarchize them (in place),
evaluate the interpolant at the P points.  Inputs are resident in HBM before the timed
region.  ``value`` is evaluated points per second over the whole timed steps;
hierarchized grid points per second and the per-kernel times come from HIP events
recorded on the stream the kernels run on.

Prints one JSON line (see the fields below).  ``--gpus`` above 1 is refused: this build
runs on one GPU.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "eval pts/s & hierarchize GP/s at d=10,n=8; served p50/p99 vs plain FP64"
HBM_PEAK_GBS = 8000.0
VARIANT_NAMES = {0: "gather, 3 tabled dims, batches of 8", 1: "gather, 2 tabled dims",
                 2: "gather, batches of 4, 4 waves/SIMD",
                 3: "staged in LDS, 3 tabled dims, FP64 batches of 4 from d=9 (8 below), "
                    "dimension 3 unrolled below small budgets"}


def cpu_baseline(d: int, n: int, P_sample: int, seed: int) -> dict:
    """The reference's own OpenMP/SSE CPU path (omp_evaluate5 and omp_hierarchize3 of
    adaptive_sparse_grid_bench.c, compiled from source into oracle/_ref/), timed on this
    host's cores on a bounded sample of the same workload; the oracle's FP64 restatement
    of evaluate0 if that library is absent."""
    import numpy as np
    from adaptive_sparse_grid_bench_amd import workloads as W
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
    pts = W.uniform_points(P_sample, d, seed=seed)
    try:
        from oracle import refbind
        if not refbind.available():
            raise ImportError("oracle/_ref/libasgref.so not built")
        g = refbind.RefGrid(d, n)
        x = pts.astype(np.float32)
        vals = g.get_values()
        hier = []
        for _ in range(5):
            g.set_values(vals)
            t0 = time.perf_counter()
            g.hierarchize("omp_hierarchize3")
            hier.append(time.perf_counter() - t0)
        g.evaluate(x[: min(P_sample, 256)], "omp_evaluate5")          # warm-up
        t0 = time.perf_counter()
        g.evaluate(x, "omp_evaluate5")
        dt = time.perf_counter() - t0
        return {"value": P_sample / dt, "unit": "points/s", "cores": threads,
                "kind": "reference",
                "sample": f"reference omp_evaluate5 (FP32, as the reference computes) on "
                          f"{P_sample} of the workload's query points at d={d} n={n}; "
                          f"omp_hierarchize3 median of 5 on the full grid",
                "hierarchize_gp_per_s": g.N / sorted(hier)[2],
                "host_cpu": _cpu_model()}
    except (ImportError, OSError) as exc:
        from oracle import oracle as O
        sh = O.tables(d, n)
        alpha = O.hierarchize(sh, W.f_gauss(O.grid_coords(sh)))
        t0 = time.perf_counter()
        O.evaluate(sh, alpha, pts)
        dt = time.perf_counter() - t0
        return {"value": P_sample / dt, "unit": "points/s", "cores": threads, "kind": "port",
                "sample": f"oracle restatement of evaluate0 (FP64) on {P_sample} points "
                          f"at d={d} n={n} ({exc})", "host_cpu": _cpu_model()}


def precision_form(args) -> dict:
    """d=20, n=6 (N=1,515,009): a seeded sum of 8 anisotropic Gaussians (centres U[0.2,
    0.8]^20, widths log-uniform [0.05, 0.5]) hierarchized in FP64 on the GPU, evaluated at
    2^24 uniform points in FP32 with Kahan accumulation and in FP64; max and RMS error of
    the FP32 result against the FP64 walk (which the tests hold to the CPU FP64 oracle
    within 1e-12 relative to sum |terms|)."""
    import torch
    from adaptive_sparse_grid_bench_amd import sgrid
    from adaptive_sparse_grid_bench_amd import workloads as W
    d, n, P = 20, 6, 1 << args.log2_points
    g = sgrid.Grid(d, n)
    x = torch.empty((g.N, d), dtype=torch.float64, device="cuda")
    g.grid_coords(x)
    c, w = W.gaussian_sum_params(d, 8, seed=7)
    f = torch.zeros(g.N, dtype=torch.float64, device="cuda")
    for cc, ww in zip(c, w):
        cc = torch.from_numpy(cc).cuda()
        ww = torch.from_numpy(ww).cuda()
        f += torch.exp(-(((x - cc) ** 2) / (2 * ww * ww)).sum(1))
    del x
    g.hierarchize(f)
    pts = torch.empty((P, d), dtype=torch.float64, device="cuda")
    sgrid.uniform_points(pts, seed=11)
    o64 = torch.empty(P, dtype=torch.float64, device="cuda")
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(4)]
    ev[0].record()
    g.evaluate(f, pts, o64)
    ev[1].record()
    a32, p32 = f.float(), pts.float()
    del pts
    o32 = torch.empty(P, dtype=torch.float32, device="cuda")
    ev[2].record()
    g.evaluate(a32, p32, o32)
    ev[3].record()
    torch.cuda.synchronize()
    err = (o32.double() - o64).abs()
    umax = o64.abs().max().item()
    res = {"workload": f"d=20 n=6 (N={g.N}, S={g.S}), sum of 8 Gaussians, 2^{args.log2_points} "
                       "points; FP32 Kahan vs FP64 walk",
           "f64_ms": ev[0].elapsed_time(ev[1]), "f32c_ms": ev[2].elapsed_time(ev[3]),
           "max_abs_err": err.max().item(), "rms_err": err.pow(2).mean().sqrt().item(),
           "max_abs_u": umax}
    res["max_rel_to_max_u"] = res["max_abs_err"] / umax if umax > 0 else None
    del o64, o32, a32, p32, f, err
    torch.cuda.empty_cache()
    return res


def other_forms(g, alpha, pts, out64, P, args) -> dict:
    """FP32-Kahan evaluation of the same workload, the batched form (256 grids, one
    hipGraph launch vs per-grid launches) and the served form (continuous batching vs
    plain FP64 per-request launches, p50/p99)."""
    import torch
    from adaptive_sparse_grid_bench_amd import batched, serve
    res = {}
    a32, p32 = alpha.float(), pts.float()
    o32 = torch.empty(P, dtype=torch.float32, device=out64.device)
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    g.evaluate(a32, p32, o32)
    torch.cuda.synchronize()
    ev[0].record()
    g.evaluate(a32, p32, o32)
    ev[1].record()
    torch.cuda.synchronize()
    ms = ev[0].elapsed_time(ev[1])
    err = (o32.double() - out64).abs()
    res["f32c"] = {"evaluate_ms": ms, "pts_per_s": P / (ms * 1e-3),
                   "max_abs_err_vs_f64": err.max().item(),
                   "rms_err_vs_f64": err.pow(2).mean().sqrt().item()}
    del a32, p32, o32
    res["precision"] = precision_form(args)
    parts = batched.build(256, tune=True)
    res["batched"] = batched.measure(*parts)
    del parts
    grids = serve.make_grids(64)
    res["served"] = serve.experiment(grids, args.served_requests, 0.8)
    del grids
    torch.cuda.empty_cache()
    return res


def _cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--d", type=int, default=10)
    ap.add_argument("--n", type=int, default=8)
    ap.add_argument("--log2-points", type=int, default=24)
    ap.add_argument("--seed", type=int, default=42)
    ap.add_argument("--cpu-sample", type=int, default=1 << 16)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-forms", action="store_true",
                    help="skip the FP32, batched and served forms (headline FP64 line only)")
    ap.add_argument("--served-requests", type=int, default=600)
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run one plain device copy of MIB MiB (the reference kernel "
                         "for scaling rocprofv3 FETCH_SIZE / WRITE_SIZE) and exit")
    args = ap.parse_args()

    if args.copy_calibration:
        import torch
        nbytes = args.copy_calibration << 20
        src = torch.ones(nbytes // 8, dtype=torch.float64, device="cuda:0")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes_read": nbytes, "bytes_written": nbytes}))
        return 0

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path "
              "(evaluation would shard by query points with the grid replicated).",
              file=sys.stderr)
        return 2

    import torch
    from adaptive_sparse_grid_bench_amd import sgrid

    d, n, P = args.d, args.n, 1 << args.log2_points
    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    stream = torch.cuda.current_stream()
    g = sgrid.Grid(d, n)
    x = torch.empty((g.N, d), dtype=torch.float64, device=dev)
    g.grid_coords(x)
    nodal = torch.exp(-((x - 0.5) ** 2).sum(1))
    del x
    alpha = torch.empty_like(nodal)
    pts = torch.empty((P, d), dtype=torch.float64, device=dev)
    sgrid.uniform_points(pts, seed=args.seed)
    out = torch.empty(P, dtype=torch.float64, device=dev)
    torch.cuda.synchronize()

    def step(events=None):
        alpha.copy_(nodal)
        if events:
            events[0].record(stream)
        g.hierarchize(alpha, stream)
        if events:
            events[1].record(stream)
        g.evaluate(alpha, pts, out, stream)
        if events:
            events[2].record(stream)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    evs = [[torch.cuda.Event(enable_timing=True) for _ in range(3)] for _ in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        step(evs[k])
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0

    hier_ms = sorted(e[0].elapsed_time(e[1]) for e in evs)
    eval_ms = sorted(e[1].elapsed_time(e[2]) for e in evs)
    eval_avg = sum(eval_ms) / len(eval_ms)
    hier_med = hier_ms[len(hier_ms) // 2]
    ms_per_step = 1e3 * wall / args.steps
    value = P * args.steps / wall

    # algorithmic bytes of the evaluate kernel: one coefficient gather per subspace,
    # the point's coordinates and its output value
    bytes_per_point = g.S * 8 + d * 8 + 8
    achieved = P * bytes_per_point / (eval_avg * 1e-3) / 1e9
    # HBM bytes per launch of this kernel, measured with rocprofv3 counters
    # (FETCH_SIZE scaled by a plain copy's reading, plus WRITE_SIZE)
    traffic = None
    tpath = os.path.join(ROOT, "profiles", f"eval_d{d}n{n}_traffic.json")
    if os.path.exists(tpath):
        try:
            tj = json.load(open(tpath))
            if tj.get("num_points") == P:
                traffic = tj.get("hbm_bytes_per_launch")
        except (OSError, ValueError):
            traffic = None
    variant = g.get_tuning()["eval_variant_f64"]

    # achievable gather rate of this grid's shape: one random in-block 8-byte gather per
    # subspace per point, blocks in level-sum order, no hats (the gather roofline)
    probe_out = torch.empty(P, dtype=torch.float64, device=dev)
    pev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    probe_ms = []
    for _ in range(3):
        pev[0].record(stream)
        sgrid.gather_probe(g, alpha, probe_out, seed=7, stream=stream)
        pev[1].record(stream)
        torch.cuda.synchronize()
        probe_ms.append(pev[0].elapsed_time(pev[1]))
    del probe_out
    probe_best = min(probe_ms[1:])
    gather_bytes = P * g.S * 8
    achievable = gather_bytes / (probe_best * 1e-3) / 1e9

    rec = {
        "metric": METRIC,
        "value": value,
        "unit": "points/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": f"regular sparse grid d={d} n={n} (N={g.N}, S={g.S}), "
                        f"f=exp(-sum (x-1/2)^2), hierarchize + evaluate 2^{args.log2_points} "
                        f"seeded uniform points, FP64",
            "d": d, "n": n, "num_grid_points": g.N, "num_subspaces": g.S,
            "num_points": P, "parallelism": "single GPU",
        },
        "hierarchize_gp_per_s": g.N / (hier_med * 1e-3),
        "hierarchize_ms": hier_med,
        # algorithmic bytes of a full hierarchization: every value read and written once
        # per dimension sweep (2 d N 8 bytes); the pair kernel moves each value once per
        # two dimensions, so its HBM bytes are about half of these
        "hierarchize_roofline": {
            "kernels": "k_hierarchize_pair (two dimensions per launch), k_hierarchize_dim",
            "bound": "hbm",
            "algorithmic_bytes": 2 * d * g.N * 8,
            "achieved": 2 * d * g.N * 8 / (hier_med * 1e-3) / 1e9,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": 2 * d * g.N * 8 / (hier_med * 1e-3) / 1e9 / HBM_PEAK_GBS,
        },
        "evaluate_ms": eval_avg,
        "evaluate_pts_per_s_kernel": P / (eval_avg * 1e-3),
        "roofline": {
            "kernel": f"k_eval<{d},...,double> walk variant {variant} ({VARIANT_NAMES[variant]})",
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "bytes_per_point": bytes_per_point,
            "achievable_gather": {
                "GBps": achievable,
                "probe_ms": probe_best,
                "evaluate_gather_GBps": gather_bytes / (eval_avg * 1e-3) / 1e9,
                "frac": probe_best / eval_avg,
                "pattern": "sg_gather_probe_f64: per point one random 8-byte gather inside "
                           "every subspace block, level-sum order, 8 in flight per lane",
            },
        },
        "cpu_baseline": None,
    }
    if not args.no_forms:
        rec.update(other_forms(g, alpha, pts, out, P, args))
    if not args.no_cpu_baseline:
        rec["cpu_baseline"] = cpu_baseline(d, n, args.cpu_sample, args.seed)
    print(json.dumps(rec), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
