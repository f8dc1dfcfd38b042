"""Benchmark of the compact sparse grid hot path on one MI355X.

Workload (the drop-in FP64 configuration): regular sparse grid d=10, n=8 (N=1,862,145
grid points, S=19,448 subspaces), nodal values f(x) = exp(-sum_i (x_i - 1/2)^2) sampled
at the grid points, P = 2^24 seeded uniform query points (k * 2^-24 lattice), FP64.

One step = restore the nodal values into the grid buffer, hierarchize them (in place),
evaluate the interpolant at the P points.  Inputs are resident in HBM before the timed
region.  ``value`` is evaluated points per second over the whole timed steps;
hierarchized grid points per second and the per-kernel times come from HIP events
recorded on the stream the kernels run on.  The reference's own throughput convention
(HI_FLOPS = 3 d N per hierarchization, EV_FLOPS = P S (11 d + 2) per evaluation,
adaptive_sparse_grid_bench.c:48-52, printed as GFLOPS at :1795) is reported beside it.

The evaluate kernel's roofline is the staged-pattern probe (sg_stage_probe_*): the same
stage plan, copies and barriers as the walk, one read per subspace per point from LDS
(global memory for the blocks the walk does not stage) and one FMA, no hats and no walk.
The walk does all of that work and more, so probe time / walk time is at most 1.

The CPU leg (rank 0 only) times the reference's own OpenMP/SSE path (omp_evaluate5 and
omp_hierarchize3, compiled from its source into oracle/_ref/) on a bounded sample, and
checks samples of every timed GPU output against the CPU oracle (max and RMS errors).

Prints one JSON line.  ``--gpus`` above 1 is refused: this build runs on one GPU.
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
FP64_PEAK_TFLOPS = 78.6
VARIANT_NAMES = {0: "gather, 3 tabled dims, batches of 8", 1: "gather, 2 tabled dims",
                 2: "gather, batches of 4, 4 waves/SIMD",
                 3: "staged in LDS, 3 tabled dims, leaves at table-given LDS offsets",
                 4: "block-staged in LDS, compile-time n, leaves at compile-time LDS offsets",
                 5: "block-staged in LDS, two points per lane"}
SAMPLE = 4096          # points of each timed output checked against the CPU oracle


def _cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _host(t, k=None):
    import torch
    torch.cuda.synchronize()
    return (t if k is None else t[:k]).cpu().numpy()


def _timed(torch, fn, stream, reps=3):
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    fn()
    torch.cuda.synchronize()
    best = float("inf")
    for _ in range(reps):
        ev[0].record(stream)
        fn()
        ev[1].record(stream)
        torch.cuda.synchronize()
        best = min(best, ev[0].elapsed_time(ev[1]))
    return best


def _profile_json(d, n, P, variant):
    """rocprofv3 figures of the evaluate kernel recorded in profiles/ (HBM bytes from
    FETCH_SIZE / WRITE_SIZE scaled by a plain copy; unit counters), if they are for this
    workload and walk variant."""
    path = os.path.join(ROOT, "profiles", f"eval_d{d}n{n}_traffic.json")
    try:
        tj = json.load(open(path))
    except (OSError, ValueError):
        return None
    if tj.get("num_points") != P or tj.get("variant") != variant:
        return None
    return tj


# --------------------------------------------------------------------- forms --

def precision_form(args, stream):
    """d=20, n=6 (N=1,515,009): a seeded sum of 8 anisotropic Gaussians (centres U[0.2,
    0.8]^20, widths log-uniform [0.05, 0.5]) hierarchized in FP64 on the GPU, evaluated at
    2^24 uniform points in FP32 with Kahan accumulation and in FP64.  Returns the
    timings and a sample for the CPU oracle check."""
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
    t64 = _timed(torch, lambda: g.evaluate(f, pts, o64, stream), stream, reps=2)
    a32 = f.float()
    sample = {"d": d, "n": n, "alpha": _host(f), "points": _host(pts, args.precision_sample),
              "f64": _host(o64, args.precision_sample)}
    p32 = pts.float()
    del pts
    o32 = torch.empty(P, dtype=torch.float32, device="cuda")
    t32 = _timed(torch, lambda: g.evaluate(a32, p32, o32, stream), stream, reps=2)
    sample["f32c"] = _host(o32, args.precision_sample)
    v64, v32 = g.get_tuning()["eval_variant_f64"], g.get_tuning()["eval_variant_f32"]
    res = {"workload": f"d=20 n=6 (N={g.N}, S={g.S}), sum of 8 Gaussians, 2^{args.log2_points} "
                       "points; FP64 walk and FP32 Kahan walk",
           "f64_ms": t64, "f32c_ms": t32, "f64_variant": v64, "f32c_variant": v32,
           "f64_pts_per_s": P / (t64 * 1e-3), "f32c_pts_per_s": P / (t32 * 1e-3),
           "max_abs_u": o64.abs().max().item()}
    del o64, o32, a32, p32, f
    torch.cuda.empty_cache()
    return res, sample


def other_forms(g, alpha, pts, P, args, stream):
    """FP32-Kahan evaluation of the headline workload, the precision form, the batched form
    (256 grids, one hipGraph launch vs per-grid launches) and the served form (continuous
    batching vs plain FP64 per-request launches at two loads, p50/p99).  Returns the
    results and the samples the CPU leg checks."""
    import torch
    from adaptive_sparse_grid_bench_amd import batched, serve
    res, samples = {}, {}
    a32, p32 = alpha.float(), pts.float()
    o32 = torch.empty(P, dtype=torch.float32, device="cuda")
    ms = _timed(torch, lambda: g.evaluate(a32, p32, o32, stream), stream, reps=2)
    res["f32c"] = {"evaluate_ms": ms, "pts_per_s": P / (ms * 1e-3),
                   "variant": g.get_tuning()["eval_variant_f32"]}
    samples["f32c"] = _host(o32, SAMPLE)
    del a32, p32, o32
    res["precision"], samples["precision"] = precision_form(args, stream)
    parts = batched.build(256, tune=True)
    res["batched"] = batched.measure(*parts)
    specs, grids, b, nodal, coords = parts
    # every grid's first 16 points after a graph run in each precision
    bs = {"grids": [(gr.d, gr.n) for gr in grids], "k": 16}
    for prec, dt in (("f64", torch.float64), ("f32c", torch.float32)):
        a = nodal.to(dt).clone()
        x = coords.to(dt)
        out = torch.empty(int(b.point_offsets[-1]), dtype=dt, device="cuda")
        b.run_graph(a, x, out, hierarchize=True)
        a_h, x_h, o_h = _host(a), _host(x), _host(out)
        bs[prec] = [(a_h[int(b.value_offsets[i]):int(b.value_offsets[i + 1])],
                     x_h[int(b.coord_offsets[i]):int(b.coord_offsets[i]) + 16 * gr.d]
                     .reshape(16, gr.d),
                     o_h[int(b.point_offsets[i]):int(b.point_offsets[i]) + 16])
                    for i, gr in enumerate(grids)]
        del a, x, out
    samples["batched"] = bs
    del parts, specs, grids, b, nodal, coords
    grids = serve.make_grids(64)
    sv = serve.experiment(grids, args.served_requests, (0.5, 0.8), keep_outputs=48)
    kept = sv.pop("_kept")
    samples["served"] = {"requests": kept,
                         "grids": [(gr["grid"].subspaces()[0] if gr["grid"].is_explicit else None,
                                    gr["grid"].d, gr["grid"].n, _host(gr["a64"]), _host(gr["a32"]))
                                   for gr in grids]}
    res["served"] = sv
    del grids
    torch.cuda.empty_cache()
    return res, samples


# --------------------------------------------------------------------- CPU leg --

def _errs(out, ref, absref=None, u=None):
    import numpy as np
    out = np.asarray(out, np.float64)
    e = np.abs(out - ref)
    r = {"max_abs": float(e.max()), "rms": float(np.sqrt(np.mean(e * e))), "points": int(e.size)}
    if absref is not None:
        r["max_rel_to_sum_abs_terms"] = float(np.max(e / np.maximum(absref, 1e-300)))
    if u is not None:
        r["max_rel_to_abs_u"] = float(np.max(e / np.maximum(np.abs(u), 1e-300)))
    return r


def cpu_leg(d, n, P_sample, seed, head, samples):
    """The reference's OpenMP/SSE CPU path timed on this host, and the CPU oracle's check
    of the timed GPU outputs (the oracle is used here only as the checker)."""
    import numpy as np
    from adaptive_sparse_grid_bench_amd import workloads as W
    from oracle import oracle as O
    cores = len(os.sched_getaffinity(0))
    baseline, reason = None, None
    try:
        from oracle import refbind
        if not refbind.available():
            raise ImportError("oracle/_ref/libasgref.so is not built")
        g = refbind.RefGrid(d, n)
        threads = g.omp_threads()
        pts = W.uniform_points(P_sample, d, seed=seed)
        x = pts.astype(np.float32)
        vals = g.get_values()
        hier = []
        for _ in range(5):
            g.set_values(vals)
            t0 = time.perf_counter()
            g.hierarchize("omp_hierarchize3")
            hier.append(time.perf_counter() - t0)
        g.evaluate(x[:256], "omp_evaluate5")                       # warm-up
        t0 = time.perf_counter()
        g.evaluate(x, "omp_evaluate5")
        dt = time.perf_counter() - t0
        baseline = {"value": P_sample / dt, "unit": "points/s", "cores": threads,
                    "kind": "reference",
                    "sample": f"reference omp_evaluate5 (FP32, as the reference computes) on "
                              f"{P_sample} of the workload's query points at d={d} n={n} "
                              f"({dt:.1f} s); omp_hierarchize3 median of 5 on the full grid",
                    "omp_threads": threads, "host_cores_available": cores,
                    "hierarchize_gp_per_s": g.N / sorted(hier)[2],
                    "evaluate_gflops_ref_convention": P_sample * g.S * (11 * d + 2) / dt / 1e9,
                    "host_cpu": _cpu_model()}
    except (ImportError, OSError) as exc:
        reason = f"reference CPU library unavailable: {exc}"
    checks = {}
    sh = O.tables(d, n)
    alpha_ref = O.hierarchize(sh, head["nodal"])
    checks["hierarchize_bit_exact"] = bool(np.array_equal(alpha_ref, head["alpha"]))
    ref = O.evaluate(sh, alpha_ref, head["points"])
    absref = O.evaluate(sh, np.abs(alpha_ref), head["points"])
    checks["f64"] = _errs(head["f64"], ref, absref, ref)
    if "f32c" in samples:
        checks["f32c"] = _errs(samples["f32c"], ref, absref, ref)
    if "precision" in samples:
        s = samples["precision"]
        sp = O.tables(s["d"], s["n"])
        r = O.evaluate(sp, s["alpha"], s["points"])
        ar = O.evaluate(sp, np.abs(s["alpha"]), s["points"])
        checks["precision"] = {"f64": _errs(s["f64"], r, ar, r), "f32c": _errs(s["f32c"], r, ar, r)}
    if "batched" in samples:
        s = samples["batched"]
        out = {}
        for prec in ("f64", "f32c"):
            es, rs, worst = [], [], 0.0
            for (dd, nn), (a, x, o) in zip(s["grids"], s[prec]):
                sb = O.tables(dd, nn)
                r = O.evaluate(sb, a.astype(np.float64), x.astype(np.float64))
                ar = O.evaluate(sb, np.abs(a.astype(np.float64)), x.astype(np.float64))
                e = np.abs(o.astype(np.float64) - r)
                es.append(e)
                worst = max(worst, float(np.max(e / np.maximum(ar, 1e-300))))
            e = np.concatenate(es)
            out[prec] = {"max_abs": float(e.max()), "rms": float(np.sqrt(np.mean(e * e))),
                         "max_rel_to_sum_abs_terms": worst, "points": int(e.size),
                         "note": "FP32 outputs against the FP64 oracle on the FP32 surpluses"
                         if prec == "f32c" else "FP64 outputs against the FP64 oracle"}
        checks["batched"] = out
    if "served" in samples:
        s = samples["served"]
        shapes = {}
        e64, e32 = [], []
        for (gi, p64, _), o64, o32 in s["requests"]:
            k = min(16, len(p64))
            if k == 0:
                continue
            lv, dd, nn, a64, _ = s["grids"][gi]
            if gi not in shapes:
                shapes[gi] = O.xshape(lv) if lv is not None else O.tables(dd, nn)
            ev = O.x_evaluate if lv is not None else O.evaluate
            r = ev(shapes[gi], a64, p64[:k])
            e64.append(np.abs(o64[:k] - r))
            e32.append(np.abs(o32[:k].astype(np.float64) - r))
        for name, es in (("f64_server", e64), ("f32c_server", e32)):
            e = np.concatenate(es)
            checks[f"served_{name}"] = {"max_abs": float(e.max()),
                                        "rms": float(np.sqrt(np.mean(e * e))),
                                        "points": int(e.size), "requests": len(es)}
    checks["oracle"] = "CPU oracle (oracle/sg_oracle.c, FP64 restatement of evaluate0)"
    return baseline, reason, checks


# ------------------------------------------------------------------------ main --

def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--d", type=int, default=10)
    ap.add_argument("--n", type=int, default=8)
    ap.add_argument("--log2-points", type=int, default=24)
    ap.add_argument("--seed", type=int, default=42)
    ap.add_argument("--cpu-sample", type=int, default=1 << 19)
    ap.add_argument("--precision-sample", type=int, default=1024)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-forms", action="store_true",
                    help="skip the FP32, batched and served forms (headline FP64 line only)")
    ap.add_argument("--served-requests", type=int, default=5000)
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run one plain device copy of MIB MiB (the reference kernel "
                         "for scaling rocprofv3 FETCH_SIZE / WRITE_SIZE) and exit")
    ap.add_argument("--probe-only", action="store_true",
                    help="only run the headline evaluate and its staged-pattern probe "
                         "(for profiling both kernels)")
    ap.add_argument("--precision-only", action="store_true",
                    help="only run the precision form (d=20, n=6, FP64 and FP32-Kahan "
                         "evaluation of 2^log2-points points; for profiling)")
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

    if args.precision_only:
        import torch
        torch.cuda.set_device(0)
        res, _ = precision_form(args, torch.cuda.current_stream())
        print(json.dumps({"precision_only": True, **res}), flush=True)
        return 0

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
    variant = g.get_tuning()["eval_variant_f64"]
    probe_out = torch.empty(P, dtype=torch.float64, device=dev)
    torch.cuda.synchronize()

    if args.probe_only:
        alpha.copy_(nodal)
        g.hierarchize(alpha, stream)
        g.evaluate(alpha, pts, out, stream)
        sgrid.stage_probe(g, alpha, probe_out, seed=7, stream=stream)
        torch.cuda.synchronize()
        print(json.dumps({"probe_only": True, "variant": variant}), flush=True)
        return 0

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
    head = {"nodal": _host(nodal), "alpha": _host(alpha), "points": _host(pts, SAMPLE),
            "f64": _host(out, SAMPLE)}

    # the roofline: the staged-pattern probe of the walk this grid runs (its floor)
    probe_ms = _timed(torch, lambda: sgrid.stage_probe(g, alpha, probe_out, seed=7,
                                                       stream=stream), stream, reps=3)
    del probe_out
    gather_bytes = P * g.S * 8                   # one 8-byte read per subspace per point
    prof = _profile_json(d, n, P, variant)
    traffic = prof.get("hbm_bytes_per_launch") if prof else None
    roofline = {
        "kernel": f"k_eval<{d},...,double> walk variant {variant} ({VARIANT_NAMES[variant]})",
        "bound": "lds-staged",
        "achieved": gather_bytes / (eval_avg * 1e-3) / 1e9,
        "peak": gather_bytes / (probe_ms * 1e-3) / 1e9,
        "unit": "GB/s",
        "frac": probe_ms / eval_avg,
        "traffic": traffic,
        "ceiling": "k_stage_probe: the walk's stage plan, copies and barriers, one read per "
                   "subspace per point (LDS, or global memory for unstaged blocks) and one "
                   "FMA; no hats, no walk",
        "probe_ms": probe_ms,
        "evaluate_ms": eval_avg,
        "bytes_per_point_read": g.S * 8,
        "algorithmic_bytes_per_point": g.S * 8 + d * 8 + 8,
        "hbm": {"achieved": traffic / (eval_avg * 1e-3) / 1e9 if traffic else None,
                "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": traffic / (eval_avg * 1e-3) / 1e9 / HBM_PEAK_GBS if traffic else None,
                "note": "rocprofv3 FETCH_SIZE and WRITE_SIZE of this kernel, scaled by a plain "
                        "copy's reading (profiles/)"},
        "units": prof.get("units") if prof else None,
    }

    ev_flops = P * g.S * (11 * d + 2)
    hi_flops = 3 * d * g.N
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
        # a full hierarchization reads and writes every value once per dimension sweep;
        # the pair kernel sweeps two dimensions per pass, so d/2 passes move 2 N 8 bytes
        # each: these are the HBM bytes (the grid is far larger than the LDS)
        "hierarchize_roofline": {
            "kernels": "k_hierarchize_pair (two dimensions per launch), k_hierarchize_dim",
            "bound": "hbm",
            "bytes": d * g.N * 8,
            "achieved": d * g.N * 8 / (hier_med * 1e-3) / 1e9,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": d * g.N * 8 / (hier_med * 1e-3) / 1e9 / HBM_PEAK_GBS,
        },
        "evaluate_ms": eval_avg,
        "evaluate_pts_per_s_kernel": P / (eval_avg * 1e-3),
        "gflops_reference_convention": {
            "hierarchize": hi_flops / (hier_med * 1e-3) / 1e9,
            "evaluate": ev_flops / (eval_avg * 1e-3) / 1e9,
            "note": "HI_FLOPS = 3 d N, EV_FLOPS = P S (11 d + 2) (adaptive_sparse_grid_bench.c:"
                    "48-52); the walk computes far fewer flops than the convention counts",
        },
        "roofline": roofline,
        "cpu_baseline": None,
    }
    samples = {}
    if not args.no_forms:
        forms, samples = other_forms(g, alpha, pts, P, args, stream)
        rec.update(forms)
    if not args.no_cpu_baseline:
        baseline, reason, checks = cpu_leg(d, n, args.cpu_sample, args.seed, head, samples)
        rec["cpu_baseline"] = baseline
        if reason:
            rec["cpu_baseline_unavailable"] = reason
        elif baseline:
            rec["speedup_vs_cpu_baseline"] = value / baseline["value"]
        rec["checks_vs_cpu_oracle"] = checks
    print(json.dumps(rec), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
