"""The batched workload: 256 independent regular grids of uneven dimension and level
(d ~ U{2..12}, n ~ U{3..8}, values prod_i sin(pi x_i) * (1 + 0.1 c_g)), hierarchized and
evaluated at floor(2^U[12,18]) uniform points each, as one hipGraph launch, against the
same work as per-grid launches.  Launch tuning per grid shape comes from the autotuner
(walk variant, workgroup size and hierarchization tile size).

Usage: python -m adaptive_sparse_grid_bench_amd.batched [--grids 256] [--no-tune]
"""
from __future__ import annotations

import argparse
import json
import time

import numpy as np

from . import autotune, sgrid
from . import workloads as W


def build(count: int = 256, seed: int = 5, tune: bool = True, cache=autotune.DEFAULT_CACHE):
    import torch
    specs = W.batched_specs(count, seed)
    grids, nodal = [], []
    tuned = {}
    for s in specs:
        g = sgrid.Grid(s["d"], s["n"])
        key = (s["d"], s["n"])
        if tune:
            if key not in tuned:
                if not autotune.apply_cached(g, cache):
                    autotune.tune(g, points=1 << 15, reps=2, cache_path=cache)
                tuned[key] = g.get_tuning()
            g.set_tuning(**tuned[key])
        x = torch.empty((g.N, g.d), dtype=torch.float64, device="cuda")
        g.grid_coords(x)
        nodal.append(torch.sin(np.pi * x).prod(1) * (1.0 + 0.1 * s["c"]))
        del x
        grids.append(g)
    b = sgrid.Batch(grids)
    counts = [s["P"] for s in specs]
    b.set_points(counts)
    nodal_all = torch.cat(nodal)
    coords = torch.empty(int(b.coord_offsets[-1]), dtype=torch.float64, device="cuda")
    sgrid.uniform_points(coords, seed=seed + 1)
    torch.cuda.synchronize()
    return specs, grids, b, nodal_all, coords


def _timed(torch, fn, reps=3):
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    fn()
    torch.cuda.synchronize()
    best = float("inf")
    for _ in range(reps):
        ev[0].record()
        fn()
        ev[1].record()
        torch.cuda.synchronize()
        best = min(best, ev[0].elapsed_time(ev[1]))
    return best


def measure(specs, grids, b, nodal_all, coords):
    import torch
    N = int(b.total_values)
    P = int(b.point_offsets[-1])
    res = {"grids": len(grids), "grid_points": N, "points": P}
    for prec, dt in (("f64", torch.float64), ("f32c", torch.float32)):
        nod = nodal_all.to(dt)
        x = coords.to(dt)
        a = torch.empty_like(nod)
        out = torch.empty(P, dtype=dt, device="cuda")

        def graph():
            a.copy_(nod)
            b.run_graph(a, x, out, hierarchize=True)

        def eager():
            a.copy_(nod)
            b.hierarchize(a)
            b.evaluate(a, x, out)

        def per_grid():
            a.copy_(nod)
            for i, g in enumerate(grids):
                lo, hi = int(b.value_offsets[i]), int(b.value_offsets[i + 1])
                g.hierarchize(a[lo:hi])
            for i, g in enumerate(grids):
                lo, hi = int(b.value_offsets[i]), int(b.value_offsets[i + 1])
                c0, c1 = int(b.coord_offsets[i]), int(b.coord_offsets[i + 1])
                p0, p1 = int(b.point_offsets[i]), int(b.point_offsets[i + 1])
                g.evaluate(a[lo:hi], x[c0:c1].view(p1 - p0, g.d), out[p0:p1])

        t_graph = _timed(torch, graph)
        ref = out.clone()
        ref_a = a.clone()
        t_eager = _timed(torch, eager)
        same_eager = torch.equal(out, ref) and torch.equal(a, ref_a)
        t_per = _timed(torch, per_grid)
        same_per = torch.equal(out, ref) and torch.equal(a, ref_a)

        def hier_only():
            a.copy_(nod)
            b.hierarchize(a)
        t_h = _timed(torch, hier_only)
        res[prec] = {"graph_ms": t_graph, "eager_ms": t_eager, "per_grid_ms": t_per,
                     "hierarchize_ms": t_h, "hierarchize_gp_per_s": N / (t_h * 1e-3),
                     "points_per_s_graph": P / (t_graph * 1e-3),
                     "speedup_graph_vs_per_grid": t_per / t_graph,
                     "graph_equals_eager": bool(same_eager),
                     "graph_equals_per_grid": bool(same_per)}
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--grids", type=int, default=256)
    ap.add_argument("--no-tune", action="store_true")
    args = ap.parse_args()
    t = time.time()
    parts = build(args.grids, tune=not args.no_tune)
    setup = time.time() - t
    res = measure(*parts)
    res["setup_s"] = setup
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
