"""The served workload: resident grids, Poisson-arrival evaluation requests, latency
percentiles and throughput of the continuous-batching server against plain per-request
FP64 launches.

Workload: ``count`` resident grids, a mix of regular ones (d ~ U{2..16}, n ~ U{3..8},
f = exp(-sum (x-1/2)^2)) and refined-subspace ones (d = 8, base level 3, up to 4 extra
levels per dimension, refined towards a seeded Gaussian peak of width 0.15, f that peak).
Requests pick a grid uniformly and carry floor(2^U[0,16]) uniform points (1 to 65,536,
log-uniform).  Arrivals are Poisson at a rate set as a fraction of the capacity measured
by a closed-loop run.  Latency = completion - arrival (queueing included); throughput =
points / (last completion - first arrival).

Usage: python -m adaptive_sparse_grid_bench_amd.serve [--grids 64] [--requests 5000]
       [--grid-dir DIR]   (resident grids from compact-grid files, see sgrid.save/load)
"""
from __future__ import annotations

import argparse
import json
import os
import threading
import time

import numpy as np

from . import autotune, sgrid
from . import workloads as W


def load_grids(grid_dir: str):
    """Resident grids read from compact-grid files (grid_NNN.sgrid, FP64 surpluses)."""
    import torch
    grids = []
    for name in sorted(os.listdir(grid_dir)):
        if not name.endswith(".sgrid"):
            continue
        g, a = sgrid.load(os.path.join(grid_dir, name))
        a = torch.from_numpy(a).cuda()
        kind = "regular" if name.endswith("_r.sgrid") else "adaptive"
        grids.append({"grid": g, "a64": a, "a32": a.float(), "kind": kind})
    torch.cuda.synchronize()
    return grids


def save_grids(grids, grid_dir: str) -> None:
    os.makedirs(grid_dir, exist_ok=True)
    for i, g in enumerate(grids):
        tag = "r" if g["kind"] == "regular" else "a"
        sgrid.save(os.path.join(grid_dir, f"grid_{i:03d}_{tag}.sgrid"), g["grid"], g["a64"])


def make_grids(count: int = 64, seed: int = 21, adaptive_every: int = 4,
               max_points: int = 1 << 25, grid_dir: str | None = None,
               keep_nodal: bool = False):
    """Resident grids with FP64 surpluses and FP32 copies (torch CUDA tensors).  With
    grid_dir, grids are read from its compact-grid files when present and written there
    after being built otherwise.  keep_nodal also keeps each grid's nodal values (a host
    array, "nodal") for checking the hierarchization."""
    import torch
    if grid_dir and os.path.isdir(grid_dir) and any(f.endswith(".sgrid") for f in os.listdir(grid_dir)):
        return load_grids(grid_dir)
    rng = np.random.default_rng(seed)
    grids = []
    k = 0
    while len(grids) < count:
        k += 1
        if len(grids) % adaptive_every == adaptive_every - 1:
            lv, c = W.adaptive_level_set(d=8, base_n=3, extra=4, seed=seed * 1000 + k)
            g = sgrid.Grid.explicit(8, lv)
            kind = "adaptive"
            centre = torch.tensor(c, dtype=torch.float64, device="cuda")
        else:
            d, n = int(rng.integers(2, 17)), int(rng.integers(3, 9))
            g = sgrid.Grid(d, n)
            if g.N > max_points:
                continue
            autotune.apply_cached(g)             # the server takes each grid's variant
            kind = "regular"
            centre = None
        x = torch.empty((g.N, g.d), dtype=torch.float64, device="cuda")
        g.grid_coords(x)
        if centre is None:
            a = torch.exp(-((x - 0.5) ** 2).sum(1))
        else:
            a = torch.exp(-((x - centre) ** 2).sum(1) / (2 * 0.15 ** 2))
        del x
        nodal = a.cpu().numpy() if keep_nodal else None
        g.hierarchize(a)
        grids.append({"grid": g, "a64": a, "a32": a.float(), "kind": kind})
        if keep_nodal:
            grids[-1]["nodal"] = nodal
    torch.cuda.synchronize()
    if grid_dir:
        save_grids(grids, grid_dir)
    return grids


def make_requests(grids, count: int, seed: int = 5):
    rng = np.random.default_rng(seed)
    reqs = []
    for r in range(count):
        gi = int(rng.integers(len(grids)))
        P = int(np.floor(2.0 ** rng.uniform(0, 16)))
        pts = W.uniform_points(P, grids[gi]["grid"].d, seed=100000 + r)
        reqs.append((gi, pts, pts.astype(np.float32)))
    return reqs


def _percentiles(lat):
    a = np.sort(np.asarray(lat))
    return float(np.percentile(a, 50)), float(np.percentile(a, 99))


def run_server(grids, reqs, precision: str, arrivals=None, slots: int = 4,
               max_batch_points: int = 1 << 19):
    """Serve ``reqs`` (all at once if ``arrivals`` is None, else at those offsets in s).

    Four slots of at most 2^19 points per launch: with 3 slots of 2^20 points the
    closed-loop capacity is higher (43-46 against 34 M points/s FP32-Kahan) but Poisson
    arrivals at 0.8 of it ran either 31 or 23 M points/s (p50 72 or 228 ms) from one run
    to the next, as single launches of up to 10^6 points held the slots; with 4 x 2^19
    two runs agree within 5 % (p50 49-52 ms)."""
    srv = sgrid.Server(precision, slots=slots, max_batch_points=max_batch_points,
                       max_requests=2048)
    key = "a64" if precision == "f64" else "a32"
    ids = [srv.add_grid(g["grid"], g[key]) for g in grids]
    dt = np.float64 if precision == "f64" else np.float32
    outs = [np.empty(len(p64), dt) for _, p64, _ in reqs]
    rids = [None] * len(reqs)
    t_sub = [0.0] * len(reqs)
    ready = threading.Event()
    t0 = [0.0]

    def client():
        ready.wait()
        for i, (gi, p64, p32) in enumerate(reqs):
            if arrivals is not None:
                delay = t0[0] + arrivals[i] - time.perf_counter()
                if delay > 0:
                    time.sleep(delay)
            t_sub[i] = time.perf_counter()
            rids[i] = srv.submit(ids[gi], p64 if precision == "f64" else p32, outs[i])

    th = threading.Thread(target=client)
    th.start()
    t0[0] = time.perf_counter()
    ready.set()
    th.join()
    lat = [srv.wait(r) for r in rids]
    t_end = max(t_sub[i] + lat[i] for i in range(len(reqs)))
    stats = srv.stats()
    srv.close()
    pts = sum(len(o) for o in outs)
    p50, p99 = _percentiles(lat)
    return {"points": pts, "seconds": t_end - t0[0], "points_per_s": pts / (t_end - t0[0]),
            "p50_s": p50, "p99_s": p99, "launches": stats["launches"],
            "requests": len(reqs)}, outs


def run_plain_f64(grids, reqs, arrivals=None):
    """The plain FP64 path: one request at a time, copy in, per-grid evaluate, copy out."""
    import torch
    stream = torch.cuda.current_stream()
    lat, outs = [], []
    t0 = time.perf_counter()
    t_end = t0
    for i, (gi, p64, _) in enumerate(reqs):
        arrive = t0 + (arrivals[i] if arrivals is not None else 0.0)
        now = time.perf_counter()
        if now < arrive:
            time.sleep(arrive - now)
        x = torch.from_numpy(p64).cuda(non_blocking=False)
        o = torch.empty(len(p64), dtype=torch.float64, device="cuda")
        grids[gi]["grid"].evaluate(grids[gi]["a64"], x, o, stream)
        out = o.cpu().numpy()
        t_end = time.perf_counter()
        lat.append(t_end - arrive)
        outs.append(out)
    pts = sum(len(o) for o in outs)
    p50, p99 = _percentiles(lat)
    return {"points": pts, "seconds": t_end - t0, "points_per_s": pts / (t_end - t0),
            "p50_s": p50, "p99_s": p99, "requests": len(reqs)}, outs


def experiment(grids, n_requests: int = 5000, loads=(0.5, 0.8), seed: int = 5,
               keep_outputs: int = 0):
    """Closed-loop capacity of the FP32 Kahan server, then for each load Poisson arrivals
    at load x capacity for the FP32 Kahan server, the FP64 server and the plain FP64
    launches (the same requests and arrival times for all three).  With keep_outputs,
    the first keep_outputs requests and their outputs of the last load are returned too
    (res["_kept"]) for checking."""
    reqs = make_requests(grids, n_requests, seed)
    run_server(grids, reqs[: min(50, len(reqs))], "f32c")          # warm-up
    cap, _ = run_server(grids, reqs, "f32c")
    cap_f64, _ = run_server(grids, reqs, "f64")
    res = {"requests": len(reqs), "points": int(sum(len(p) for _, p, _ in reqs)),
           "capacity_closed_loop": {"f32c_server": cap, "f64_server": cap_f64},
           "loads": [],
           "grids": {"count": len(grids),
                     "adaptive": sum(g["kind"] == "adaptive" for g in grids),
                     "grid_points": int(sum(g["grid"].N for g in grids))}}
    for li, load in enumerate(loads):
        rate = load * cap["requests"] / cap["seconds"]               # requests / s
        rng = np.random.default_rng(seed + 1 + li)
        arrivals = np.cumsum(rng.exponential(1.0 / rate, len(reqs)))
        f32, o32 = run_server(grids, reqs, "f32c", arrivals)
        f64, o64 = run_server(grids, reqs, "f64", arrivals)
        plain, op = run_plain_f64(grids, reqs, arrivals)
        err = [np.abs(a.astype(np.float64) - b) for a, b in zip(o32, o64) if len(a)]
        allerr = np.concatenate(err) if err else np.zeros(1)
        same = all(np.array_equal(a, b) for a, b in zip(o64, op))
        res["loads"].append({
            "load": load, "arrival_rate_per_s": rate,
            "f32c_server": f32, "f64_server": f64, "plain_f64": plain,
            "f32c_vs_f64_server": {"max_abs": float(allerr.max()),
                                   "rms": float(np.sqrt(np.mean(allerr ** 2)))},
            "f64_server_equals_plain": bool(same),
            "speedup_f32c_server_vs_plain_f64": plain["seconds"] / f32["seconds"],
            "speedup_f64_server_vs_plain_f64": plain["seconds"] / f64["seconds"]})
        if keep_outputs and li == len(loads) - 1:
            res["_kept"] = [(reqs[i], o64[i], o32[i]) for i in range(min(keep_outputs, len(reqs)))]
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--grids", type=int, default=64)
    ap.add_argument("--requests", type=int, default=5000)
    ap.add_argument("--loads", default="0.5,0.8")
    ap.add_argument("--grid-dir", default=None,
                    help="read the resident grids from compact-grid files here (written on "
                         "the first run)")
    args = ap.parse_args()
    t = time.time()
    grids = make_grids(args.grids, grid_dir=args.grid_dir)
    res = experiment(grids, args.requests, tuple(float(x) for x in args.loads.split(",")))
    res["setup_s"] = time.time() - t
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
