"""Autotuner of the launch parameters of one grid shape.

For a grid it times, on the GPU and with HIP events on the launching stream:
  * the evaluate kernel over the compiled walk variants (tabled dimensions, gathers per
    batch, waves per SIMD, the staged and block-staged walks, one or two points per
    lane), the workgroup size and spatial binning on or off, in FP64 and FP32 (Kahan);
    a variant the grid does not cover runs another one (``Grid.walk_variant``) and is
    recorded as such instead of being timed again;
  * the hierarchization over the pole-family tile size (values per workgroup).
The fastest choice is applied with ``Grid.set_tuning`` and stored in a JSON cache keyed
by the grid shape (d, n, restriction vector or a digest of the explicit level set), so a
later process applies it without timing again.  Timings use synthetic data of the
grid's shape; results do not depend on the data (the walk and the sweep are
data-independent).
"""
from __future__ import annotations

import hashlib
import json
import os
import time

import numpy as np

from . import sgrid

DEFAULT_CACHE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "autotune_cache.json")

EVAL_BLOCKS = (64, 128, 256)
HIER_TILES = (256, 512, 1024, 2048, 4096)


def shape_key(g: "sgrid.Grid") -> str:
    lv, _ = g.subspaces()
    digest = hashlib.sha1(np.ascontiguousarray(lv).tobytes()).hexdigest()[:16]
    return f"d{g.d}_n{g.n}_S{g.S}_{digest}"


def _load_cache(path: str) -> dict:
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _store_cache(path: str, cache: dict) -> None:
    tmp = f"{path}.{os.getpid()}.tmp"
    with open(tmp, "w") as f:
        json.dump(cache, f, indent=1, sort_keys=True)
    os.replace(tmp, path)


def apply_cached(g: "sgrid.Grid", cache_path: str = DEFAULT_CACHE) -> bool:
    """Apply a stored choice for this grid's shape; False if there is none (or it was
    timed on another version of the kernels)."""
    entry = _load_cache(cache_path).get(shape_key(g))
    if not entry or entry.get("library_version") != sgrid.lib.sg_version():
        return False
    g.set_tuning(**entry["tuning"])
    return True


def _time(torch, fn, reps: int) -> float:
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    fn()                                                   # warm-up
    best = float("inf")
    for _ in range(reps):
        torch.cuda.synchronize()
        ev[0].record()
        fn()
        ev[1].record()
        torch.cuda.synchronize()
        best = min(best, ev[0].elapsed_time(ev[1]))
    return best


def tune(g: "sgrid.Grid", points: int = 1 << 17, reps: int = 3, precisions=("f64", "f32"),
         cache_path: str | None = DEFAULT_CACHE, verbose: bool = False) -> dict:
    """Time the candidates for grid ``g``, apply the fastest and return the record."""
    import torch

    t_start = time.time()
    dev = torch.device("cuda")
    pts = torch.empty((points, g.d), dtype=torch.float64, device=dev)
    sgrid.uniform_points(pts, seed=1234)
    alpha = torch.empty(g.N, dtype=torch.float64, device=dev)
    sgrid.uniform_points(alpha, seed=4321)
    data = {"f64": (alpha, pts, torch.empty(points, dtype=torch.float64, device=dev)),
            "f32": (alpha.float(), pts.float(), torch.empty(points, dtype=torch.float32,
                                                            device=dev))}
    base = g.get_tuning()
    timings = {}
    runs_as = {}
    best = dict(base)
    for prec in precisions:
        a, x, o = data[prec]
        key = "eval_variant_f64" if prec == "f64" else "eval_variant_f32"
        cand = []
        for variant in range(sgrid.EVAL_VARIANTS):
            # a variant this grid does not cover runs another one: time each walk once
            walk = g.walk_variant(variant, prec)
            if walk != variant:
                runs_as[f"{prec}/v{variant}"] = walk
                continue
            staged = variant in sgrid.STAGED_VARIANTS
            # the staged walks choose their workgroup size themselves and never bin
            for block in ((256,) if staged else EVAL_BLOCKS):
                for binning in ((-1,) if staged else (-1, 4096)):
                    g.set_tuning(**{**base, key: variant, "eval_block": block,
                                    "bin_min_points": binning})
                    ms = _time(torch, lambda: g.evaluate(a, x, o), reps)
                    cand.append((ms, variant, block, binning))
                    timings[f"{prec}/v{variant}/b{block}/bin{binning}"] = ms
        cand.sort()
        ms, variant, block, binning = cand[0]
        best[key] = variant
        if prec == "f64" or "f64" not in precisions:
            best["eval_block"] = block
            best["bin_min_points"] = binning
        if verbose:
            print(f"[autotune] {shape_key(g)} {prec}: variant {variant} block {block} "
                  f"bin {binning} {ms:.3f} ms for {points} points")
    hcand = []
    v = torch.empty(g.N, dtype=torch.float64, device=dev)
    for tile in HIER_TILES:
        g.set_tuning(**{**best, "hier_tile_values": tile})
        ms = _time(torch, lambda: (v.copy_(alpha), g.hierarchize(v)), reps)
        hcand.append((ms, tile))
        timings[f"hier/tile{tile}"] = ms
    hcand.sort()
    best["hier_tile_values"] = hcand[0][1]
    g.set_tuning(**best)
    record = {"tuning": best, "timings_ms": timings, "runs_as": runs_as, "points": points,
              "tuned_in_s": round(time.time() - t_start, 2),
              "library_version": sgrid.lib.sg_version()}
    if cache_path:
        cache = _load_cache(cache_path)
        cache[shape_key(g)] = record
        _store_cache(cache_path, cache)
    return record
