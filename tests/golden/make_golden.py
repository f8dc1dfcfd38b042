"""Write the golden fixtures under tests/golden/ from the reference C benchmark itself.

Run in a container that holds the reference sources:

    make -C oracle ref && python tests/golden/make_golden.py

It loads oracle/_ref/libasgref.so (adaptive_sparse_grid_bench.c compiled from source,
main() renamed) and records, in FP32 exactly as the reference computes:

  * idx2gp tables (levels, indices per offset) for small shapes, sampled offsets at
    d=10 n=8 and d=20 n=6, and shapes with a restriction vector (asg.c:82-83, 210-243);
  * idx2l for every (rank, level sum), l2idx / gp2idx round trips (asg.c:187-279);
  * grid values from the reference's own init() (asg.c:143-152) and their
    hierarchize0 surpluses (asg.c:557-571), plus cross-checks against the other
    hierarchize variants where they are valid;
  * evaluate0 outputs (asg.c:1237-1274) at fixed FP32 points;
  * for the d=10 n=8 shape, where whole arrays would be too large to commit, an FP64
    checksum of the inputs and surpluses, sampled surpluses, and evaluate0 outputs.

Fixtures are .npz files without pickled objects.  The reference itself never leaves this
container in any form; only these input/output vectors are committed.
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.abspath(os.path.join(HERE, "..", "..")))

from oracle import refbind  # noqa: E402
from adaptive_sparse_grid_bench_amd import workloads as W  # noqa: E402


def full_bijection(g: refbind.RefGrid):
    lv = np.zeros((g.N, g.d), np.int32)
    iv = np.zeros((g.N, g.d), np.int32)
    for o in range(g.N):
        lv[o], iv[o] = g.idx2gp(o)
    return lv, iv


def write_small(d, n, limits, tag, funcs=("reference", "parabola", "gauss"), P=512):
    g = refbind.RefGrid(d, n, limits)
    lv, iv = full_bijection(g)
    ranks = [(s, r, *g.idx2l(r, s)) for s in range(n) for r in range(int(g.dp[d - 1][s]))]
    rank_tab = np.array(ranks, np.int32)
    rt = np.array([g.gp2idx(lv[o], iv[o]) for o in range(g.N)], np.int32)
    l2 = np.array([g.l2idx(row[2:]) for row in rank_tab], np.int32)
    coords = ((2.0 * iv + 1.0) / np.exp2(lv + 1.0)).astype(np.float32)
    init_values = g.get_values()
    pts = W.uniform_points(P, d, seed=1234, dtype=np.float32)
    out = {"d": d, "n": n, "limits": g._limits.copy(), "N": g.N, "S": g.S,
           "levels": lv, "indices": iv, "rank_table": rank_tab, "gp2idx_roundtrip": rt,
           "l2idx": l2, "points": pts, "init_values": init_values,
           "dp": g.dp, "lidx": g.lidx}
    for fname in funcs:
        if fname == "reference":
            vals = init_values
        elif fname == "parabola":
            vals = W.f_prod_parabola(coords.astype(np.float64)).astype(np.float32)
        else:
            vals = W.f_gauss(coords.astype(np.float64)).astype(np.float32)
        g.set_values(vals)
        g.hierarchize("hierarchize0")
        alpha = g.get_values()
        # cross-variant agreement inside the reference (asg.c:1810-1928); the loop-
        # interchange variants 7/8 are only valid without a restriction (asg.c:1009-1176)
        variants = refbind.HIERARCHIZE[1:7] + ["omp_hierarchize3"]
        if limits is None:
            variants += ["hierarchize7", "hierarchize8"]
        for var in variants:
            g.set_values(vals)
            g.hierarchize(var)
            assert (g.get_values() == alpha).all(), (tag, fname, var)
        g.set_values(alpha)
        ev = g.evaluate(pts, "evaluate0")
        for var in refbind.EVALUATE[1:]:
            assert (g.evaluate(pts, var) == ev).all(), (tag, fname, var)
        out[f"values_{fname}"] = vals
        out[f"alpha_{fname}"] = alpha
        out[f"eval_{fname}"] = ev
    path = os.path.join(HERE, f"ref_{tag}.npz")
    np.savez_compressed(path, **out)
    print("wrote", path, g.N, "points")


def write_sampled(d, n, tag, samples=4096, P=64):
    g = refbind.RefGrid(d, n)
    rng = np.random.default_rng(99)
    offs = np.sort(rng.choice(g.N, size=samples, replace=False)).astype(np.int32)
    offs[0], offs[-1] = 0, g.N - 1
    lv = np.zeros((samples, d), np.int32)
    iv = np.zeros((samples, d), np.int32)
    for j, o in enumerate(offs):
        lv[j], iv[j] = g.idx2gp(int(o))
    vals = g.get_values()
    out = {"d": d, "n": n, "N": g.N, "S": g.S, "offsets": offs, "levels": lv,
           "indices": iv, "init_checksum": np.float64(vals.astype(np.float64).sum()),
           "init_sample": vals[offs]}
    if d <= 10:
        g.hierarchize("omp_hierarchize3")
        alpha = g.get_values()
        pts = W.uniform_points(P, d, seed=4321, dtype=np.float32)
        out.update({"alpha_checksum": np.float64(alpha.astype(np.float64).sum()),
                    "alpha_abs_checksum": np.float64(np.abs(alpha.astype(np.float64)).sum()),
                    "alpha_sample": alpha[offs], "points": pts,
                    "eval": g.evaluate(pts, "omp_evaluate5")})
    path = os.path.join(HERE, f"ref_{tag}.npz")
    np.savez_compressed(path, **out)
    print("wrote", path, g.N, "points")


def main():
    if not refbind.available():
        sys.exit("oracle/_ref/libasgref.so is missing: run `make -C oracle ref` first")
    write_small(2, 3, None, "d2n3")
    write_small(3, 5, None, "d3n5")
    write_small(4, 4, None, "d4n4")
    write_small(5, 3, None, "d5n3")
    write_small(2, 6, None, "d2n6")
    write_small(2, 3, [2, 1], "d2n3_lim21")
    write_small(3, 5, [4, 2, 3], "d3n5_lim423")
    write_small(4, 5, [4, 1, 4, 2], "d4n5_lim4142")
    write_sampled(10, 8, "d10n8_sampled")
    write_sampled(20, 6, "d20n6_sampled")


if __name__ == "__main__":
    main()
