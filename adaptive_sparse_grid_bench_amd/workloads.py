"""Synthetic inputs of the benchmark workloads: seeded query points and the test
functions sampled at grid points.

Query points are FP32-representable values k * 2^-24 with k in [0, 2^24), so x = 1 never
occurs (the reference has no guard for it, asg.c:1259) and FP32 and FP64 paths see the
same x.  The generator is counter-based splitmix64: value j of stream `seed` is
mix64(seed * 2^32 + j), so any slice can be regenerated independently (and on the GPU by
the same formula).
"""
from __future__ import annotations

import numpy as np

_GAMMA = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def mix64(z: np.ndarray) -> np.ndarray:
    """splitmix64 finaliser on uint64 (wrap-around arithmetic)."""
    z = np.asarray(z, np.uint64)
    with np.errstate(over="ignore"):
        z = (z + _GAMMA)
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        z = z ^ (z >> np.uint64(31))
    return z


def uniform_u24(count: int, seed: int, start: int = 0) -> np.ndarray:
    """count integers k in [0, 2^24): the top 24 bits of mix64(seed<<32 + j)."""
    j = np.arange(start, start + count, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = mix64((np.uint64(seed) << np.uint64(32)) + j)
    return (z >> np.uint64(40)).astype(np.int64)


def uniform_points(P: int, d: int, seed: int = 42, dtype=np.float64) -> np.ndarray:
    """[P][d] row-major points k * 2^-24 (exact in FP32 and FP64)."""
    k = uniform_u24(P * d, seed)
    return (k.astype(np.float64) * 2.0 ** -24).reshape(P, d).astype(dtype)


# --- test functions (FP64 unless stated) -------------------------------------------

def f_prod_parabola(x: np.ndarray) -> np.ndarray:
    """prod_i 4 x_i (1 - x_i): hierarchical surpluses are exactly 4^-|l|_1 (0-based l)."""
    return np.prod(4.0 * x * (1.0 - x), axis=1)


def f_gauss(x: np.ndarray) -> np.ndarray:
    """exp(-sum_i (x_i - 1/2)^2), the d=10, n=8 drop-in workload."""
    return np.exp(-np.sum((x - 0.5) ** 2, axis=1))


def f_reference(x: np.ndarray) -> np.ndarray:
    """sum_i x_i^2 (1 - x_i)^2, the reference benchmark's own function (asg.c:87-94)."""
    return np.sum(x * x * (1.0 - x) * (1.0 - x), axis=1)


def gaussian_sum_params(d: int, count: int = 8, seed: int = 7):
    """Centres uniform in [0.2, 0.8]^d, widths log-uniform in [0.05, 0.5], unit weights."""
    u = uniform_u24(count * d * 2, seed).astype(np.float64) * 2.0 ** -24
    c = 0.2 + 0.6 * u[: count * d].reshape(count, d)
    w = np.exp(np.log(0.05) + (np.log(0.5) - np.log(0.05)) * u[count * d:].reshape(count, d))
    return c, w


def f_gaussian_sum(x: np.ndarray, centers: np.ndarray, widths: np.ndarray) -> np.ndarray:
    """sum_g prod_k exp(-(x_k - c_gk)^2 / (2 s_gk^2)), the d=20, n=6 precision workload."""
    out = np.zeros(x.shape[0])
    for c, s in zip(centers, widths):
        out += np.exp(-np.sum((x - c) ** 2 / (2.0 * s * s), axis=1))
    return out


def f_sin_prod(x: np.ndarray, c: float) -> np.ndarray:
    """prod_i sin(pi x_i) * (1 + 0.1 c), the batched-grid workload."""
    return np.prod(np.sin(np.pi * x), axis=1) * (1.0 + 0.1 * c)


# --- refined-subspace (dimension-adaptive) level sets ---------------------------------

def _surplus_1d_gauss(l: int, t: int, c: float, sigma: float) -> float:
    """1-D hierarchical surplus of exp(-(x-c)^2 / (2 sigma^2)) at point (l, t)."""
    g = lambda x: np.exp(-(x - c) ** 2 / (2.0 * sigma * sigma))
    x = (2 * t + 1) / 2.0 ** (l + 1)
    h = 2.0 ** -(l + 1)
    left = g(x - h) if t > 0 else 0.0
    right = g(x + h) if t < (1 << l) - 1 else 0.0
    return g(x) - 0.5 * (left + right)


def adaptive_peak(d: int, seed: int):
    """Seeded peak centre, uniform in [0.3, 0.7]^d."""
    u = uniform_u24(d, seed).astype(np.float64) * 2.0 ** -24
    return 0.3 + 0.4 * u


def adaptive_level_set(d: int = 8, base_n: int = 3, extra: int = 4, seed: int = 11,
                       sigma: float = 0.15, max_subspaces: int = 2000, tol: float = 1e-4):
    """Dimension-adaptive refinement of a regular base grid towards a Gaussian peak.

    Starts from the regular set |l|_1 <= base_n - 1 and greedily adds admissible forward
    neighbours l + e_k (every backward neighbour present, l_k + 1 <= base_n - 1 + extra)
    in order of the surplus magnitude of f = exp(-|x - c|^2 / (2 sigma^2)) at the cell of
    subspace l + e_k containing the peak c (the largest surplus of that subspace, a
    product of 1-D surpluses), until max_subspaces or the indicator drops below tol.
    Returns (levels [S][d] uint8, centre c).  The set is downward closed by construction.
    """
    import heapq
    c = adaptive_peak(d, seed)
    cap = base_n - 1 + extra
    cache = {}

    def indicator(l):
        v = 1.0
        for k, lk in enumerate(l):
            key = (k, lk)
            if key not in cache:
                t = min(int(np.floor(c[k] * 2 ** lk)), (1 << lk) - 1)
                cache[key] = abs(_surplus_1d_gauss(lk, t, c[k], sigma))
            v *= cache[key]
        return v

    from itertools import product
    current = {l for l in product(range(base_n), repeat=d) if sum(l) <= base_n - 1}
    heap, seen = [], set()

    def push_neighbours(l):
        for k in range(d):
            if l[k] + 1 > cap:
                continue
            m = list(l)
            m[k] += 1
            m = tuple(m)
            if m in current or m in seen:
                continue
            if all(m[j] == 0 or tuple(m[i] - (i == j) for i in range(d)) in current
                   for j in range(d)):
                seen.add(m)
                heapq.heappush(heap, (-indicator(m), m))

    for l in sorted(current):
        push_neighbours(l)
    while heap and len(current) < max_subspaces:
        neg, m = heapq.heappop(heap)
        if -neg < tol:
            break
        current.add(m)
        push_neighbours(m)      # every newly admissible set member has m as a parent
    return np.array(sorted(current), np.uint8), c


# --- the batched workload ---------------------------------------------------------------

def batched_specs(count: int = 256, seed: int = 5):
    """Per grid: d ~ U{2..12}, n ~ U{3..8}, c_g ~ U[0, 1), points floor(2^U[12, 18])."""
    u = uniform_u24(count * 4, seed).astype(np.float64) * 2.0 ** -24
    u = u.reshape(count, 4)
    d = 2 + np.floor(u[:, 0] * 11).astype(int)
    n = 3 + np.floor(u[:, 1] * 6).astype(int)
    c = u[:, 2]
    P = np.floor(2.0 ** (12 + 6 * u[:, 3])).astype(np.int64)
    return [dict(d=int(d[i]), n=int(n[i]), c=float(c[i]), P=int(P[i])) for i in range(count)]
