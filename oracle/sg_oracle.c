/*
 * CPU oracle for the compact sparse grid's index bijection, hierarchization and
 * evaluation.  TEST INFRASTRUCTURE ONLY: the test suite, the smoke check and the
 * benchmark's CPU-baseline leg load this library as the checker.  No product code
 * path links it or calls into it.
 *
 * It restates the algorithm of the reference C benchmark (adaptive_sparse_grid_bench.c,
 * cited below as asg.c) in plain C, in FP64 and in FP32:
 *
 *   tables        asg.c:101-140  (dp matrix under a per-dimension restriction vector,
 *                                 per-level-sum start offsets, N and S)
 *   gp2idx        asg.c:187-207
 *   idx2gp        asg.c:210-243
 *   idx2l/l2idx   asg.c:245-279
 *   parents       asg.c:282-344, 828-854
 *   hierarchize   asg.c:557-571 (hierarchize0, descending offsets)
 *                 asg.c:660-708 (hierarchize3, level groups from the top level sum down)
 *   evaluate      asg.c:1237-1274 (evaluate0, reference summation order)
 *
 * Differences from asg.c, all deliberate:
 *   - dp[0][s] for s > limits[0] is 0 (asg.c:110-115 leaves it uninitialised);
 *   - offsets are int64 so that overflow cannot go unnoticed;
 *   - the library is stateless: every call takes the shape tables explicitly.
 *
 * Parity status of each function and what pins it is listed in the test files
 * (tests/test_oracle.py) and in the project design notes.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>

typedef int64_t i64;

/* ---------------------------------------------------------------- tables --- */

/* dp[i*n + s] = number of level vectors with i+1 components, sum s, component j <=
 * limits[j].  lidx[s] = first offset of level sum s; lidx[n] = N.  Returns 0, or -1
 * if N or S overflows int64 (cannot happen for the shapes the reference supports).
 * asg.c:108-140; the recurrence is the direct form of asg.cu:79-89. */
int sgo_tables(int d, int n, const int *limits, i64 *dp, i64 *lidx, i64 *N, i64 *S)
{
    int i, s, k;
    if (d < 1 || n < 1) return -1;
    for (s = 0; s < n; ++s) dp[s] = (s <= limits[0]) ? 1 : 0;
    for (i = 1; i < d; ++i) {
        for (s = 0; s < n; ++s) {
            i64 acc = 0;
            for (k = 0; k <= s && k <= limits[i]; ++k) acc += dp[(i - 1) * n + (s - k)];
            dp[i * n + s] = acc;
        }
    }
    i64 off = 0, cnt = 0;
    for (s = 0; s < n; ++s) {
        lidx[s] = off;
        off += dp[(d - 1) * n + s] << s;
        cnt += dp[(d - 1) * n + s];
    }
    lidx[n] = off;
    *N = off;
    *S = cnt;
    return 0;
}

/* ------------------------------------------------------------- bijection --- */

/* rank of a level vector among the vectors of its level sum; asg.c:266-279 */
i64 sgo_l2idx(int d, int n, const i64 *dp, const int *levels)
{
    i64 rank = 0;
    int sum = levels[0];
    for (int i = 1; i < d; ++i) {
        sum += levels[i];
        for (int j = 0; j < levels[i]; ++j) rank += dp[(i - 1) * n + (sum - j)];
    }
    return rank;
}

/* level vector of the given rank within level sum `sum`; asg.c:245-263 */
void sgo_idx2l(int d, int n, const i64 *dp, i64 rank, int sum, int *levels)
{
    for (int i = d - 1; i >= 1; --i) {
        int l = 0;
        while (rank >= dp[(i - 1) * n + (sum - l)]) {
            rank -= dp[(i - 1) * n + (sum - l)];
            ++l;
        }
        levels[i] = l;
        sum -= l;
    }
    levels[0] = sum;
}

/* (levels, indices) -> compact offset; asg.c:187-207 */
i64 sgo_gp2idx(int d, int n, const i64 *dp, const i64 *lidx, const int *levels,
               const int *indices)
{
    i64 in_block = indices[0];
    int sum = levels[0];
    for (int i = 1; i < d; ++i) {
        in_block = (in_block << levels[i]) + indices[i];
        sum += levels[i];
    }
    return lidx[sum] + (sgo_l2idx(d, n, dp, levels) << sum) + in_block;
}

/* compact offset -> (levels, indices); asg.c:210-243 */
void sgo_idx2gp(int d, int n, const i64 *dp, const i64 *lidx, i64 index, int *levels,
                int *indices)
{
    int s = 0;
    while (s + 1 < n && index >= lidx[s + 1]) ++s;
    i64 r = index - lidx[s];
    i64 bits = r & ((((i64)1) << s) - 1);
    i64 rank = r >> s;
    int sum = s;
    for (int i = d - 1; i >= 1; --i) {
        int l = 0;
        while (rank >= dp[(i - 1) * n + (sum - l)]) {
            rank -= dp[(i - 1) * n + (sum - l)];
            ++l;
        }
        levels[i] = l;
        indices[i] = (int)(bits & ((((i64)1) << l) - 1));
        bits >>= l;
        sum -= l;
    }
    levels[0] = sum;
    indices[0] = (int)(bits & ((((i64)1) << sum) - 1));
}

/* idx2gp over every offset 0..N-1, row-major [N][d] outputs */
void sgo_idx2gp_all(int d, int n, const i64 *dp, const i64 *lidx, i64 N, int *levels,
                    int *indices)
{
    #pragma omp parallel for schedule(static)
    for (i64 o = 0; o < N; ++o)
        sgo_idx2gp(d, n, dp, lidx, o, levels + o * d, indices + o * d);
}

/* gp2idx over M (levels, indices) rows */
void sgo_gp2idx_many(int d, int n, const i64 *dp, const i64 *lidx, i64 M,
                     const int *levels, const int *indices, i64 *out)
{
    #pragma omp parallel for schedule(static)
    for (i64 m = 0; m < M; ++m)
        out[m] = sgo_gp2idx(d, n, dp, lidx, levels + m * d, indices + m * d);
}

/* hierarchical parents in one dimension; asg.c:828-854.  Returns 0 when the parent
 * does not exist (index 0 has no left parent, index 2^l-1 no right parent), else 1. */
int sgo_left_parent(int l, int i, int *pl, int *pi)
{
    if (i == 0) return 0;
    int ll = l - 1, li = i;
    while ((li & 1) == 0) { li >>= 1; --ll; }
    *pl = ll;
    *pi = li >> 1;
    return 1;
}

int sgo_right_parent(int l, int i, int *pl, int *pi)
{
    if (i == (1 << l) - 1) return 0;
    int rl = l - 1, ri = i + 1;
    while ((ri & 1) == 0) { ri >>= 1; --rl; }
    *pl = rl;
    *pi = ri >> 1;
    return 1;
}

/* ------------------------------------------------- FP64 and FP32 numerics --- */

#define REAL double
#define SFX(name) name##_f64
#include "sg_oracle_fp.h"
#undef REAL
#undef SFX

#define REAL float
#define SFX(name) name##_f32
#include "sg_oracle_fp.h"
#undef REAL
#undef SFX

/* --------------------------------------------- explicit (refined) level sets --- */
/* The same operations over an explicit downward-closed set of level vectors given in
 * compact order (level sum ascending, then l_{d-1}, l_{d-2}, ... ascending), the
 * generalisation of the restriction vector (asg.c:82-83, 113-129).  Offsets follow the
 * same compact rule: subspace blocks of 2^|l| values in list order. */

static int sgo_x_cmp(const unsigned char *a, const unsigned char *b, int d)
{
    int sa = 0, sb = 0;
    for (int j = 0; j < d; ++j) { sa += a[j]; sb += b[j]; }
    if (sa != sb) return sa < sb ? -1 : 1;
    for (int j = d - 1; j >= 0; --j)
        if (a[j] != b[j]) return a[j] < b[j] ? -1 : 1;
    return 0;
}

/* index of level vector l in the sorted list, or -1 */
i64 sgo_x_find(int d, i64 S, const unsigned char *levels, const unsigned char *l)
{
    i64 lo = 0, hi = S;
    while (lo < hi) {
        i64 mid = (lo + hi) / 2;
        int c = sgo_x_cmp(levels + mid * d, l, d);
        if (c < 0) lo = mid + 1;
        else hi = mid;
    }
    if (lo < S && sgo_x_cmp(levels + lo * d, l, d) == 0) return lo;
    return -1;
}

/* base offsets of the subspaces and N; returns -1 if the list is not sorted */
i64 sgo_x_offsets(int d, i64 S, const unsigned char *levels, i64 *base)
{
    i64 off = 0;
    for (i64 q = 0; q < S; ++q) {
        if (q > 0 && sgo_x_cmp(levels + (q - 1) * d, levels + q * d, d) >= 0) return -1;
        int s = 0;
        for (int j = 0; j < d; ++j) s += levels[q * d + j];
        base[q] = off;
        off += ((i64)1) << s;
    }
    return off;
}

#define REAL double
#define SFX(name) name##_f64
#include "sg_oracle_x.h"
#undef REAL
#undef SFX

#define REAL float
#define SFX(name) name##_f32
#include "sg_oracle_x.h"
#undef REAL
#undef SFX
