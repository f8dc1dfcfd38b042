/*
 * Floating-point part of the CPU oracle, included twice by sg_oracle.c with
 * REAL = double (suffix _f64) and REAL = float (suffix _f32).  TEST INFRASTRUCTURE
 * ONLY (see sg_oracle.c).  Built with -ffp-contract=off so that every operation is
 * rounded exactly as written, as the reference's SSE build does.
 */

/* hierarchize0 semantics, asg.c:557-571: for each dimension, walk the offsets from
 * N-1 down to 0 and subtract the mean of the two parents read through the full
 * bijection.  O(d*N*(d+n)); for small grids and for pinning the faster form below. */
void SFX(sgo_hierarchize0)(int d, int n, const i64 *dp, const i64 *lidx, i64 N, REAL *v)
{
    int *lv = (int *)malloc(sizeof(int) * d), *iv = (int *)malloc(sizeof(int) * d);
    for (int k = 0; k < d; ++k) {
        for (i64 j = N - 1; j >= 0; --j) {
            sgo_idx2gp(d, n, dp, lidx, j, lv, iv);
            int l = lv[k], i = iv[k], pl, pi;
            REAL left = (REAL)0, right = (REAL)0;
            if (sgo_left_parent(l, i, &pl, &pi)) {
                lv[k] = pl; iv[k] = pi;
                left = v[sgo_gp2idx(d, n, dp, lidx, lv, iv)];
                lv[k] = l; iv[k] = i;
            }
            if (sgo_right_parent(l, i, &pl, &pi)) {
                lv[k] = pl; iv[k] = pi;
                right = v[sgo_gp2idx(d, n, dp, lidx, lv, iv)];
                lv[k] = l; iv[k] = i;
            }
            v[j] -= (left + right) * (REAL)0.5;
        }
    }
    free(lv);
    free(iv);
}

/* Level-group hierarchization, the order of hierarchize3/omp_hierarchize3
 * (asg.c:660-708, 1180-1233): per dimension, level sums from n-1 down to 0, every
 * subspace of a level sum independently (parents always lie in lower level sums, so
 * they are read before this sweep updates them).  The parent address is formed in
 * O(1) from the in-block index as in asg.c:878-918.  Same single operation per point
 * per dimension as hierarchize0, hence bit-identical to it. */
void SFX(sgo_hierarchize)(int d, int n, const i64 *dp, const i64 *lidx, i64 N, REAL *v)
{
    (void)N;
    for (int k = 0; k < d; ++k) {
        for (int s = n - 1; s >= 0; --s) {
            i64 cnt = dp[(d - 1) * n + s];
            #pragma omp parallel for schedule(dynamic, 16)
            for (i64 rank = 0; rank < cnt; ++rank) {
                int lv[64];
                i64 par_base[64];
                sgo_idx2l(d, n, dp, rank, s, lv);
                int lk = lv[k];
                int below = 0;                       /* sum of levels of dims > k */
                for (int j = k + 1; j < d; ++j) below += lv[j];
                /* base offset of the subspace with l_k replaced by pl, for pl < lk */
                for (int pl = 0; pl < lk; ++pl) {
                    lv[k] = pl;
                    int ps = s - lk + pl;
                    par_base[pl] = lidx[ps] + (sgo_l2idx(d, n, dp, lv) << ps);
                }
                lv[k] = lk;
                i64 base = lidx[s] + (rank << s);
                i64 npts = ((i64)1) << s;
                for (i64 r = 0; r < npts; ++r) {
                    i64 lo = r & ((((i64)1) << below) - 1);
                    int ik = (int)((r >> below) & ((((i64)1) << lk) - 1));
                    i64 hi = r >> (below + lk);
                    int pl, pi;
                    REAL left = (REAL)0, right = (REAL)0;
                    if (sgo_left_parent(lk, ik, &pl, &pi))
                        left = v[par_base[pl] + (hi << (below + pl)) + ((i64)pi << below) + lo];
                    if (sgo_right_parent(lk, ik, &pl, &pi))
                        right = v[par_base[pl] + (hi << (below + pl)) + ((i64)pi << below) + lo];
                    v[base + r] -= (left + right) * (REAL)0.5;
                }
            }
        }
    }
}

/* evaluate0 semantics, asg.c:1237-1274: for each point, sum over level sums ascending
 * and ranks ascending of  (prod_k hat_k) * alpha[gather], in exactly the reference's
 * operation order.  The level vectors are listed once up front (asg.c calls idx2l per
 * point and subspace; the values are the same).  Points are independent, so they are
 * split over OpenMP threads without changing any result.
 * Preconditions as in the reference: 0 <= x < 1. */
void SFX(sgo_evaluate)(int d, int n, const i64 *dp, const i64 *lidx, const REAL *alpha,
                       const REAL *coords, i64 P, REAL *out)
{
    i64 S = 0;
    for (int s = 0; s < n; ++s) S += dp[(d - 1) * n + s];
    int *levels = (int *)malloc(sizeof(int) * (size_t)(S * d));
    int *sums = (int *)malloc(sizeof(int) * (size_t)S);
    i64 q = 0;
    for (int s = 0; s < n; ++s)
        for (i64 rank = 0; rank < dp[(d - 1) * n + s]; ++rank, ++q) {
            sgo_idx2l(d, n, dp, rank, s, levels + q * d);
            sums[q] = s;
        }
    (void)lidx;
    #pragma omp parallel for schedule(static)
    for (i64 c = 0; c < P; ++c) {
        const REAL *x = coords + c * d;
        REAL acc = (REAL)0;
        i64 base = 0;
        for (i64 j = 0; j < S; ++j) {
            const int *lv = levels + j * d;
            REAL prod = (REAL)1;
            i64 index2 = 0;
            for (int k = 0; k < d; ++k) {
                REAL div = ((REAL)1 - (REAL)0) / (REAL)(1 << lv[k]);
                int t = (int)((x[k] - (REAL)0) / div);
                index2 = index2 * (1 << lv[k]) + t;
                REAL left = (REAL)t * div;
                REAL m = ((REAL)2 * (x[k] - left) - div) / div;
                prod *= (REAL)1 + m * ((REAL)1 - (REAL)((m >= (REAL)0) << 1));
            }
            prod *= alpha[base + index2];
            acc += prod;
            base += ((i64)1) << sums[j];
        }
        out[c] = acc;
    }
    free(levels);
    free(sums);
}

/* Brute-force interpolant: sum over every grid point (offset order) of
 * alpha_j * prod_k max(0, 1 - |2^(l+1) x_k - 2 i_k - 1|).  Uses the bijection only to
 * name each point's (l, i), not the evaluation walk.  O(P*N*d): tiny grids only. */
void SFX(sgo_evaluate_brute)(int d, int n, const i64 *dp, const i64 *lidx, i64 N,
                             const REAL *alpha, const REAL *coords, i64 P, REAL *out)
{
    int *lv = (int *)malloc(sizeof(int) * (size_t)(N * d));
    int *iv = (int *)malloc(sizeof(int) * (size_t)(N * d));
    for (i64 j = 0; j < N; ++j) sgo_idx2gp(d, n, dp, lidx, j, lv + j * d, iv + j * d);
    #pragma omp parallel for schedule(static)
    for (i64 c = 0; c < P; ++c) {
        const REAL *x = coords + c * d;
        REAL acc = (REAL)0;
        for (i64 j = 0; j < N; ++j) {
            REAL prod = alpha[j];
            for (int k = 0; k < d && prod != (REAL)0; ++k) {
                REAL y = x[k] * (REAL)(((i64)2) << lv[j * d + k]);
                REAL h = (REAL)1 - (REAL)fabs((double)(y - (REAL)(2 * iv[j * d + k] + 1)));
                prod *= (h > (REAL)0) ? h : (REAL)0;
            }
            acc += prod;
        }
        out[c] = acc;
    }
    free(lv);
    free(iv);
}
