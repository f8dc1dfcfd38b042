/*
 * Floating-point part of the explicit-level-set oracle, included twice by
 * sg_oracle.c.  TEST INFRASTRUCTURE ONLY.
 */

/* Level-group hierarchization over an explicit set (asg.c:660-708 order: subspaces
 * from the last to the first, i.e. level sums descending; parents lie in lower sums).
 * Identical operation per point to the restricted-regular form. */
void SFX(sgo_x_hierarchize)(int d, i64 S, const unsigned char *levels, const i64 *base,
                            REAL *v)
{
    for (int k = 0; k < d; ++k) {
        for (i64 q = S - 1; q >= 0; --q) {
            const unsigned char *lv = levels + q * d;
            int lk = lv[k];
            if (lk == 0) continue;
            int s = 0, below = 0;
            for (int j = 0; j < d; ++j) {
                s += lv[j];
                if (j > k) below += lv[j];
            }
            unsigned char pl_lv[64];
            i64 par_base[64];
            for (int j = 0; j < d; ++j) pl_lv[j] = lv[j];
            for (int pl = 0; pl < lk; ++pl) {
                pl_lv[k] = (unsigned char)pl;
                i64 id = sgo_x_find(d, S, levels, pl_lv);
                par_base[pl] = id < 0 ? -1 : base[id];
            }
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
                v[base[q] + r] -= (left + right) * (REAL)0.5;
            }
        }
    }
}

/* evaluate0 operation order (asg.c:1250-1268) over the list in compact order */
void SFX(sgo_x_evaluate)(int d, i64 S, const unsigned char *levels, const i64 *base,
                         const REAL *alpha, const REAL *coords, i64 P, REAL *out)
{
    #pragma omp parallel for schedule(static)
    for (i64 c = 0; c < P; ++c) {
        const REAL *x = coords + c * d;
        REAL acc = (REAL)0;
        for (i64 q = 0; q < S; ++q) {
            const unsigned char *lv = levels + q * d;
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
            prod *= alpha[base[q] + index2];
            acc += prod;
        }
        out[c] = acc;
    }
}
