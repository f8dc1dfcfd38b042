/*
 * asg_compat.h — drop-in replacement of the reference benchmark's global-state C
 * interface (adaptive_sparse_grid_bench.c, "asg.c"), served by libasg_compat.so on top
 * of libsgrid.so.  A program written against asg.c's functions and globals links this
 * library instead; its host buffers stay where they are and the work runs on the GPU.
 *
 * Semantics follow asg.c exactly, FP32 as the reference computes:
 *   - the caller sets num_dims, num_levels and limits, then calls init(), which fills
 *     dp_mat, lidx_vect, num_grid_points, num_subspaces, par_lili and sg1d (the
 *     reference's test function sampled at the grid points, asg.c:143-152);
 *   - hierarchize*() turn sg1d into surpluses in place; the result is bit-identical to
 *     hierarchize0 (asg.c:557-571), as every CPU variant of the reference is
 *     (asg.c:1810-1928);
 *   - evaluate*(sg1d, coord_mat, out, num_evals) overwrite out with the interpolant at
 *     the row-major [num_evals][num_dims] points, bit-identical to evaluate0
 *     (asg.c:1237-1274), as every CPU variant is (asg.c:1946-2023).
 * Return values are those of asg.c (0).  A GPU failure prints the library's message to
 * stderr and returns -1 (asg.c has no failure path).  Like asg.c the interface is
 * process-global and not re-entrant.
 */
#ifndef ASG_COMPAT_H
#define ASG_COMPAT_H

#ifdef __cplusplus
extern "C" {
#endif

extern int num_levels, num_dims, num_evals, num_subspaces;   /* asg.c:77 */
extern int num_grid_points;                                  /* asg.c:78 */
extern float *sg1d;                                          /* asg.c:79 */
extern int **par_lili;                                       /* asg.c:80 */
extern int *limits;                                          /* asg.c:83 */
extern int **dp_mat;                                         /* asg.c:84 */
extern int *lidx_vect;                                       /* asg.c:85 */

void init(void);                                             /* asg.c:101 */
int gp2idx(int *levels, int *indices);                       /* asg.c:187 */
int idx2gp(int index, int *levels, int *indices);            /* asg.c:210 */
int idx2l(int index, int sum, int *levels);                  /* asg.c:245 */
int l2idx(int *levels);                                      /* asg.c:266 */
int get_left_li(int l, int i, int *ll, int *li);             /* asg.c:828 */
int get_right_li(int l, int i, int *rl, int *ri);            /* asg.c:842 */

int hierarchize0(void);                                      /* asg.c:557 */
int hierarchize1(void);
int hierarchize2(void);
int hierarchize3(void);
int hierarchize4(void);
int hierarchize5(void);
int hierarchize6(void);
int hierarchize7(void);
int hierarchize8(void);
int omp_hierarchize3(void);                                  /* asg.c:1180 */

int evaluate0(float *sg1d, float *coord_mat, float *out, int num_evals);  /* asg.c:1237 */
int evaluate1(float *sg1d, float *coord_mat, float *out, int num_evals);
int evaluate2(float *sg1d, float *coord_mat, float *out, int num_evals);
int evaluate3(float *sg1d, float *coord_mat, float *out, int num_evals);
int evaluate4(float *sg1d, float *coord_mat, float *out, int num_evals);
int evaluate5(float *sg1d, float *coord_mat, float *out, int num_evals);
int omp_evaluate2(float *sg1d, float *coord_mat, float *out, int num_evals); /* asg.c:1691 */
int omp_evaluate5(float *sg1d, float *coord_mat, float *out, int num_evals); /* asg.c:1708 */

/* FP64 twins on the grid init() built (the reference computes in FP32 only):
 *   init_values_f64     values[N] = fct at the grid points in FP64 (asg.c:87-94, 143-152)
 *   hierarchize_f64     values[N] in place; hierarchize0 (asg.c:557-571) in FP64, exactly
 *   evaluate_f64        evaluate0's operation order (asg.c:1237-1274) in FP64, exactly
 *   evaluate_walk_f64   the depth-first walk: within 1e-12 of evaluate_f64 relative to
 *                       the sum of |terms|, much faster
 * Same argument layouts and return values as the FP32 functions. */
int init_values_f64(double *values);
int hierarchize_f64(double *values);
int evaluate_f64(double *values, double *coord_mat, double *out, int num_evals);
int evaluate_walk_f64(double *values, double *coord_mat, double *out, int num_evals);

#ifdef __cplusplus
}
#endif

#endif /* ASG_COMPAT_H */
