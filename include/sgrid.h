/*
 * sgrid.h — C ABI of the MI355X compact sparse grid library (libsgrid.so).
 *
 * The library holds a compact sparse grid (no boundary points, 0-based levels) in
 * HBM and runs the reference benchmark's hot path on one AMD Instinct MI355X
 * (gfx950): the offset <-> (levels, indices) bijection, hierarchization and
 * evaluation.  Each entry point cites the reference interface it replaces; paths are
 * relative to the reference repository adeftu/adaptive_sparse_grid_bench:
 *   asg.c  = adaptive_sparse_grid_bench.c
 *   asg.cu = adaptive_sparse_grid_bench_cuda.cu
 *
 * Conventions
 *   - Every function returns an int status: SG_OK (0) or a negative SG_E_* code.
 *     sg_last_error() returns a thread-local message for the last failure.
 *   - d_* arguments are device pointers on the device that was current when the
 *     grid was created; other pointers are host memory.  Host buffers stay owned by
 *     the caller, device tables by the grid handle.
 *   - stream is a hipStream_t passed as void* (NULL = the null stream).  Device calls
 *     are asynchronous on that stream and perform no device-wide synchronisation.
 *   - Grid values use the reference's compact order (asg.c:187-243): level sum
 *     ascending, then rank within the level sum (asg.c:266-279), then the row-major
 *     in-block index with dimension 0 most significant.  Query points are row-major
 *     [P][d] (asg.c:1259) with 0 <= x < 1.
 *   - A handle may be used from several host threads at once: after creation the
 *     calls only read it.
 */
#ifndef SGRID_H
#define SGRID_H

#ifdef __cplusplus
extern "C" {
#endif

#define SG_OK 0
#define SG_E_INVALID (-1)     /* bad argument or precondition */
#define SG_E_NOMEM (-2)       /* device or host allocation failed */
#define SG_E_HIP (-3)         /* HIP runtime error */
#define SG_E_RANGE (-4)       /* offset, level or size out of range */
#define SG_E_UNSUPPORTED (-5) /* shape outside the kernels' limits */

#define SG_MAX_DIMS 20        /* asg.cu:17 */

typedef struct sg_grid sg_grid;

const char *sg_last_error(void);
const char *sg_status_string(int status);
int sg_version(void);   /* changes whenever kernel behaviour or timing may change */

/* ------------------------------------------------------------------ grids ---- */

/* Restricted-regular grid { l : |l|_1 <= num_levels-1, l_k <= limits[k] } with its
 * shape tables and launch tables, uploaded to the current device.
 * Replaces init() on the globals num_dims/num_levels/limits (asg.c:101-184) and the
 * CUDA host set-up init_restr_vec/init_dp_mat/init_lidx_vect (asg.cu:72-102).
 * limits == NULL means num_levels-1 in every dimension (asg.c:1760-1763). */
int sg_grid_create(int num_dims, int num_levels, const int *limits, sg_grid **out);

/* Refined-subspace grid: an explicit downward-closed set of `count` level vectors
 * (levels[count][num_dims], 0-based, each <= 15).  Generalises the restriction vector
 * (asg.c:82-83); a restricted-regular set given this way gets the identical order. */
int sg_grid_create_explicit(int num_dims, const unsigned char *levels, long long count,
                            sg_grid **out);

int sg_grid_destroy(sg_grid *grid);

/* num_dims, num_levels (level sums 0..num_levels-1), N and S (asg.c:131-140). */
int sg_grid_info(const sg_grid *grid, int *num_dims, int *num_levels,
                 long long *num_grid_points, long long *num_subspaces);

/* dp matrix [d][n] (restricted-regular grids only; NULL to skip) and the level-sum
 * start offsets lidx[n+1] (asg.c:108-140 dp_mat, lidx_vect). */
int sg_grid_tables(const sg_grid *grid, long long *dp, long long *lidx);

/* level vectors [S][d] and base offsets [S] of the subspaces in compact order. */
int sg_grid_subspaces(const sg_grid *grid, unsigned char *levels, long long *base);

/* Launch tuning of one grid.  sg_grid_get_tuning returns the current values (defaults
 * chosen from the grid's shape); sg_grid_set_tuning replaces them (not concurrently
 * with other calls on the same grid).  The autotuner (package module autotune.py)
 * times the candidates on the GPU and stores its choice per grid shape.
 *   eval_block        threads per workgroup of the evaluate kernels: 64, 128 or 256
 *                     (the staged walks choose their own: 256, and 512 for the
 *                     block-staged walk of n = 8 grids at d = 9..11 and n = 7 grids at
 *                     d = 10, 11, which runs at 4 waves per SIMD);
 *   eval_variant_f64  walk variant of the FP64 evaluate,
 *   eval_variant_f32  and of the FP32 one: 0..SG_EVAL_VARIANTS-1.  0..2 gather each
 *                     coefficient per lane from HBM/cache (they differ in tabled
 *                     dimensions, gathers per batch and waves per SIMD); 3 is the staged
 *                     walk, where the workgroup copies surplus blocks into LDS and the
 *                     lanes read them there (blocks too large for a 26 KiB stage are
 *                     gathered as in 0..2); 4 is the block-staged walk of regular grids
 *                     (blocks at compile-time LDS offsets; a sparse block order from
 *                     d = 12 in FP64 and d = 15 in FP32); 5 is the block-staged walk with
 *                     two query points per lane (FP64 below d = 12; elsewhere it runs as
 *                     4).  A grid a variant does not cover runs the next more general
 *                     one (5 -> 4 -> 3 -> 0; 1 and 2 -> 0 on explicit sets and where
 *                     their tabled dimensions reach level 8);
 *   bin_min_points    evaluate launches with at least this many points bin the points
 *                     spatially first (<= 0: never; the staged variants never bin);
 *   hier_tile_values  target values per hierarchization workgroup tile (>= 64). */
#define SG_EVAL_VARIANTS 6
typedef struct sg_tuning {
    int eval_block;
    int eval_variant_f64;
    int eval_variant_f32;
    int bin_min_points;
    int hier_tile_values;
} sg_tuning;
int sg_grid_get_tuning(const sg_grid *grid, sg_tuning *tuning);
int sg_grid_set_tuning(sg_grid *grid, const sg_tuning *tuning);
/* The walk variant a single-grid evaluate of this grid runs when its tuning names
 * `variant` (elem_bytes 8: sg_evaluate_f64, 4: sg_evaluate_f32c), after the fallback
 * above: *walk = variant where the grid is covered, else the more general one that runs.
 * Host only.  Lets a tuner time each distinct walk once.  No counterpart in the
 * reference (it has one kernel per variant name). */
int sg_grid_walk_variant(const sg_grid *grid, int elem_bytes, int variant, int *walk);

/* ------------------------------------------------------ host bijection ------- */

int sg_gp2idx(const sg_grid *grid, const int *levels, const int *indices,
              long long *index);                                 /* asg.c:187-207 */
int sg_idx2gp(const sg_grid *grid, long long index, int *levels,
              int *indices);                                     /* asg.c:210-243 */
int sg_idx2l(const sg_grid *grid, long long rank, int level_sum,
             int *levels);                                       /* asg.c:245-263 */
int sg_l2idx(const sg_grid *grid, const int *levels, long long *rank); /* asg.c:266-279 */

/* ---------------------------------------------------- device bijection ------- */

/* idx2gp for `count` offsets (d_index == NULL: offsets 0..count-1) into row-major
 * [count][d] int32 outputs; an offset outside [0, N) yields -1 entries.
 * Device form of asg.cu:218-259 (idx2gp). */
int sg_idx2gp_dev(const sg_grid *grid, const int *d_index, long long count, int *d_levels,
                  int *d_indices, void *stream);

/* gp2idx for `count` (levels, indices) rows; a point not in the grid yields -1.
 * Device form of asg.cu:189-214 (gp2idx). */
int sg_gp2idx_dev(const sg_grid *grid, const int *d_levels, const int *d_indices,
                  long long count, int *d_index, void *stream);

/* Coordinates (2 i_k + 1) / 2^(l_k + 1) of every grid point, [N][d] FP64
 * (asg.c:143-152, asg.cu:105-117). */
int sg_grid_coords_f64(const sg_grid *grid, double *d_coords, void *stream);

/* ------------------------------------------------------- hierarchize -------- */

/* Nodal values -> hierarchical surpluses in place, d_values[N].  One sweep per
 * dimension, v -= (left parent + right parent) / 2, parents read before their own
 * update.  Bit-identical to hierarchize0 (asg.c:557-571) and every other CPU variant
 * (asg.c:1810-1928); replaces hierarchize0..8 / omp_hierarchize3 (asg.c:557-1233) and
 * the CUDA launches hierarchize0..7 (asg.cu:448-1160, 1720-1806). */
int sg_hierarchize_f64(const sg_grid *grid, double *d_values, void *stream);
int sg_hierarchize_f32(const sg_grid *grid, float *d_values, void *stream);

/* --------------------------------------------------------- evaluate --------- */

/* Interpolant at P points: d_out[p] = sum over subspaces of prod_k hat_k * surplus.
 * Replaces evaluate0..5 / omp_evaluate2/5 (asg.c:1237-1722; out overwritten as in
 * asg.c:1243-1244) and the CUDA evaluate0..5 (asg.cu:1169-1598).
 *   _f64   FP64, depth-first walk with incremental hat products; agrees with
 *          evaluate0's value within 1e-12 relative to sum |terms|.
 *   _f32c  FP32 surpluses and points, Kahan-compensated accumulation.
 *   _ref_* the reference's exact operation order: bit-identical to evaluate0 in the
 *          same precision (slower; what the asg_compat shims call). */
int sg_evaluate_f64(const sg_grid *grid, const double *d_alpha, const double *d_coords,
                    long long num_evals, double *d_out, void *stream);
int sg_evaluate_f32c(const sg_grid *grid, const float *d_alpha, const float *d_coords,
                     long long num_evals, float *d_out, void *stream);
int sg_evaluate_ref_f64(const sg_grid *grid, const double *d_alpha, const double *d_coords,
                        long long num_evals, double *d_out, void *stream);
int sg_evaluate_ref_f32(const sg_grid *grid, const float *d_alpha, const float *d_coords,
                        long long num_evals, float *d_out, void *stream);

/* sg_evaluate_f64 without any device allocation.  No counterpart in the reference: it is
 * for callers that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  The point-binning workspace that
 * sg_evaluate_f64 takes from stream-ordered allocations comes from a fixed device array
 * of the library instead (about 16.3 MiB, held by each device the library runs on until
 * the process ends), so:
 *   - num_evals is at most SG_NOALLOC_MAX_POINTS, else SG_E_RANGE;
 *   - the grid's device tables must already be uploaded by an earlier device call on
 *     the grid (any of sg_hierarchize_*, sg_evaluate_*, ...), else SG_E_INVALID;
 *   - every call on one device shares that device's array: calls that overlap on one
 *     device (different streams, or host threads) race and give wrong results.
 * Results are bit-identical to sg_evaluate_f64's. */
#define SG_NOALLOC_MAX_POINTS 2097152LL
int sg_evaluate_f64_noalloc(const sg_grid *grid, const double *d_alpha,
                            const double *d_coords, long long num_evals, double *d_out,
                            void *stream);

/* ---------------------------------------------------------- batches --------- */
/* A batch of independent grids resident together (the batched form).  Surpluses of
 * grid g live at value_offsets[g] of one buffer; after sg_batch_set_points(counts),
 * grid g's points are row-major [counts[g]][d_g] starting at element coord_offsets[g]
 * of one coordinate buffer and its outputs at point_offsets[g] of one output buffer.
 * Hierarchization runs one launch per dimension sweep over the pole-family tiles of
 * every grid; evaluation one launch per group of grids with the same dimensions, over
 * device-resident work lists.  sg_batch_run_graph captures that sequence once per
 * buffer triple into a hipGraph (evaluation groups on parallel branches) and launches
 * it; later calls with the same buffers relaunch the same graph.  The grids must stay
 * alive while the batch exists; each grid's walk variants are taken from its tuning
 * when the batch is created.  Results are bit-identical to the per-grid calls. */
typedef struct sg_batch sg_batch;
int sg_batch_create(sg_grid *const *grids, int count, sg_batch **out);
int sg_batch_destroy(sg_batch *batch);
/* total grid points and value_offsets[count+1] (may be NULL) */
int sg_batch_values(const sg_batch *batch, long long *total, long long *value_offsets);
/* per-grid point counts; point_offsets[count+1] and coord_offsets[count+1] out (NULL ok) */
int sg_batch_set_points(sg_batch *batch, const long long *counts, long long *point_offsets,
                        long long *coord_offsets);
int sg_batch_hierarchize_f64(sg_batch *batch, double *d_values, void *stream);
int sg_batch_hierarchize_f32(sg_batch *batch, float *d_values, void *stream);
int sg_batch_evaluate_f64(sg_batch *batch, const double *d_alpha, const double *d_coords,
                          double *d_out, void *stream);
int sg_batch_evaluate_f32c(sg_batch *batch, const float *d_alpha, const float *d_coords,
                           float *d_out, void *stream);
/* hierarchize (if with_hierarchize) then evaluate, as one hipGraph launch */
int sg_batch_run_graph_f64(sg_batch *batch, double *d_alpha, const double *d_coords,
                           double *d_out, int with_hierarchize, void *stream);
int sg_batch_run_graph_f32c(sg_batch *batch, float *d_alpha, const float *d_coords,
                            float *d_out, int with_hierarchize, void *stream);

/* ---------------------------------------------------------- serving --------- */
/* Evaluation server with continuous batching (the served form).  Grids and their
 * surpluses (device, in the server's precision) are registered once and stay
 * resident.  Requests (grid, host points [P][d], host output [P]) may be submitted
 * from any thread; a worker thread admits waiting requests in arrival order while
 * their staging bytes (points + outputs) fit the device budget, packs them into one
 * launch per free slot (a work list across grids and requests, one evaluation kernel
 * per dimension count), stages points through pinned memory and copies results back
 * asynchronously.  Requests arriving while launches run join the next launch as soon
 * as a slot and budget free up; up to `slots` launches overlap on the device.  Each
 * request's results are the same alone and in any mix (every point runs the same
 * walk).  Host buffers must stay valid until sg_server_wait returns; every request is
 * waited for exactly once. */
#define SG_SERVE_F64 0
#define SG_SERVE_F32C 1
typedef struct sg_server sg_server;
typedef struct sg_server_config {
    int precision;              /* SG_SERVE_F64 or SG_SERVE_F32C (points, outputs, surpluses) */
    int slots;                  /* launches in flight at once, 1..16 */
    long long max_batch_points; /* points per launch (>= 256); also the largest request */
    int max_requests;           /* requests per launch */
    long long budget_bytes;     /* device bytes admitted requests may hold; <= 0: a quarter
                                   of the free device memory at creation */
} sg_server_config;
int sg_server_create(const sg_server_config *cfg, sg_server **out);
int sg_server_destroy(sg_server *server);   /* finishes the requests in flight */
int sg_server_add_grid(sg_server *server, const sg_grid *grid, const void *d_alpha,
                       int *grid_id);
int sg_server_submit(sg_server *server, int grid_id, const void *points, long long num_points,
                     void *out, long long *request_id);
int sg_server_poll(sg_server *server, long long request_id, int *done);
int sg_server_wait(sg_server *server, long long request_id, double *latency_seconds);
/* Fault injection for the server's error path (a test hook; it has no counterpart in
 * the reference).  The next `launches` launch groups that hold requests of grid_id report
 * a launch failure instead of running: those requests finish with SG_E_HIP, every other
 * request of the same launch completes normally.  launches = 0 clears it. */
int sg_server_inject_launch_failure(sg_server *server, int grid_id, int launches);
int sg_server_stats(sg_server *server, long long *launches, long long *points_done,
                    long long *waiting, long long *budget_used);

/* ----------------------------------------------------------- inputs --------- */

/* Sample a test function at every grid point on the device: d_values[o] = f(x(o)) with
 * x_k = (2 i_k + 1) / 2^(l_k + 1) and (l, i) = idx2gp(o), compact order.  Replaces the
 * host sampling loop of init() (asg.c:143-152) and init_sg1d (asg.cu:105-117).
 *   SG_FN_REFERENCE     sum_k x_k^2 (1 - x_k)^2, the reference's fct (asg.c:87-94) in the
 *                       same operation order: sg_sample_f32 is bit-identical to init();
 *                       params: none
 *   SG_FN_PARABOLA      prod_k 4 x_k (1 - x_k); params: none
 *   SG_FN_GAUSS         exp(-sum_k (x_k - c_k)^2); params: none (c = 1/2) or c[d]
 *   SG_FN_GAUSSIAN_SUM  sum_g exp(-sum_k (x_k - c_gk)^2 / (2 w_gk^2)), g < G <= 8;
 *                       params: [G, c[G][d], w[G][d]]
 *   SG_FN_SIN_PRODUCT   a * prod_k sin(pi x_k); params: none (a = 1) or [a]
 * All but SG_FN_REFERENCE are computed in FP64 and rounded to the output type; exp and
 * sin are the device library's, within a few ulp of the host's. */
#define SG_FN_REFERENCE 0
#define SG_FN_PARABOLA 1
#define SG_FN_GAUSS 2
#define SG_FN_GAUSSIAN_SUM 3
#define SG_FN_SIN_PRODUCT 4
int sg_sample_f64(const sg_grid *grid, int function, const double *params, long long num_params,
                  double *d_values, void *stream);
int sg_sample_f32(const sg_grid *grid, int function, const double *params, long long num_params,
                  float *d_values, void *stream);

/* count values k * 2^-24, k = top 24 bits of splitmix64(seed * 2^32 + j), j = 0..count-1:
 * seeded uniform query coordinates exact in FP32 and FP64 (replaces the time-seeded
 * init_coord_mat, asg.cu:152-165). */
int sg_uniform_points_f64(long long count, unsigned long long seed, double *d_out,
                          void *stream);
int sg_uniform_points_f32(long long count, unsigned long long seed, float *d_out,
                          void *stream);
/* Query points held in another layout -> row-major [P][d], on the device (d_out must
 * not overlap d_in; it holds P * num_dims values):
 *   SG_POINTS_ROW_MAJOR     [P][d] (a copy; asg.c:1259)
 *   SG_POINTS_COLUMN_MAJOR  [d][P]
 *   SG_POINTS_TILED32       [ceil(P/32)][d][32], the CUDA reference's transposed layout
 *                           (init_coord_mat_trans, asg.cu:167-180; its tail tile padded)
 * Replaces the host transpose of the CUDA driver (asg.cu:1679). */
#define SG_POINTS_ROW_MAJOR 0
#define SG_POINTS_COLUMN_MAJOR 1
#define SG_POINTS_TILED32 2
int sg_points_to_row_major_f64(const double *d_in, int layout, int num_dims, long long num_points,
                               double *d_out, void *stream);
int sg_points_to_row_major_f32(const float *d_in, int layout, int num_dims, long long num_points,
                               float *d_out, void *stream);
/* FP64 -> FP32 rounding copy (surpluses for the FP32 evaluate). */
int sg_convert_f64_to_f32(const double *d_in, float *d_out, long long count, void *stream);

/* ------------------------------------------------------------ files --------- */
/* A compact-grid file: a 64-byte header (magic "SGRIDv01", d, n, kind, element size,
 * N, S, FNV-1a 64 checksum of the rest), the level set (restricted-regular: int32
 * limits[d]; explicit: uint8 levels[S][d] in compact order) and N values in compact
 * order, little-endian.  It lets a server load resident grids without rebuilding them.
 * The reference rebuilds its grid every run (asg.c:1767) and has no file format.
 * Values are host memory; elem_bytes is 8, 4, or 0 (a shape-only file). */
int sg_file_write(const char *path, const sg_grid *grid, const void *values, int elem_bytes);
int sg_file_info(const char *path, int *num_dims, int *num_levels, long long *num_grid_points,
                 long long *num_subspaces, int *elem_bytes);
/* Rebuilds the grid (restricted-regular or explicit) and, if values is not NULL, copies
 * the values into it (values_bytes >= N * elem_bytes); the checksum is verified when
 * values are read. */
int sg_file_read(const char *path, sg_grid **grid, void *values, long long values_bytes);

/* ------------------------------------------------------ measurement --------- */

/* The achievable-gather pattern of this grid's shape, the gather roofline the evaluate
 * kernels are measured against.  Per point, one 8-byte (_f64) or 4-byte (_f32) gather at
 * a random position (xorshift32 stream seeded by seed and the point index) inside every
 * subspace block, blocks in compact (level-sum) order, 8 gathers in flight per lane, no
 * hat products and no walk.  d_out[p] = sum of point p's gathered values.  No
 * counterpart in the reference (it has no roofline measurement). */
int sg_gather_probe_f64(const sg_grid *grid, const double *d_alpha, long long num_points,
                        unsigned long long seed, double *d_out, void *stream);
int sg_gather_probe_f32(const sg_grid *grid, const float *d_alpha, long long num_points,
                        unsigned long long seed, float *d_out, void *stream);

/* The staged-pattern probe of the staged walk the grid runs in that precision (its tuned
 * variant; SG_E_UNSUPPORTED if that walk is not a staged one): the walk's stage plan (the
 * same copies into LDS and stage barriers) and, per point, one read per subspace at a
 * random cell of its block (from LDS where the walk stages it, else from global memory)
 * accumulated with one FMA; no hat products and no walk.  Its time is a floor for the
 * staged walk on that grid (the walk does this work and more): the evaluate kernels'
 * roofline.  d_out[p] = point p's sum.  No counterpart in the reference. */
int sg_stage_probe_f64(const sg_grid *grid, const double *d_alpha, long long num_points,
                       unsigned long long seed, double *d_out, void *stream);
int sg_stage_probe_f32(const sg_grid *grid, const float *d_alpha, long long num_points,
                       unsigned long long seed, float *d_out, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* SGRID_H */
