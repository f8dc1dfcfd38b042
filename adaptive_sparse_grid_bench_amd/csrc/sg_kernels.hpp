// Launchers of the HIP kernels (sg_kernels.hip).  Plain pointers and sizes; every
// launcher checks its launch shape against the buffers it is given and returns a
// status from sg::Status.
#pragma once

#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cstdint>

#include "sg_grid.hpp"

namespace sg {

namespace eval {
// One grid of a launch: where its surpluses, points and outputs start in the launch's
// buffers, its walk tables and its level-set bounds.  Device-resident.
struct GridDesc {
    int64_t alpha_off;          // element offset of the grid's surpluses in the launch's
                                // surplus buffer, or
    const void* alpha_ptr;      // their address when the launch passes no buffer
    int64_t coord_off;          // element offset of its first point's coordinates
    int64_t out_off;            // element offset of its first output
    const int32_t* leaf_base;   // subspace base offsets in walk order
    const uint8_t* bounds;      // explicit sets: one code per walk node
    int64_t N;                  // grid points (the buffer range of its surpluses)
    int32_t n;                  // level sums 0..n-1
    int32_t limits[kMaxDims];
    // LDS stage tables of the staged walk ([0]: 8-byte surpluses, [1]: 4-byte), see
    // StagePlan
    struct Stages {
        const int32_t* leaf_lds;
        const int32_t* stage_first;
        const int32_t* seg_begin;
        const int32_t* segs;
        int32_t num_stages;
        int32_t pad;
    } stage[2];
    // copy plans of the block-staged walk ([0]: 8-byte surpluses, [1]: 4-byte), see
    // BlockStagePlan
    struct BlkStages {
        const int32_t* seg_begin;
        const int32_t* segs;
        const int32_t* leaf_base;   // subspace base offsets in this walk's leaf order
        int32_t num_stages;
        int32_t pad;
    } blk[2];
};

// A run of at most one workgroup's points of one grid.
struct WorkItem {
    int32_t grid;
    int32_t first;              // first point within the grid
    int32_t count;              // points (<= threads per workgroup)
    int32_t pad;
};

}  // namespace eval

// Device copy of the tables one grid's kernels read.  Owned by the grid handle.
struct DeviceTables {
    int device = -1;
    int32_t* dp = nullptr;          // [d][n] int32 (restricted-regular grids)
    int32_t* lidx = nullptr;        // [n+1]
    int32_t* cnt = nullptr;         // [n+1]
    uint8_t* sub_levels = nullptr;  // [S][d]
    int32_t* sub_base = nullptr;    // [S]
    uint8_t* sub_lsum = nullptr;    // [S] level sum of each subspace (compact order)
    int32_t* dfs_base = nullptr;    // [S]
    uint8_t* dfs_bound = nullptr;   // explicit sets: walk-node bounds
    // tables of the staged-pattern probe per plan ([0] staged, [1] block-staged) and
    // element size: per stage (staged leaves, cell mask), per gathered leaf (base, level sum)
    struct Probe {
        int32_t* info = nullptr;
        int32_t* glob = nullptr;
        int num_stages = 0;
        int num_glob = 0;
    } probe[2][2];
    int32_t* stage_tabs[2] = {};    // StagePlan arrays per element size, one allocation each
    int32_t* blk_tabs[2] = {};      // BlockStagePlan arrays per element size
    void* desc = nullptr;           // the grid's evaluation descriptor (eval::GridDesc)
    eval::GridDesc host_desc{};     // host copy: tables set, launch offsets zero
    struct Dim {
        int32_t* tiles = nullptr;
        int32_t* fams = nullptr;
        int32_t* blocks = nullptr;
        int num_tiles = 0;
        int max_families = 0;
        int tile_values = 0;
    } dims[kMaxDims];
    struct Pair {                   // plane families of dimensions (2p, 2p+1)
        bool ok = false;
        int32_t* tiles = nullptr;
        int32_t* fams = nullptr;
        int32_t* subs = nullptr;
        int32_t* blocks = nullptr;
        int num_tiles = 0;
        int max_families = 0;
        int max_subs = 0;
        int max_blocks = 0;
        int tile_values = 0;
    } pairs[kMaxDims / 2];
    int hier_tile_values = 0;
};

int upload_tables(const Grid& g, DeviceTables& t);
void free_tables(DeviceTables& t);

int launch_idx2gp(const Grid& g, const DeviceTables& t, const int32_t* index, int64_t count,
                  int32_t* levels, int32_t* indices, hipStream_t s);
int launch_gp2idx(const Grid& g, const DeviceTables& t, const int32_t* levels,
                  const int32_t* indices, int64_t count, int32_t* index, hipStream_t s);
int launch_grid_coords_f64(const Grid& g, const DeviceTables& t, double* coords, hipStream_t s);

int launch_hierarchize_f64(const Grid& g, const DeviceTables& t, double* v, hipStream_t s);
// one dimension sweep over a set of pole-family tiles (of one grid or of a batch)
template <class T>
int launch_hierarchize_tiles(T* v, const int32_t* tiles, const int32_t* fams,
                             const int32_t* blocks, int num_tiles, int tile_values,
                             int max_families, hipStream_t s);
int launch_hierarchize_f32(const Grid& g, const DeviceTables& t, float* v, hipStream_t s);
// sweep of the dimension pair (k, k+1) over a set of plane-family tiles (see PairTables)
template <class T>
int launch_hierarchize_pairs(T* v, const int32_t* tiles, const int32_t* fams, const int32_t* subs,
                             const int32_t* blocks, int num_tiles, int tile_values,
                             int max_families, int max_subs, int max_blocks, hipStream_t s);

struct EvalTuning {
    int block = 256;            // threads per workgroup
    int variant = 0;            // walk variant (sg_eval.hpp kVariants)
    int bin_min_points = 4096;  // spatially bin the points of launches at least this big
};

// the walk variant that reads surpluses from LDS stages
constexpr int kStagedVariant = 3;
// the block-staged walk (regular grids, n - 1 in kBlkMinNM1..kBlkMaxNM1)
constexpr int kBlkVariant = 4;
// the block-staged walk with two points per lane (blk2_built: FP64, depth-first order)
constexpr int kBlk2Variant = 5;

// default variants by dimension, n - 1 and precision, from timings at 2^22-2^24 points
// on d = 5..16 and 20 (the autotuner refines them per grid shape); a grid the
// block-staged walk does not cover runs the staged walk.  FP64 below d = 12 takes the
// block-staged walk where it runs at 4 waves per SIMD (blk_quad: n = 8 at d = 9..11, n = 7 at d = 10, 11;
// 2^22 points at d = 9 and 11, 2^24 at d = 10: 16.7 / 127.0 / 58.7 ms against the staged
// walk's 19.4 / 143.9 / 65.9 ms); at 3 waves per SIMD it ties with the staged walk there.
inline int default_variant(int d, int nm1, bool f32) {
    if (f32) return d >= 8 ? kBlkVariant : kStagedVariant;
    return d >= 12 || blk_quad(d, nm1, 8) ? kBlkVariant : kStagedVariant;
}
// the variant a launch runs: the staged walk falls back to variant 0 for a grid whose
// blocks cannot be staged in this element size
inline int effective_variant(const Grid& g, int variant, int elem_bytes) {
    const int e = elem_bytes == 8 ? 0 : 1;
    // the gathering walks 1 and 2 differ from 0 only in their tabled dimensions and batch:
    // explicit sets, and grids whose tabled dimensions reach level 8, run the one
    // table-free gathering walk (sg_eval.hpp launch_variant / launch_d)
    if (variant == 1 || variant == 2) {
        bool tabled = !g.explicit_set;
        for (int j = 0; j < std::min(g.d, variant == 1 ? 2 : 3) && tabled; ++j)
            tabled = g.limits[j] <= 7;
        if (!tabled) return 0;
    }
    if (variant == kBlk2Variant) {
        if (!blk2_built(g.d, elem_bytes)) variant = kBlkVariant;
        else if (g.explicit_set || !g.blk[e].ok || g.n - 1 < kBlkMinNM1 || g.n - 1 > kBlkMaxNM1)
            variant = kStagedVariant;
        else
            return variant;
    }
    if (variant == kBlkVariant &&
        (g.explicit_set || !g.blk[e].ok || g.n - 1 < kBlkMinNM1 || g.n - 1 > kBlkMaxNM1 ||
         !blk_built(g.d, elem_bytes)))
        variant = kStagedVariant;
    if (variant == kStagedVariant && !g.stage[e].ok) return 0;
    return variant;
}
// the variant of a work-list launch (batches, the server); work items hold one point per
// lane, so the two-point walk runs as the block-staged one
inline int work_list_variant(const Grid& g, int variant, int elem_bytes) {
    return effective_variant(g, variant == kBlk2Variant ? kBlkVariant : variant, elem_bytes);
}
// points per work item of a walk in a work list: its workgroup size (512 for the
// block-staged walk at 4 waves per SIMD, blk_quad; 256 for every other walk)
inline int work_item_points(const Grid& g, int variant, int elem_bytes) {
    return variant == kBlkVariant ? blk_threads(g.d, g.n - 1, elem_bytes) : kBlkThreads;
}

// perm: lane order of the points (from sort_points) or null for the given order
int launch_evaluate_f64(const Grid& g, const DeviceTables& t, const double* alpha,
                        const double* coords, int64_t P, double* out, const EvalTuning& tu,
                        const uint32_t* perm, hipStream_t s);
int launch_evaluate_f32c(const Grid& g, const DeviceTables& t, const float* alpha,
                         const float* coords, int64_t P, float* out, const EvalTuning& tu,
                         const uint32_t* perm, hipStream_t s);

// Spatial binning of query points: a key of k interleaved leading coordinate bits
// (bit 1 of dims 0..d-1, then bit 2, ...), counting-sorted into perm[P].  Points that
// share a key share their cell in every subspace whose levels stay within those bits.
// Workspace: 4 (P + 2^k + 1) bytes, from sort_workspace_bytes.
int key_bits_for(int64_t P, int d);
size_t sort_workspace_bytes(int64_t P, int key_bits);
int launch_sort_points_f64(const double* coords, int64_t P, int d, int key_bits, void* ws,
                           uint32_t* perm, hipStream_t s);
int launch_sort_points_f32(const float* coords, int64_t P, int d, int key_bits, void* ws,
                           uint32_t* perm, hipStream_t s);

// Fixed device scratch of the allocation-free evaluate (sg_evaluate_f64_noalloc): the
// current device's address of a file-scope device array holding the binning workspace
// and permutation of up to kNoAllocMaxPoints points (looked up once per device).
constexpr int64_t kNoAllocMaxPoints = int64_t(1) << 21;
int noalloc_scratch(void** ptr, size_t* bytes);

int launch_evaluate_ref_f64(const Grid& g, const DeviceTables& t, const double* alpha,
                            const double* coords, int64_t P, double* out, hipStream_t s);
int launch_evaluate_ref_f32(const Grid& g, const DeviceTables& t, const float* alpha,
                            const float* coords, int64_t P, float* out, hipStream_t s);

int launch_uniform_points_f64(int64_t count, uint64_t seed, double* out, hipStream_t s);
int launch_uniform_points_f32(int64_t count, uint64_t seed, float* out, hipStream_t s);
int launch_f64_to_f32(const double* in, float* out, int64_t count, hipStream_t s);

// Query points in another layout -> row-major [P][d] (the layout the kernels read).
enum PointLayout : int {
    kPointsRowMajor = 0,      // [P][d]
    kPointsColumnMajor = 1,   // [d][P]
    kPointsTiled32 = 2,       // [ceil(P/32)][d][32], asg.cu:167-180
};
int launch_points_to_row_major_f64(const double* in, int layout, int d, int64_t P, double* out,
                                   hipStream_t s);
int launch_points_to_row_major_f32(const float* in, int layout, int d, int64_t P, float* out,
                                   hipStream_t s);

// Test functions sampled at the grid points on the device (values in compact order).
enum SampleFn : int {
    kFnReference = 0,     // sum_k x_k^2 (1 - x_k)^2, asg.c:87-94, same operation order
    kFnParabola = 1,      // prod_k 4 x_k (1 - x_k)
    kFnGauss = 2,         // exp(-sum_k (x_k - c_k)^2)
    kFnGaussianSum = 3,   // sum_g exp(-sum_k (x_k - c_gk)^2 / (2 w_gk^2))
    kFnSinProduct = 4,    // scale * prod_k sin(pi x_k)
};
constexpr int kMaxSampleTerms = 8;
struct SampleParams {
    int fn = 0;
    int terms = 1;                    // Gaussians of kFnGaussianSum
    double scale = 1.0;               // kFnSinProduct factor
    double centre[kMaxSampleTerms][kMaxDims] = {};
    double width[kMaxSampleTerms][kMaxDims] = {};
};
int launch_sample_f64(const Grid& g, const DeviceTables& t, const SampleParams& sp, double* out,
                      hipStream_t s);
int launch_sample_f32(const Grid& g, const DeviceTables& t, const SampleParams& sp, float* out,
                      hipStream_t s);

// The achievable-gather pattern of grid g: per point, one gather at a random position
// inside every subspace block, blocks in compact (level-sum) order, in batches of 8;
// no hat products and no walk.  out[p] = the sum of the values gathered for point p.
int launch_gather_probe_f64(const Grid& g, const DeviceTables& t, const double* alpha,
                            int64_t P, uint32_t seed, double* out, hipStream_t s);
int launch_gather_probe_f32(const Grid& g, const DeviceTables& t, const float* alpha,
                            int64_t P, uint32_t seed, float* out, hipStream_t s);
// The staged-pattern probe of the staged walk grid g runs under `variant` (block-staged or
// staged): the walk's stage plan and one LDS read (global read for unstaged leaves) plus
// one FMA per leaf at a random cell, no hats and no walk.  out[p] = point p's sum.
int launch_stage_probe_f64(const Grid& g, const DeviceTables& t, int variant,
                           const double* alpha, int64_t P, uint32_t seed, double* out,
                           hipStream_t s);
int launch_stage_probe_f32(const Grid& g, const DeviceTables& t, int variant, const float* alpha,
                           int64_t P, uint32_t seed, float* out, hipStream_t s);

}  // namespace sg
