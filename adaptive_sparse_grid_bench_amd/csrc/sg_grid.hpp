// Host-side description of one compact sparse grid: shape tables, the subspace list in
// compact order, the offset <-> (levels, indices) bijection, and the launch tables the
// HIP kernels walk.
//
// Compact order (adaptive_sparse_grid_bench.c, "asg.c", 187-243): subspaces sorted by
// level sum s ascending, then by rank within s (asg.c:266-279: l_{d-1} most significant,
// ascending); a subspace l holds 2^s points stored row-major over its per-dimension
// indices with dimension 0 most significant.  Levels are 0-based; point (l, i) sits at
// x_k = (2 i_k + 1) / 2^(l_k + 1) (asg.c:150).
//
// Two kinds of level sets are supported:
//   * restricted-regular: { l : |l|_1 <= n-1, l_k <= limits[k] } (asg.c:82-83, 108-129);
//     limits = n-1 everywhere is the regular grid.  Ranks come from the dp matrix.
//   * explicit: any downward-closed set of level vectors given by the caller (the
//     refined-subspace form for adaptive grids).  Ranks are positions in the sorted list,
//     sorted by the same key, so a restricted-regular set given explicitly lands in the
//     identical order.
#pragma once

#include <cstdint>
#include <string>
#include <unordered_map>
#include <vector>

namespace sg {

constexpr int kMaxDims = 20;        // asg.cu:17 (MAX_NUM_DIMS)
constexpr int kMaxLevel = 15;       // per-dimension level cap of the kernels' tables
constexpr int kMaxLevelSum = 26;    // keeps 2^s * sizeof(double) inside 32-bit offsets
constexpr int kExplicitTabDims = 3; // register-tabled dims of the explicit-set walk

#if defined(__HIPCC__)
#define SG_HOST_DEVICE __host__ __device__
#else
#define SG_HOST_DEVICE
#endif

// Staged evaluation: the workgroup copies surplus blocks into LDS and each lane reads its
// coefficient there.  Two stage buffers of kStageBytes per workgroup; blocks are copied
// in segments of at most kStageSegBytes (one wave instruction of 16 bytes per lane).
// The staged walk tables kStageJ dimensions (its batch size is the kernel's choice).
// What is staged: the static blocks of the walk (all leaves below a depth-J node whose
// remaining level budget R fits the tabled dimensions) small enough to share a stage
// with others, stage_whole(); other leaves are gathered from global memory.
constexpr int kStageBytes = 26624;   // 3 workgroups x 2 stages fill 156 KiB of a CU's 160
constexpr int kStageSegBytes = 1024;
// a copy segment's second word: LDS byte offset (below 2^24) | lanes << kSegLanesShift
constexpr int kSegLanesShift = 24;
constexpr int32_t kSegLdsMask = (1 << kSegLanesShift) - 1;
constexpr int kStageJ = 3;

SG_HOST_DEVICE constexpr int64_t stage_binom(int a, int b) {
    int64_t r = 1;
    for (int i = 1; i <= b; ++i) r = r * (a - b + i) / i;
    return r;
}
// Bound on the LDS bytes of the static block of budget R over J tabled dimensions below
// a node whose outer levels sum to hsum: sum over its C(R+J, J) leaves of
// 2^(hsum + |c|) elements, plus 16 bytes of alignment padding per leaf.
SG_HOST_DEVICE constexpr int64_t stage_block_bytes(int J, int R, int hsum, int elem) {
    int64_t units = 0;                 // sum_c 2^|c| = sum_s C(s+J-1, J-1) 2^s
    for (int s = 0; s <= R; ++s) units += stage_binom(s + J - 1, J - 1) << s;
    return ((int64_t(elem) << hsum) * units) + 16 * stage_binom(R + J, J);
}
SG_HOST_DEVICE constexpr bool stage_whole(int J, int R, int hsum, int elem) {
    return hsum <= 24 && stage_block_bytes(J, R, hsum, elem) <= kStageBytes;
}
// stage_whole grows false with hsum: the largest hsum for which it holds (-1: none)
SG_HOST_DEVICE constexpr int stage_hmax(int J, int R, int elem) {
    int h = -1;
    for (int x = 0; x <= 24; ++x)
        if (stage_whole(J, R, x, elem)) h = x;
    return h;
}

// Block-staged evaluation (the walk templated on n - 1 of a regular grid).  A
// static block of budget R sits below outer levels summing to H = n - 1 - R, so on the
// device its size and the LDS offset of each of its leaves are compile-time constants.
// The blocks are laid out in LDS in walk order, one after another; inside a block the
// leaves of level sum s (relative to the block) form one run, the runs s = 0..R follow
// each other, each padded to 16 bytes, and inside a run the leaves keep their HBM order
// (l_{J-1}, ..., l_1 ascending, the compact order of one level sum).  Every run is one
// contiguous range in HBM as well.  A workgroup of blk_threads(d, n - 1, elem) threads
// double-buffers stages of blk_stage_bytes(d, n - 1, elem); a stage takes the next
// blocks in walk order while they fit.
// the block-staged kernels are built for n - 1 in kBlkMinNM1..kBlkMaxNM1 (the tabled
// dimensions hold levels 0..7) and for the dimensions where they beat the staged walk
// (2^22-2^24 points, n = 5..8: FP64 from d = 12 on, up to 1.5x at d = 16..20 in the
// depth-first order, 2.5x in the sparse one; FP32 from d = 8 on, up to 12 % at d = 10..12
// and 13-22 % at d = 17..20 in the depth-first order)
constexpr int kBlkMinNM1 = 2;
constexpr int kBlkMaxNM1 = 7;
#ifndef SG_BLK32_DMAX
#define SG_BLK32_DMAX 20
#endif
SG_HOST_DEVICE constexpr bool blk_built(int d, int elem) {
    return elem == 8 ? (d >= 9 && d <= 20) : (d >= 7 && d <= SG_BLK32_DMAX);
}

// bytes of run s (level sum s inside the block) of a static block over J tabled
// dimensions below outer level sum H: C(s+J-1, J-1) subspaces of 2^(H+s) elements
SG_HOST_DEVICE constexpr int64_t blk_run_bytes(int J, int s, int H, int elem) {
    return (stage_binom(s + J - 1, J - 1) * (int64_t(elem) << (H + s)) + 15) & ~int64_t(15);
}
SG_HOST_DEVICE constexpr int64_t blk_run_offset(int J, int s, int H, int elem) {
    int64_t o = 0;
    for (int r = 0; r < s; ++r) o += blk_run_bytes(J, r, H, elem);
    return o;
}
SG_HOST_DEVICE constexpr int64_t blk_bytes(int J, int R, int H, int elem) {
    return blk_run_offset(J, R + 1, H, elem);
}
// position of the leaf with levels lv[J-1..0] (sum s) inside its run: the number of
// level vectors of sum s that precede it in the order (l_{J-1}, ..., l_1) ascending
SG_HOST_DEVICE constexpr int64_t blk_run_pos(int J, const int* lv) {
    int s = 0;
    for (int j = 0; j < J; ++j) s += lv[j];
    int64_t pos = 0;
    int left = s;                       // sum still to place in dims j..0
    for (int j = J - 1; j >= 1; --j) {
        for (int a = 0; a < lv[j]; ++a) pos += stage_binom(left - a + j - 1, j - 1);
        left -= lv[j];
    }
    return pos;
}
SG_HOST_DEVICE constexpr int64_t blk_leaf_offset(int J, const int* lv, int H, int elem) {
    int s = 0;
    for (int j = 0; j < J; ++j) s += lv[j];
    return blk_run_offset(J, s, H, elem) + blk_run_pos(J, lv) * (int64_t(elem) << (H + s));
}

// The block of budget R below outer level sum H fits a stage whole (1), by slices (2:
// slice a = the leaves with l_{J-1} = a, a block over the J-1 inner tabled dimensions
// below outer level sum H + a), or is gathered from global memory (0).  The kernel and
// the host plan apply the same rule.
SG_HOST_DEVICE constexpr int blk_staging(int J, int R, int H, int elem, int stage_bytes) {
    if (H < 0) return 0;
    if (blk_bytes(J, R, H, elem) <= stage_bytes) return 1;
    if (J < 2) return 0;
    for (int a = 0; a <= R; ++a)
        if (blk_bytes(J - 1, R - a, H + a, elem) > stage_bytes) return 0;
    return 2;
}

// The block-staged walk visits the static blocks in one of two orders.
//   depth first (the other walks' order): dimension d-1 outermost, levels ascending;
//   sparse: from a node whose dimensions >= c are fixed, for k = c-1 down to J and
//     l = 1..rem the child with l_k = l (dimensions k+1..c-1 at level 0), then the node's
//     own static block (dimensions J..c-1 all at level 0).  The walk nests only as deep
//     as the number of nonzero outer levels (at most n - 1) instead of d - J loops, and
//     visits each outer level vector once; at high d with small n that removes most of
//     the depth-first walk's loop iterations and the per-level state it spills.
// Both orders give every leaf the same weight, computed in the same order; only the
// order in which blocks are summed differs.  blk_sparse names the dimensions that take
// the sparse order, per element size (host plan and kernel agree on it).
// Measured (2^22 points, ms, depth-first -> sparse): FP64 d = 12, n = 7: 50.0 -> 43.4;
// d = 14, n = 6: 39.3 -> 30.9; d = 16, n = 5: 23.3 -> 14.7; d = 20, n = 6: 307.9 -> 185.8;
// d = 10, n = 8: 35.8 -> 38.7.  FP32 (Kahan) d = 12: 29.3 -> 35.4; d = 14: 25.8 -> 26.3;
// d = 16: 17.5 -> 12.6; d = 20: 225.4 -> 143.9.
#ifndef SG_BLK_SPARSE64_DMIN
#define SG_BLK_SPARSE64_DMIN 12
#endif
#ifndef SG_BLK_SPARSE32_DMIN
#define SG_BLK_SPARSE32_DMIN 15
#endif
SG_HOST_DEVICE constexpr bool blk_sparse(int d, int elem) {
    return d >= (elem == 8 ? SG_BLK_SPARSE64_DMIN : SG_BLK_SPARSE32_DMIN);
}

// the two-points-per-lane kernels (variant 5): FP64, the dimensions with the block-staged
// walk in the depth-first order
//   Measured slower than one point per lane everywhere (2^22-2^24 points, ms, 2 / 1 points
//   per lane): FP64 d = 9, n = 8: 21.6 / 18.7; d = 10, n = 8: 171.6 / 143.4 (2^24);
//   d = 11, n = 7: 72.3 / 29.6; the same lean walk in FP32: d = 10: 137.9 / 99.3 (2^24).
//   Two points' lane state needs the 256 VGPRs of 2 waves per SIMD and still spills
//   (300 bytes per lane at d = 10); the halved stage copies (its probe: 90.3 -> 75.5 ms
//   at d = 10) do not pay for that.  The autotuner keeps it as a candidate.
SG_HOST_DEVICE constexpr bool blk2_built(int d, int elem) {
    return elem == 8 && blk_built(d, elem) && !blk_sparse(d, elem);
}

// Workgroup shape of the block-staged walk.  The depth-first walk of n = 8 grids at
// d = 9..11 and n = 7 grids at d = 10, 11 keeps the offset table of tabled dimension 0 only and computes the hats of
// the outermost tabled dimension once per run (packed lane tables), which fits it in 128
// VGPRs: 512-thread workgroups at 4 waves per SIMD, two per CU, each double-buffering
// stages of 40448 bytes, so one copied stage serves 512 points.  d = 10, n = 8, 2^24
// points, FP64: 144.1 ms at 256 threads and 3 waves per SIMD (26 KiB stages), 135.8 ms
// at 256 threads and 4 waves (19.75 KiB stages), 126.9 ms as here (one workgroup of
// 1024 threads per CU with 80896-byte stages: 147.2 ms); FP32-Kahan: 99.4 -> 85.4 ms.
// n = 7 grids spill 44-64 bytes per lane at 128 VGPRs and still gain at d = 10 and 11
// (2^22 points, FP64: 16.4 -> 15.5 ms, 29.7 -> 27.6 ms; FP32-Kahan: 10.4 -> 9.7 ms,
// 17.6 -> 16.6 ms), not at d = 9 (FP64 8.24 ms against the staged walk's 7.48); at
// n <= 6 it is even (d = 10, n = 6 FP64: 6.80 / 6.83 ms).  Every other block-staged walk: 256 threads, 3 waves per SIMD,
// 26 KiB stages.
constexpr int kBlkThreads = 256, kBlkStageBytes = 26624, kBlkWaves = 3;
constexpr int kBlkQuadThreads = 512, kBlkQuadStageBytes = 40448, kBlkQuadWaves = 4;
SG_HOST_DEVICE constexpr bool blk_quad(int d, int nm1, int elem) {
    return d >= 9 && d <= 11 && !blk_sparse(d, elem) && nm1 <= kBlkMaxNM1 &&
           nm1 >= (d >= 10 ? kBlkMaxNM1 - 1 : kBlkMaxNM1);
}
SG_HOST_DEVICE constexpr int blk_threads(int d, int nm1, int elem) {
    return blk_quad(d, nm1, elem) ? kBlkQuadThreads : kBlkThreads;
}
SG_HOST_DEVICE constexpr int blk_stage_bytes(int d, int nm1, int elem) {
    return blk_quad(d, nm1, elem) ? kBlkQuadStageBytes : kBlkStageBytes;
}
SG_HOST_DEVICE constexpr int blk_waves(int d, int nm1, int elem) {
    return blk_quad(d, nm1, elem) ? kBlkQuadWaves : kBlkWaves;
}

// How the block-staged walk copies one grid's blocks into LDS, for one element size.
struct BlockStagePlan {
    bool ok = false;
    int num_stages = 0;
    std::vector<int32_t> seg_begin;   // [2K+1] as StagePlan::seg_begin
    std::vector<int32_t> segs;        // [nseg + 4][2] as StagePlan::segs
    // subspace base offset of every leaf in this walk's order (the order blk_sparse
    // picks); the kernel gathers the leaves of unstaged blocks through it
    std::vector<int32_t> leaf_base;
    // for the staged-pattern probe (the kernel itself needs neither): LDS byte offset of
    // every staged leaf in walk order (-1: gathered), first leaf of every stage ([K] = S)
    std::vector<int32_t> leaf_lds;
    std::vector<int32_t> stage_first;
};

struct HierTables {
    // One entry per tile / family / block of the pole-family decomposition of one
    // dimension.  A family is the set of subspaces that differ only in l_k; its blocks
    // are those subspaces, l_k = 0..m.
    std::vector<int32_t> tiles;     // [T][3]: first family, family count, value count
    std::vector<int32_t> fams;      // [F][3]: first value in the tile, m|c<<8|B<<16, first block
    std::vector<int32_t> blocks;    // base offset of each block
    int max_family_values = 0;
    int max_families_per_tile = 0;
};

// Plane families of the dimension pair (k, k+1): the subspaces that share every level
// but l_k and l_{k+1}.  A family is stored as its dimension-(k+1) poles for l_k = a =
// 0..A, one after another ("sub-poles"); sub-pole a holds blocks b = 0..mb(a), block b
// at 2^(c+a) (2^b - 1) within it, c the family's level sum outside k and k+1.  Both
// dimensions' parents of every point lie in its family, so one workgroup sweeps k and
// then k+1 in LDS.
constexpr int kPairMaxValues = 4096;  // largest tile (and family) of the pair sweep
constexpr int kPairTileInts = 8;       // ints per pair tile entry
struct PairTables {
    bool ok = false;                  // every family fits kPairMaxValues
    std::vector<int32_t> tiles;       // [T][8]: first family, family count, value count,
                                      //   first sub-pole, sub-pole count, first block,
                                      //   block count, 0 (a tile's sub-poles and blocks
                                      //   are consecutive)
    std::vector<int32_t> fams;        // [F][3]: first value in the tile, c|Bfix<<8|A<<16,
                                      //          first sub-pole
    std::vector<int32_t> subs;        // [P][2]: first value in the family, first block
    std::vector<int32_t> blocks;      // base offset of each block
    int max_families_per_tile = 0;
    int max_subs_per_tile = 0;
    int max_blocks_per_tile = 0;
    int max_tile_values = 0;
};

// How the staged walk copies one grid's surplus blocks into LDS, for one element size.
// The whole static blocks (stage_whole) are packed in walk order into stages; stage k
// lives in LDS buffer k % 2 and holds whole blocks only, so the kernel checks for a stage
// change once per block.  Offsets are absolute LDS byte offsets (buffer included).
struct StagePlan {
    bool ok = false;                  // the staged walk may run on this grid
    std::vector<int32_t> leaf_lds;    // [S] LDS byte offset of each staged leaf
    std::vector<int32_t> stage_first; // [K+1] first leaf of stage k; [K] = S
    std::vector<int32_t> seg_begin;   // [2K+1] stage k's full segments (64 lanes) run
                                      //   from entry 2k to 2k+1, its partial ones to 2k+2
    std::vector<int32_t> segs;        // [nseg + 4][2]: source byte offset,
                                      //   lds offset | lanes << kSegLanesShift (4 zero
                                      //   entries at the end)
};

struct Grid {
    int d = 0;
    int n = 0;                      // number of level sums: 0..n-1
    bool explicit_set = false;
    int limits[kMaxDims] = {};
    int max_level[kMaxDims] = {};   // largest l_k that occurs
    std::vector<int64_t> dp;        // [d][n] (restricted-regular only)
    std::vector<int64_t> lidx;      // [n+1]: first offset of level sum s; lidx[n] = N
    std::vector<int64_t> cnt;       // [n+1]: subspaces before level sum s; cnt[n] = S
    int64_t N = 0;
    int64_t S = 0;
    std::vector<uint8_t> sub_levels;  // [S][d] in compact order
    std::vector<int32_t> sub_base;    // [S] first offset of each subspace
    std::unordered_map<uint64_t, int64_t> index_of;  // explicit sets: key(l) -> subspace id

    // evaluation walk: subspaces in depth-first order (dims d-1 outermost ... 0 innermost)
    std::vector<int32_t> dfs_base;    // [S] base offset per leaf, in walk order
    std::vector<uint8_t> dfs_bound;   // explicit sets: per walk node, the largest l_k
    std::vector<HierTables> hier;     // [d]
    std::vector<PairTables> hier2;    // [d/2]: dimensions (2p, 2p+1)
    StagePlan stage[2];               // [0]: 8-byte surpluses, [1]: 4-byte
    BlockStagePlan blk[2];            // block-staged walk, [0]: 8-byte, [1]: 4-byte
};

// status codes shared with the C ABI
enum Status : int {
    kOk = 0,
    kInvalid = -1,
    kNoMem = -2,
    kHip = -3,
    kRange = -4,
    kUnsupported = -5,
};

const char* status_name(int st);
void set_error(const std::string& msg);
const std::string& last_error();

int build_regular(Grid& g, int d, int n, const int* limits);
int build_explicit(Grid& g, int d, const uint8_t* levels, int64_t count);
int build_launch_tables(Grid& g, int hier_tile_values);
int explicit_tab_dims(const Grid& g);   // tabled dims the explicit-set node codes assume

int64_t subspace_id(const Grid& g, const int* levels);   // -1 if absent
int64_t l2idx(const Grid& g, const int* levels);        // rank within its level sum
int idx2l(const Grid& g, int64_t rank, int sum, int* levels);
int64_t gp2idx(const Grid& g, const int* levels, const int* indices);  // -1 if absent
int idx2gp(const Grid& g, int64_t offset, int* levels, int* indices);

}  // namespace sg
