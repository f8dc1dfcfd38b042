// C ABI of libsgrid.so (include/sgrid.h): argument checks, handle ownership and
// dispatch to the host tables (sg_grid.cpp) and the HIP launchers (sg_kernels.hip).
#include <hip/hip_runtime_api.h>

#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "../../include/sgrid.h"
#include "sg_batch.hpp"
#include "sg_grid.hpp"
#include "sg_server.hpp"
#include "sg_kernels.hpp"

// Device tables are uploaded on the first device call, to the device current then, so
// that creating a grid and the host bijection need no GPU.
struct sg_grid {
    sg::Grid g;
    sg::DeviceTables t;
    sg::EvalTuning tune;        // FP64 evaluate
    sg::EvalTuning tune32;      // FP32 evaluate
    int hier_tile_values = 1024;
    std::mutex mu;
    bool uploaded = false;
};

namespace {

int fail(int st, const std::string& msg) {
    sg::set_error(msg);
    return st;
}

hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

// Every device entry point runs on the device that holds the grid's tables; the first
// one uploads them.
int check_device(const sg_grid* hc) {
    sg_grid* h = const_cast<sg_grid*>(hc);
    int cur = -1;
    if (hipGetDevice(&cur) != hipSuccess) return fail(sg::kHip, "hipGetDevice failed");
    std::lock_guard<std::mutex> lock(h->mu);
    if (!h->uploaded) {
        int st = sg::upload_tables(h->g, h->t);
        if (st != sg::kOk) {
            sg::free_tables(h->t);
            return st;
        }
        h->uploaded = true;
    }
    if (cur != h->t.device)
        return fail(sg::kInvalid, "current device differs from the grid's device");
    return sg::kOk;
}

int finish_create(sg_grid* h, sg_grid** out) {
    h->tune.variant = sg::default_variant(h->g.d, h->g.n - 1, false);
    h->tune32.variant = sg::default_variant(h->g.d, h->g.n - 1, true);
    int st = sg::build_launch_tables(h->g, h->hier_tile_values);
    if (st != sg::kOk) {
        delete h;
        return st;
    }
    *out = h;
    return sg::kOk;
}

// Launches of at least tune.bin_min_points points first bin the points spatially (a
// permutation in stream-ordered scratch memory, freed on the same stream), then walk
// them in binned order; every point's arithmetic is the same either way.
int eval_one(const sg_grid* h, const double* a, const double* x, int64_t P, double* o,
             const uint32_t* perm, hipStream_t s) {
    return sg::launch_evaluate_f64(h->g, h->t, a, x, P, o, h->tune, perm, s);
}
int eval_one(const sg_grid* h, const float* a, const float* x, int64_t P, float* o,
             const uint32_t* perm, hipStream_t s) {
    return sg::launch_evaluate_f32c(h->g, h->t, a, x, P, o, h->tune32, perm, s);
}
int sort_one(const double* x, int64_t P, int d, int kb, void* ws, uint32_t* perm, hipStream_t s) {
    return sg::launch_sort_points_f64(x, P, d, kb, ws, perm, s);
}
int sort_one(const float* x, int64_t P, int d, int kb, void* ws, uint32_t* perm, hipStream_t s) {
    return sg::launch_sort_points_f32(x, P, d, kb, ws, perm, s);
}

// ws / ws_bytes: a fixed workspace for the binning (the allocation-free entry point), or
// null for stream-ordered memory allocated and freed per call.
template <class T>
int evaluate_binned(const sg_grid* h, const T* a, const T* x, int64_t P, T* o, void* stream,
                    void* ws = nullptr, size_t ws_cap = 0) {
    hipStream_t s = as_stream(stream);
    // binning pays when the walk per point is long: skip it for small grids, and for the
    // staged walk, where it only removes LDS bank conflicts: at d=10, n=8 and 2^24 points
    // the sort costs 2.5 ms and the binned walk is 0.6 ms faster (FP64, 187.4 ms unbinned)
    const int v = sg::effective_variant(h->g, sizeof(T) == 8 ? h->tune.variant : h->tune32.variant,
                                        int(sizeof(T)));
    if (P < h->tune.bin_min_points || h->tune.bin_min_points <= 0 || h->g.S < 1024 ||
        v == sg::kStagedVariant || v == sg::kBlkVariant || v == sg::kBlk2Variant)
        return eval_one(h, a, x, P, o, nullptr, s);
    const int kb = sg::key_bits_for(P, h->g.d);
    const size_t ws_bytes = sg::sort_workspace_bytes(P, kb);
    if (ws) {
        if (ws_bytes + size_t(P) * 4 > ws_cap)
            return fail(sg::kRange, "evaluate: binning workspace exceeds the fixed scratch");
        uint32_t* perm = reinterpret_cast<uint32_t*>(static_cast<char*>(ws) + ws_bytes);
        int st = sort_one(x, P, h->g.d, kb, ws, perm, s);
        return st == sg::kOk ? eval_one(h, a, x, P, o, perm, s) : st;
    }
    void* mem = nullptr;
    hipError_t e = hipMallocAsync(&mem, ws_bytes + size_t(P) * 4, s);
    if (e != hipSuccess)
        return fail(e == hipErrorOutOfMemory ? sg::kNoMem : sg::kHip,
                    std::string("evaluate: hipMallocAsync: ") + hipGetErrorString(e));
    uint32_t* perm = reinterpret_cast<uint32_t*>(static_cast<char*>(mem) + ws_bytes);
    int st = sort_one(x, P, h->g.d, kb, mem, perm, s);
    if (st == sg::kOk) st = eval_one(h, a, x, P, o, perm, s);
    e = hipFreeAsync(mem, s);
    if (st == sg::kOk && e != hipSuccess)
        return fail(sg::kHip, std::string("evaluate: hipFreeAsync: ") + hipGetErrorString(e));
    return st;
}

}  // namespace

struct sg_batch {
    sg::Batch b;
    int device = -1;
    bool points_set = false;
};

namespace {
int batch_device(const sg_batch* b) {
    int cur = -1;
    if (hipGetDevice(&cur) != hipSuccess) return fail(sg::kHip, "hipGetDevice failed");
    if (cur != b->device) return fail(sg::kInvalid, "current device differs from the batch's");
    if (!b->points_set) return fail(sg::kInvalid, "sg_batch_set_points has not been called");
    return sg::kOk;
}
}  // namespace

extern "C" {

const char* sg_last_error(void) { return sg::last_error().c_str(); }
const char* sg_status_string(int status) { return sg::status_name(status); }
int sg_version(void) { return 210; }

int sg_grid_create(int num_dims, int num_levels, const int* limits, sg_grid** out) {
    if (!out) return fail(sg::kInvalid, "sg_grid_create: out is NULL");
    *out = nullptr;
    sg_grid* h = new (std::nothrow) sg_grid();
    if (!h) return fail(sg::kNoMem, "sg_grid_create: host allocation failed");
    int st = sg::build_regular(h->g, num_dims, num_levels, limits);
    if (st != sg::kOk) {
        delete h;
        return st;
    }
    return finish_create(h, out);
}

int sg_grid_create_explicit(int num_dims, const unsigned char* levels, long long count,
                            sg_grid** out) {
    if (!out || !levels) return fail(sg::kInvalid, "sg_grid_create_explicit: NULL argument");
    *out = nullptr;
    sg_grid* h = new (std::nothrow) sg_grid();
    if (!h) return fail(sg::kNoMem, "sg_grid_create_explicit: host allocation failed");
    int st = sg::build_explicit(h->g, num_dims, levels, count);
    if (st != sg::kOk) {
        delete h;
        return st;
    }
    return finish_create(h, out);
}

int sg_grid_destroy(sg_grid* grid) {
    if (!grid) return sg::kOk;
    if (grid->uploaded) sg::free_tables(grid->t);
    delete grid;
    return sg::kOk;
}

int sg_grid_info(const sg_grid* grid, int* num_dims, int* num_levels,
                 long long* num_grid_points, long long* num_subspaces) {
    if (!grid) return fail(sg::kInvalid, "sg_grid_info: NULL grid");
    if (num_dims) *num_dims = grid->g.d;
    if (num_levels) *num_levels = grid->g.n;
    if (num_grid_points) *num_grid_points = grid->g.N;
    if (num_subspaces) *num_subspaces = grid->g.S;
    return sg::kOk;
}

int sg_grid_tables(const sg_grid* grid, long long* dp, long long* lidx) {
    if (!grid) return fail(sg::kInvalid, "sg_grid_tables: NULL grid");
    if (dp) {
        if (grid->g.explicit_set)
            return fail(sg::kUnsupported, "sg_grid_tables: explicit level sets have no dp matrix");
        for (size_t i = 0; i < grid->g.dp.size(); ++i) dp[i] = grid->g.dp[i];
    }
    if (lidx)
        for (size_t i = 0; i < grid->g.lidx.size(); ++i) lidx[i] = grid->g.lidx[i];
    return sg::kOk;
}

int sg_grid_subspaces(const sg_grid* grid, unsigned char* levels, long long* base) {
    if (!grid) return fail(sg::kInvalid, "sg_grid_subspaces: NULL grid");
    if (levels) std::memcpy(levels, grid->g.sub_levels.data(), grid->g.sub_levels.size());
    if (base)
        for (int64_t i = 0; i < grid->g.S; ++i) base[i] = grid->g.sub_base[i];
    return sg::kOk;
}

int sg_grid_get_tuning(const sg_grid* grid, sg_tuning* tu) {
    if (!grid || !tu) return fail(sg::kInvalid, "sg_grid_get_tuning: NULL argument");
    tu->eval_block = grid->tune.block;
    tu->eval_variant_f64 = grid->tune.variant;
    tu->eval_variant_f32 = grid->tune32.variant;
    tu->bin_min_points = grid->tune.bin_min_points;
    tu->hier_tile_values = grid->hier_tile_values;
    return sg::kOk;
}

int sg_grid_set_tuning(sg_grid* grid, const sg_tuning* tu) {
    if (!grid || !tu) return fail(sg::kInvalid, "sg_grid_set_tuning: NULL argument");
    if (tu->eval_block != 64 && tu->eval_block != 128 && tu->eval_block != 256)
        return fail(sg::kInvalid, "sg_grid_set_tuning: eval_block must be 64, 128 or 256");
    if (tu->eval_variant_f64 < 0 || tu->eval_variant_f64 >= SG_EVAL_VARIANTS ||
        tu->eval_variant_f32 < 0 || tu->eval_variant_f32 >= SG_EVAL_VARIANTS)
        return fail(sg::kInvalid, "sg_grid_set_tuning: unknown evaluate variant");
    if (tu->hier_tile_values < 64 || tu->hier_tile_values > (1 << 24))
        return fail(sg::kInvalid, "sg_grid_set_tuning: hier_tile_values out of range");
    std::lock_guard<std::mutex> lock(grid->mu);
    grid->tune.block = grid->tune32.block = tu->eval_block;
    grid->tune.variant = tu->eval_variant_f64;
    grid->tune32.variant = tu->eval_variant_f32;
    grid->tune.bin_min_points = grid->tune32.bin_min_points = tu->bin_min_points;
    if (tu->hier_tile_values != grid->hier_tile_values) {
        grid->hier_tile_values = tu->hier_tile_values;
        int st = sg::build_launch_tables(grid->g, tu->hier_tile_values);
        if (st != sg::kOk) return st;
        if (grid->uploaded) {            // re-upload on the next device call
            sg::free_tables(grid->t);
            grid->uploaded = false;
        }
    }
    return sg::kOk;
}

int sg_grid_walk_variant(const sg_grid* grid, int elem_bytes, int variant, int* walk) {
    if (!grid || !walk) return fail(sg::kInvalid, "sg_grid_walk_variant: NULL argument");
    if (elem_bytes != 8 && elem_bytes != 4)
        return fail(sg::kInvalid, "sg_grid_walk_variant: elem_bytes must be 8 or 4");
    if (variant < 0 || variant >= SG_EVAL_VARIANTS)
        return fail(sg::kInvalid, "sg_grid_walk_variant: unknown evaluate variant");
    *walk = sg::effective_variant(grid->g, variant, elem_bytes);
    return sg::kOk;
}

int sg_gp2idx(const sg_grid* grid, const int* levels, const int* indices, long long* index) {
    if (!grid || !levels || !indices || !index) return fail(sg::kInvalid, "sg_gp2idx: NULL argument");
    int64_t o = sg::gp2idx(grid->g, levels, indices);
    if (o < 0) return fail(sg::kRange, "sg_gp2idx: point is not in the grid");
    *index = o;
    return sg::kOk;
}

int sg_idx2gp(const sg_grid* grid, long long index, int* levels, int* indices) {
    if (!grid || !levels || !indices) return fail(sg::kInvalid, "sg_idx2gp: NULL argument");
    int st = sg::idx2gp(grid->g, index, levels, indices);
    if (st != sg::kOk) return fail(st, "sg_idx2gp: offset out of range");
    return sg::kOk;
}

int sg_idx2l(const sg_grid* grid, long long rank, int level_sum, int* levels) {
    if (!grid || !levels) return fail(sg::kInvalid, "sg_idx2l: NULL argument");
    int st = sg::idx2l(grid->g, rank, level_sum, levels);
    if (st != sg::kOk) return fail(st, "sg_idx2l: rank or level sum out of range");
    return sg::kOk;
}

int sg_l2idx(const sg_grid* grid, const int* levels, long long* rank) {
    if (!grid || !levels || !rank) return fail(sg::kInvalid, "sg_l2idx: NULL argument");
    if (sg::subspace_id(grid->g, levels) < 0)
        return fail(sg::kRange, "sg_l2idx: level vector is not in the grid");
    *rank = sg::l2idx(grid->g, levels);
    return sg::kOk;
}

int sg_idx2gp_dev(const sg_grid* grid, const int* d_index, long long count, int* d_levels,
                  int* d_indices, void* stream) {
    if (!grid || count < 0 || (count > 0 && (!d_levels || !d_indices)))
        return fail(sg::kInvalid, "sg_idx2gp_dev: bad argument");
    int st = check_device(grid);
    if (st != sg::kOk) return st;
    return sg::launch_idx2gp(grid->g, grid->t, d_index, count, d_levels, d_indices,
                             as_stream(stream));
}

int sg_gp2idx_dev(const sg_grid* grid, const int* d_levels, const int* d_indices,
                  long long count, int* d_index, void* stream) {
    if (!grid || count < 0 || (count > 0 && (!d_levels || !d_indices || !d_index)))
        return fail(sg::kInvalid, "sg_gp2idx_dev: bad argument");
    int st = check_device(grid);
    if (st != sg::kOk) return st;
    return sg::launch_gp2idx(grid->g, grid->t, d_levels, d_indices, count, d_index,
                             as_stream(stream));
}

int sg_grid_coords_f64(const sg_grid* grid, double* d_coords, void* stream) {
    if (!grid || !d_coords) return fail(sg::kInvalid, "sg_grid_coords_f64: NULL argument");
    int st = check_device(grid);
    if (st != sg::kOk) return st;
    return sg::launch_grid_coords_f64(grid->g, grid->t, d_coords, as_stream(stream));
}

int sg_hierarchize_f64(const sg_grid* grid, double* d_values, void* stream) {
    if (!grid || !d_values) return fail(sg::kInvalid, "sg_hierarchize_f64: NULL argument");
    int st = check_device(grid);
    if (st != sg::kOk) return st;
    return sg::launch_hierarchize_f64(grid->g, grid->t, d_values, as_stream(stream));
}

int sg_hierarchize_f32(const sg_grid* grid, float* d_values, void* stream) {
    if (!grid || !d_values) return fail(sg::kInvalid, "sg_hierarchize_f32: NULL argument");
    int st = check_device(grid);
    if (st != sg::kOk) return st;
    return sg::launch_hierarchize_f32(grid->g, grid->t, d_values, as_stream(stream));
}

#define SG_EVAL_CHECK(name)                                                            \
    if (!grid || num_evals < 0 || (num_evals > 0 && (!d_alpha || !d_coords || !d_out))) \
        return fail(sg::kInvalid, name ": bad argument");                              \
    {                                                                                  \
        int st_ = check_device(grid);                                                  \
        if (st_ != sg::kOk) return st_;                                                \
    }

int sg_evaluate_f64(const sg_grid* grid, const double* d_alpha, const double* d_coords,
                    long long num_evals, double* d_out, void* stream) {
    SG_EVAL_CHECK("sg_evaluate_f64")
    return evaluate_binned<double>(grid, d_alpha, d_coords, num_evals, d_out, stream);
}

int sg_evaluate_f64_noalloc(const sg_grid* grid, const double* d_alpha, const double* d_coords,
                            long long num_evals, double* d_out, void* stream) {
    if (!grid || num_evals < 0 || (num_evals > 0 && (!d_alpha || !d_coords || !d_out)))
        return fail(sg::kInvalid, "sg_evaluate_f64_noalloc: bad argument");
    if (num_evals > sg::kNoAllocMaxPoints)
        return fail(sg::kRange, "sg_evaluate_f64_noalloc: " + std::to_string(num_evals) +
                                    " points exceed SG_NOALLOC_MAX_POINTS (" +
                                    std::to_string(sg::kNoAllocMaxPoints) + ")");
    sg_grid* h = const_cast<sg_grid*>(grid);
    int cur = -1;
    if (hipGetDevice(&cur) != hipSuccess) return fail(sg::kHip, "hipGetDevice failed");
    {
        // uploading the grid's tables allocates: here they must already be on the device
        std::lock_guard<std::mutex> lock(h->mu);
        if (!h->uploaded)
            return fail(sg::kInvalid,
                        "sg_evaluate_f64_noalloc: the grid's device tables are not uploaded; "
                        "make one other device call on the grid first (e.g. "
                        "sg_hierarchize_f64 or sg_evaluate_f64)");
        if (cur != h->t.device)
            return fail(sg::kInvalid, "current device differs from the grid's device");
    }
    void* ws = nullptr;
    size_t ws_cap = 0;
    int st = sg::noalloc_scratch(&ws, &ws_cap);
    if (st != sg::kOk) return st;
    return evaluate_binned<double>(grid, d_alpha, d_coords, num_evals, d_out, stream, ws, ws_cap);
}

int sg_evaluate_f32c(const sg_grid* grid, const float* d_alpha, const float* d_coords,
                     long long num_evals, float* d_out, void* stream) {
    SG_EVAL_CHECK("sg_evaluate_f32c")
    return evaluate_binned<float>(grid, d_alpha, d_coords, num_evals, d_out, stream);
}

int sg_evaluate_ref_f64(const sg_grid* grid, const double* d_alpha, const double* d_coords,
                        long long num_evals, double* d_out, void* stream) {
    SG_EVAL_CHECK("sg_evaluate_ref_f64")
    return sg::launch_evaluate_ref_f64(grid->g, grid->t, d_alpha, d_coords, num_evals, d_out,
                                       as_stream(stream));
}

int sg_evaluate_ref_f32(const sg_grid* grid, const float* d_alpha, const float* d_coords,
                        long long num_evals, float* d_out, void* stream) {
    SG_EVAL_CHECK("sg_evaluate_ref_f32")
    return sg::launch_evaluate_ref_f32(grid->g, grid->t, d_alpha, d_coords, num_evals, d_out,
                                       as_stream(stream));
}

int sg_uniform_points_f64(long long count, unsigned long long seed, double* d_out,
                          void* stream) {
    if (count < 0 || (count > 0 && !d_out)) return fail(sg::kInvalid, "sg_uniform_points_f64: bad argument");
    return sg::launch_uniform_points_f64(count, seed, d_out, as_stream(stream));
}

int sg_uniform_points_f32(long long count, unsigned long long seed, float* d_out,
                          void* stream) {
    if (count < 0 || (count > 0 && !d_out)) return fail(sg::kInvalid, "sg_uniform_points_f32: bad argument");
    return sg::launch_uniform_points_f32(count, seed, d_out, as_stream(stream));
}

namespace {
int sample_params(const sg_grid* grid, int function, const double* params, long long num_params,
                  sg::SampleParams& sp) {
    const int d = grid->g.d;
    sp.fn = function;
    if (num_params < 0 || (num_params > 0 && !params))
        return fail(sg::kInvalid, "sg_sample: bad parameter array");
    switch (function) {
        case SG_FN_REFERENCE:
        case SG_FN_PARABOLA:
            if (num_params != 0) return fail(sg::kInvalid, "sg_sample: function takes no parameters");
            return sg::kOk;
        case SG_FN_GAUSS:
            if (num_params != 0 && num_params != d)
                return fail(sg::kInvalid, "sg_sample: Gaussian takes 0 or d centres");
            for (int k = 0; k < d; ++k) sp.centre[0][k] = num_params ? params[k] : 0.5;
            return sg::kOk;
        case SG_FN_GAUSSIAN_SUM: {
            if (num_params < 1) return fail(sg::kInvalid, "sg_sample: Gaussian sum needs its count");
            const double tc = params[0];
            const int terms = int(tc);
            if (double(terms) != tc || terms < 1 || terms > sg::kMaxSampleTerms ||
                num_params != 1 + 2LL * terms * d)
                return fail(sg::kInvalid, "sg_sample: Gaussian sum takes [count, centres, widths]");
            sp.terms = terms;
            for (int t = 0; t < terms; ++t)
                for (int k = 0; k < d; ++k) {
                    sp.centre[t][k] = params[1 + t * d + k];
                    sp.width[t][k] = params[1 + terms * d + t * d + k];
                    if (!(sp.width[t][k] > 0.0))
                        return fail(sg::kInvalid, "sg_sample: Gaussian widths must be positive");
                }
            return sg::kOk;
        }
        case SG_FN_SIN_PRODUCT:
            if (num_params > 1) return fail(sg::kInvalid, "sg_sample: sine product takes 0 or 1 factor");
            sp.scale = num_params ? params[0] : 1.0;
            return sg::kOk;
        default:
            return fail(sg::kInvalid, "sg_sample: unknown function");
    }
}
}  // namespace

int sg_sample_f64(const sg_grid* grid, int function, const double* params, long long num_params,
                  double* d_values, void* stream) {
    if (!grid || !d_values) return fail(sg::kInvalid, "sg_sample_f64: NULL argument");
    sg::SampleParams sp;
    int st = sample_params(grid, function, params, num_params, sp);
    if (st != sg::kOk) return st;
    if ((st = check_device(grid)) != sg::kOk) return st;
    return sg::launch_sample_f64(grid->g, grid->t, sp, d_values, as_stream(stream));
}

int sg_sample_f32(const sg_grid* grid, int function, const double* params, long long num_params,
                  float* d_values, void* stream) {
    if (!grid || !d_values) return fail(sg::kInvalid, "sg_sample_f32: NULL argument");
    sg::SampleParams sp;
    int st = sample_params(grid, function, params, num_params, sp);
    if (st != sg::kOk) return st;
    if ((st = check_device(grid)) != sg::kOk) return st;
    return sg::launch_sample_f32(grid->g, grid->t, sp, d_values, as_stream(stream));
}

// ------------------------------------------------------------- grid files --
// Layout (little-endian): 64-byte header, then the level set (restricted-regular:
// int32 limits[d]; explicit: uint8 levels[S][d] in compact order), then N values.
namespace {

constexpr char kFileMagic[8] = {'S', 'G', 'R', 'I', 'D', 'v', '0', '1'};

struct FileHeader {
    char magic[8];
    uint32_t header_bytes;
    uint32_t d, n;
    uint32_t kind;              // 0 restricted-regular, 1 explicit
    uint32_t elem_bytes;        // 8, 4, or 0 (no values)
    uint32_t reserved;
    uint64_t N, S;
    uint64_t checksum;          // FNV-1a 64 of every byte after the header
    uint64_t reserved2;
};
static_assert(sizeof(FileHeader) == 64, "file header is 64 bytes");

uint64_t fnv1a(uint64_t h, const void* p, size_t n) {
    const unsigned char* b = static_cast<const unsigned char*>(p);
    for (size_t i = 0; i < n; ++i) h = (h ^ b[i]) * 0x100000001b3ull;
    return h;
}

struct FileCloser {
    std::FILE* f;
    ~FileCloser() { if (f) std::fclose(f); }
};

int read_header(std::FILE* f, const char* path, FileHeader& h) {
    if (std::fread(&h, sizeof(h), 1, f) != 1)
        return fail(sg::kInvalid, std::string("sg_file: short header in ") + path);
    if (std::memcmp(h.magic, kFileMagic, 8) != 0 || h.header_bytes != sizeof(FileHeader))
        return fail(sg::kInvalid, std::string("sg_file: not a grid file: ") + path);
    if (h.d < 1 || h.d > sg::kMaxDims || h.n < 1 || h.kind > 1 ||
        (h.elem_bytes != 0 && h.elem_bytes != 4 && h.elem_bytes != 8) || h.N == 0 ||
        h.S == 0 || h.N > uint64_t(1) << 31 || h.S > h.N)
        return fail(sg::kInvalid, std::string("sg_file: bad header in ") + path);
    return sg::kOk;
}

}  // namespace

int sg_file_write(const char* path, const sg_grid* grid, const void* values, int elem_bytes) {
    if (!path || !grid || (elem_bytes != 0 && elem_bytes != 4 && elem_bytes != 8) ||
        (elem_bytes != 0 && !values))
        return fail(sg::kInvalid, "sg_file_write: bad argument");
    const sg::Grid& g = grid->g;
    FileHeader h{};
    std::memcpy(h.magic, kFileMagic, 8);
    h.header_bytes = sizeof(FileHeader);
    h.d = uint32_t(g.d);
    h.n = uint32_t(g.n);
    h.kind = g.explicit_set ? 1u : 0u;
    h.elem_bytes = uint32_t(elem_bytes);
    h.N = uint64_t(g.N);
    h.S = uint64_t(g.S);
    const std::vector<int32_t> lim(g.limits, g.limits + g.d);   // d <= 20 entries
    const void* shape = g.explicit_set ? static_cast<const void*>(g.sub_levels.data())
                                       : static_cast<const void*>(lim.data());
    const size_t shape_bytes = g.explicit_set ? g.sub_levels.size() : lim.size() * 4;
    const size_t value_bytes = size_t(g.N) * size_t(elem_bytes);
    uint64_t c = fnv1a(0xcbf29ce484222325ull, shape, shape_bytes);
    h.checksum = fnv1a(c, values, value_bytes);
    std::FILE* f = std::fopen(path, "wb");
    if (!f) return fail(sg::kInvalid, std::string("sg_file_write: cannot open ") + path);
    FileCloser fc{f};
    if (std::fwrite(&h, sizeof(h), 1, f) != 1 || std::fwrite(shape, 1, shape_bytes, f) != shape_bytes ||
        (value_bytes && std::fwrite(values, 1, value_bytes, f) != value_bytes))
        return fail(sg::kInvalid, std::string("sg_file_write: write failed: ") + path);
    if (std::fclose(f) != 0) {
        fc.f = nullptr;
        return fail(sg::kInvalid, std::string("sg_file_write: write failed: ") + path);
    }
    fc.f = nullptr;
    return sg::kOk;
}

int sg_file_info(const char* path, int* num_dims, int* num_levels, long long* num_grid_points,
                 long long* num_subspaces, int* elem_bytes) {
    if (!path) return fail(sg::kInvalid, "sg_file_info: NULL path");
    std::FILE* f = std::fopen(path, "rb");
    if (!f) return fail(sg::kInvalid, std::string("sg_file_info: cannot open ") + path);
    FileCloser fc{f};
    FileHeader h{};
    int st = read_header(f, path, h);
    if (st != sg::kOk) return st;
    if (num_dims) *num_dims = int(h.d);
    if (num_levels) *num_levels = int(h.n);
    if (num_grid_points) *num_grid_points = (long long)h.N;
    if (num_subspaces) *num_subspaces = (long long)h.S;
    if (elem_bytes) *elem_bytes = int(h.elem_bytes);
    return sg::kOk;
}

int sg_file_read(const char* path, sg_grid** out, void* values, long long values_bytes) {
    if (!path || !out) return fail(sg::kInvalid, "sg_file_read: NULL argument");
    *out = nullptr;
    std::FILE* f = std::fopen(path, "rb");
    if (!f) return fail(sg::kInvalid, std::string("sg_file_read: cannot open ") + path);
    FileCloser fc{f};
    FileHeader h{};
    int st = read_header(f, path, h);
    if (st != sg::kOk) return st;
    const size_t shape_bytes = h.kind == 1 ? size_t(h.S) * h.d : size_t(h.d) * 4;
    const size_t value_bytes = size_t(h.N) * h.elem_bytes;
    if (values && values_bytes < (long long)value_bytes)
        return fail(sg::kRange, "sg_file_read: value buffer smaller than the file's values");
    // the file must hold what the header announces before anything is allocated for it
    if (std::fseek(f, 0, SEEK_END) != 0)
        return fail(sg::kInvalid, std::string("sg_file_read: cannot seek in ") + path);
    const long end = std::ftell(f);
    if (end < 0 || uint64_t(end) < sizeof(FileHeader) + shape_bytes + value_bytes ||
        std::fseek(f, long(sizeof(FileHeader)), SEEK_SET) != 0)
        return fail(sg::kInvalid, std::string("sg_file_read: truncated file ") + path);
    std::vector<unsigned char> shape, vals;
    try {
        shape.resize(shape_bytes);
        vals.resize(values ? value_bytes : 0);
    } catch (const std::bad_alloc&) {
        return fail(sg::kNoMem, "sg_file_read: host allocation failed");
    }
    if (std::fread(shape.data(), 1, shape_bytes, f) != shape_bytes)
        return fail(sg::kInvalid, std::string("sg_file_read: truncated level set in ") + path);
    uint64_t c = fnv1a(0xcbf29ce484222325ull, shape.data(), shape_bytes);
    if (values) {
        if (std::fread(vals.data(), 1, value_bytes, f) != value_bytes)
            return fail(sg::kInvalid, std::string("sg_file_read: truncated values in ") + path);
        c = fnv1a(c, vals.data(), value_bytes);
        if (c != h.checksum)
            return fail(sg::kInvalid, std::string("sg_file_read: checksum mismatch in ") + path);
    }
    sg_grid* g = nullptr;
    if (h.kind == 1) {
        st = sg_grid_create_explicit(int(h.d), shape.data(), (long long)h.S, &g);
    } else {
        std::vector<int> lim(h.d);
        std::memcpy(lim.data(), shape.data(), shape_bytes);
        st = sg_grid_create(int(h.d), int(h.n), lim.data(), &g);
    }
    if (st != sg::kOk) return st;
    if (uint64_t(g->g.N) != h.N || uint64_t(g->g.S) != h.S || uint32_t(g->g.n) != h.n) {
        sg_grid_destroy(g);
        return fail(sg::kInvalid, std::string("sg_file_read: level set does not match the header in ") + path);
    }
    if (values) std::memcpy(values, vals.data(), value_bytes);
    *out = g;
    return sg::kOk;
}

int sg_gather_probe_f64(const sg_grid* grid, const double* d_alpha, long long num_points,
                        unsigned long long seed, double* d_out, void* stream) {
    if (!grid || num_points < 0 || (num_points > 0 && (!d_alpha || !d_out)))
        return fail(sg::kInvalid, "sg_gather_probe_f64: bad argument");
    int st = check_device(grid);
    if (st != sg::kOk) return st;
    return sg::launch_gather_probe_f64(grid->g, grid->t, d_alpha, num_points, uint32_t(seed),
                                       d_out, as_stream(stream));
}

int sg_gather_probe_f32(const sg_grid* grid, const float* d_alpha, long long num_points,
                        unsigned long long seed, float* d_out, void* stream) {
    if (!grid || num_points < 0 || (num_points > 0 && (!d_alpha || !d_out)))
        return fail(sg::kInvalid, "sg_gather_probe_f32: bad argument");
    int st = check_device(grid);
    if (st != sg::kOk) return st;
    return sg::launch_gather_probe_f32(grid->g, grid->t, d_alpha, num_points, uint32_t(seed),
                                       d_out, as_stream(stream));
}

int sg_stage_probe_f64(const sg_grid* grid, const double* d_alpha, long long num_points,
                       unsigned long long seed, double* d_out, void* stream) {
    if (!grid || num_points < 0 || (num_points > 0 && (!d_alpha || !d_out)))
        return fail(sg::kInvalid, "sg_stage_probe_f64: bad argument");
    int st = check_device(grid);
    if (st != sg::kOk) return st;
    return sg::launch_stage_probe_f64(grid->g, grid->t, grid->tune.variant, d_alpha, num_points,
                                      uint32_t(seed), d_out, as_stream(stream));
}

int sg_stage_probe_f32(const sg_grid* grid, const float* d_alpha, long long num_points,
                       unsigned long long seed, float* d_out, void* stream) {
    if (!grid || num_points < 0 || (num_points > 0 && (!d_alpha || !d_out)))
        return fail(sg::kInvalid, "sg_stage_probe_f32: bad argument");
    int st = check_device(grid);
    if (st != sg::kOk) return st;
    return sg::launch_stage_probe_f32(grid->g, grid->t, grid->tune32.variant, d_alpha,
                                      num_points, uint32_t(seed), d_out, as_stream(stream));
}

int sg_points_to_row_major_f64(const double* d_in, int layout, int num_dims, long long num_points,
                               double* d_out, void* stream) {
    if (num_points < 0 || (num_points > 0 && (!d_in || !d_out)) || d_in == d_out)
        return fail(sg::kInvalid, "sg_points_to_row_major_f64: bad argument");
    return sg::launch_points_to_row_major_f64(d_in, layout, num_dims, num_points, d_out,
                                              as_stream(stream));
}

int sg_points_to_row_major_f32(const float* d_in, int layout, int num_dims, long long num_points,
                               float* d_out, void* stream) {
    if (num_points < 0 || (num_points > 0 && (!d_in || !d_out)) || d_in == d_out)
        return fail(sg::kInvalid, "sg_points_to_row_major_f32: bad argument");
    return sg::launch_points_to_row_major_f32(d_in, layout, num_dims, num_points, d_out,
                                              as_stream(stream));
}

int sg_convert_f64_to_f32(const double* d_in, float* d_out, long long count, void* stream) {
    if (count < 0 || (count > 0 && (!d_in || !d_out)))
        return fail(sg::kInvalid, "sg_convert_f64_to_f32: bad argument");
    return sg::launch_f64_to_f32(d_in, d_out, count, as_stream(stream));
}

}  // extern "C"

extern "C" {

int sg_batch_create(sg_grid* const* grids, int count, sg_batch** out) {
    if (!grids || count < 1 || !out) return fail(sg::kInvalid, "sg_batch_create: bad argument");
    *out = nullptr;
    std::vector<sg::Batch::GridRef> refs;
    for (int i = 0; i < count; ++i) {
        if (!grids[i]) return fail(sg::kInvalid, "sg_batch_create: NULL grid");
        // the evaluate kernels address a grid's surpluses through a buffer of at most
        // 2 GiB; FP64 batches check their members again at evaluate
        if (grids[i]->g.N * int64_t(sizeof(float)) > int64_t(0x7fffffff))
            return fail(sg::kRange, "sg_batch_create: grid " + std::to_string(i) +
                                        "'s surpluses exceed 2 GiB");
    }
    for (int i = 0; i < count; ++i) {
        int st = check_device(grids[i]);           // uploads the grid's tables
        if (st != sg::kOk) return st;
        refs.push_back({&grids[i]->g, &grids[i]->t, grids[i]->tune.variant,
                        grids[i]->tune32.variant});
    }
    sg_batch* b = new (std::nothrow) sg_batch();
    if (!b) return fail(sg::kNoMem, "sg_batch_create: host allocation failed");
    (void)hipGetDevice(&b->device);
    int st = b->b.build(std::move(refs));
    if (st != sg::kOk) {
        b->b.release();
        delete b;
        return st;
    }
    *out = b;
    return sg::kOk;
}

int sg_batch_destroy(sg_batch* batch) {
    if (!batch) return sg::kOk;
    batch->b.release();
    delete batch;
    return sg::kOk;
}

int sg_batch_values(const sg_batch* batch, long long* total, long long* value_offsets) {
    if (!batch) return fail(sg::kInvalid, "sg_batch_values: NULL batch");
    if (total) *total = batch->b.value_off.back();
    if (value_offsets)
        for (size_t i = 0; i < batch->b.value_off.size(); ++i) value_offsets[i] = batch->b.value_off[i];
    return sg::kOk;
}

int sg_batch_set_points(sg_batch* batch, const long long* counts, long long* point_offsets,
                        long long* coord_offsets) {
    if (!batch || !counts) return fail(sg::kInvalid, "sg_batch_set_points: NULL argument");
    std::lock_guard<std::mutex> lock(batch->b.mu);
    std::vector<int64_t> c(counts, counts + batch->b.grids.size());
    int st = batch->b.set_points(c);
    if (st != sg::kOk) return st;
    batch->points_set = true;
    for (size_t i = 0; i < batch->b.point_off.size(); ++i) {
        if (point_offsets) point_offsets[i] = batch->b.point_off[i];
        if (coord_offsets) coord_offsets[i] = batch->b.coord_off[i];
    }
    return sg::kOk;
}

int sg_batch_hierarchize_f64(sg_batch* batch, double* d_values, void* stream) {
    if (!batch || !d_values) return fail(sg::kInvalid, "sg_batch_hierarchize_f64: NULL argument");
    int cur = -1;
    (void)hipGetDevice(&cur);
    if (cur != batch->device) return fail(sg::kInvalid, "current device differs from the batch's");
    return batch->b.hierarchize_f64(d_values, as_stream(stream));
}

int sg_batch_hierarchize_f32(sg_batch* batch, float* d_values, void* stream) {
    if (!batch || !d_values) return fail(sg::kInvalid, "sg_batch_hierarchize_f32: NULL argument");
    int cur = -1;
    (void)hipGetDevice(&cur);
    if (cur != batch->device) return fail(sg::kInvalid, "current device differs from the batch's");
    return batch->b.hierarchize_f32(d_values, as_stream(stream));
}

int sg_batch_evaluate_f64(sg_batch* batch, const double* d_alpha, const double* d_coords,
                          double* d_out, void* stream) {
    if (!batch || !d_alpha || !d_coords || !d_out)
        return fail(sg::kInvalid, "sg_batch_evaluate_f64: NULL argument");
    int st = batch_device(batch);
    if (st != sg::kOk) return st;
    return batch->b.evaluate_f64(d_alpha, d_coords, d_out, as_stream(stream));
}

int sg_batch_evaluate_f32c(sg_batch* batch, const float* d_alpha, const float* d_coords,
                           float* d_out, void* stream) {
    if (!batch || !d_alpha || !d_coords || !d_out)
        return fail(sg::kInvalid, "sg_batch_evaluate_f32c: NULL argument");
    int st = batch_device(batch);
    if (st != sg::kOk) return st;
    return batch->b.evaluate_f32c(d_alpha, d_coords, d_out, as_stream(stream));
}

int sg_batch_run_graph_f64(sg_batch* batch, double* d_alpha, const double* d_coords,
                           double* d_out, int with_hierarchize, void* stream) {
    if (!batch || !d_alpha || !d_coords || !d_out)
        return fail(sg::kInvalid, "sg_batch_run_graph_f64: NULL argument");
    int st = batch_device(batch);
    if (st != sg::kOk) return st;
    return batch->b.run_graph_f64(d_alpha, d_coords, d_out, with_hierarchize != 0,
                                  as_stream(stream));
}

int sg_batch_run_graph_f32c(sg_batch* batch, float* d_alpha, const float* d_coords,
                            float* d_out, int with_hierarchize, void* stream) {
    if (!batch || !d_alpha || !d_coords || !d_out)
        return fail(sg::kInvalid, "sg_batch_run_graph_f32c: NULL argument");
    int st = batch_device(batch);
    if (st != sg::kOk) return st;
    return batch->b.run_graph_f32c(d_alpha, d_coords, d_out, with_hierarchize != 0,
                                   as_stream(stream));
}

}  // extern "C"

struct sg_server {
    sg::Server s;
    int precision = SG_SERVE_F64;
};

extern "C" {

int sg_server_create(const sg_server_config* cfg, sg_server** out) {
    if (!cfg || !out) return fail(sg::kInvalid, "sg_server_create: NULL argument");
    *out = nullptr;
    if (cfg->precision != SG_SERVE_F64 && cfg->precision != SG_SERVE_F32C)
        return fail(sg::kInvalid, "sg_server_create: unknown precision");
    sg::ServerConfig c;
    c.elem_bytes = cfg->precision == SG_SERVE_F64 ? 8 : 4;
    c.slots = cfg->slots;
    c.max_batch_points = cfg->max_batch_points;
    c.max_requests = cfg->max_requests;
    c.budget_bytes = cfg->budget_bytes;
    if (c.budget_bytes <= 0) {
        size_t free_b = 0, total_b = 0;
        if (hipMemGetInfo(&free_b, &total_b) != hipSuccess)
            return fail(sg::kHip, "sg_server_create: hipMemGetInfo failed");
        c.budget_bytes = int64_t(free_b / 4);
    }
    sg_server* h = new (std::nothrow) sg_server();
    if (!h) return fail(sg::kNoMem, "sg_server_create: host allocation failed");
    h->precision = cfg->precision;
    int st = h->s.start(c);
    if (st != sg::kOk) {
        delete h;
        return st;
    }
    *out = h;
    return sg::kOk;
}

int sg_server_destroy(sg_server* server) {
    delete server;
    return sg::kOk;
}

int sg_server_add_grid(sg_server* server, const sg_grid* grid, const void* d_alpha,
                       int* grid_id) {
    if (!server || !grid || !d_alpha || !grid_id)
        return fail(sg::kInvalid, "sg_server_add_grid: NULL argument");
    int st = check_device(grid);
    if (st != sg::kOk) return st;
    const int variant = server->precision == SG_SERVE_F64 ? grid->tune.variant : grid->tune32.variant;
    return server->s.add_grid(&grid->g, &grid->t, d_alpha, variant, grid_id);
}

int sg_server_submit(sg_server* server, int grid_id, const void* points, long long num_points,
                     void* out, long long* request_id) {
    if (!server || !request_id) return fail(sg::kInvalid, "sg_server_submit: NULL argument");
    int64_t id = 0;
    int st = server->s.submit(grid_id, points, num_points, out, &id);
    *request_id = id;
    return st;
}

int sg_server_poll(sg_server* server, long long request_id, int* done) {
    if (!server || !done) return fail(sg::kInvalid, "sg_server_poll: NULL argument");
    return server->s.poll(request_id, done);
}

int sg_server_wait(sg_server* server, long long request_id, double* latency_seconds) {
    if (!server) return fail(sg::kInvalid, "sg_server_wait: NULL server");
    return server->s.wait(request_id, latency_seconds);
}

int sg_server_inject_launch_failure(sg_server* server, int grid_id, int launches) {
    if (!server) return fail(sg::kInvalid, "sg_server_inject_launch_failure: NULL server");
    return server->s.inject_launch_failure(grid_id, launches);
}

int sg_server_stats(sg_server* server, long long* launches, long long* points_done,
                    long long* waiting, long long* budget_used) {
    if (!server) return fail(sg::kInvalid, "sg_server_stats: NULL server");
    int64_t a = 0, b = 0, c = 0, d = 0;
    int st = server->s.stats(&a, &b, &c, &d);
    if (launches) *launches = a;
    if (points_done) *points_done = b;
    if (waiting) *waiting = c;
    if (budget_used) *budget_used = d;
    return st;
}

}  // extern "C"
