// A batch of independent grids resident together (sg_batch.hip).
#pragma once

#include <hip/hip_runtime_api.h>

#include <cstdint>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include "sg_grid.hpp"
#include "sg_kernels.hpp"

namespace sg {

struct Batch {
    struct GridRef {
        const Grid* g;
        const DeviceTables* t;
        int variant_f64;        // the grid's tuned walk variants
        int variant_f32;
    };
    struct HierDim {            // one launch of the hierarchization
        bool pair = false;      // plane families of (k, k+1), else pole families of k
        int32_t* tiles = nullptr;
        int32_t* fams = nullptr;
        int32_t* subs = nullptr;
        int32_t* blocks = nullptr;
        int num_tiles = 0;
        int tile_values = 0;
        int max_families = 0;
        int max_subs = 0;
        int max_blocks = 0;
    };
    struct Group {               // one evaluation launch
        bool explicit_set;
        int d;
        int lt;
        int variant_f64;
        int variant_f32;
        int nm1;                 // n - 1 of the grids (block-staged groups), else -1
        int64_t first;           // first work item
        int64_t count;           // work items (= workgroups)
    };
    using GraphKey = std::tuple<const void*, const void*, const void*, bool, int>;

    std::vector<GridRef> grids;
    std::vector<int64_t> value_off;   // [G+1]: grid g's surpluses start here
    int64_t max_grid_values = 0;      // largest member grid
    std::vector<int64_t> point_off;   // [G+1]: grid g's points / outputs start here
    std::vector<int64_t> coord_off;   // [G+1]: grid g's coordinates start here (elements)
    std::vector<HierDim> hier;        // the sweeps, in launch order
    std::vector<Group> groups;
    eval::GridDesc* d_desc = nullptr;
    eval::WorkItem* d_work = nullptr;
    std::map<GraphKey, hipGraphExec_t> graphs;
    std::mutex mu;

    int build(std::vector<GridRef> members);
    int set_points(const std::vector<int64_t>& counts);
    int hierarchize_f64(double* v, hipStream_t s) const;
    int hierarchize_f32(float* v, hipStream_t s) const;
    int evaluate_f64(const double* alpha, const double* coords, double* out, hipStream_t s) const;
    int evaluate_f32c(const float* alpha, const float* coords, float* out, hipStream_t s) const;
    int run_graph_f64(double* alpha, const double* coords, double* out, bool with_hier,
                      hipStream_t s);
    int run_graph_f32c(float* alpha, const float* coords, float* out, bool with_hier,
                       hipStream_t s);
    void invalidate_graphs();
    void release();

    template <class T>
    int hierarchize(T* v, hipStream_t s) const;
    template <class T>
    int check_range() const;
    template <class T, class Acc>
    int evaluate(const T* alpha, const T* coords, T* out, hipStream_t s) const;
    template <class T, class Acc>
    int evaluate_group(const Group& gr, const T* alpha, const T* coords, T* out,
                       hipStream_t s) const;
    template <class T, class Acc>
    int run_graph(T* alpha, const T* coords, T* out, bool with_hierarchize, hipStream_t s);
};

}  // namespace sg
