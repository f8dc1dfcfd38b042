// Batches of independent grids (the batched form): surpluses of all grids in one
// buffer, one hierarchization launch per dimension sweep over every grid's pole-family
// tiles, one evaluation launch per (dimensions, table length, level-set kind) group over
// device-resident work lists, and the whole sequence captured once into a hipGraph
// whose evaluation groups run on parallel branches.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <map>
#include <mutex>
#include <new>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/sgrid.h"
#include "sg_batch.hpp"
#include "sg_eval.hpp"

namespace sg {

namespace {

template <class T>
int upload_vec(T*& dst, const std::vector<T>& src) {
    if (dst) (void)hipFree(dst);
    dst = nullptr;
    if (src.empty()) return kOk;
    hipError_t e = hipMalloc(&dst, src.size() * sizeof(T));
    if (e != hipSuccess) {
        dst = nullptr;
        set_error(std::string("batch: hipMalloc: ") + hipGetErrorString(e));
        return e == hipErrorOutOfMemory ? kNoMem : kHip;
    }
    e = hipMemcpy(dst, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        set_error(std::string("batch: hipMemcpy: ") + hipGetErrorString(e));
        return kHip;
    }
    return kOk;
}

template <class T>
void free_ptr(T*& p) {
    if (p) (void)hipFree(p);
    p = nullptr;
}

}  // namespace

int Batch::build(std::vector<GridRef> members) {
    grids = std::move(members);
    const int G = int(grids.size());
    value_off.assign(G + 1, 0);
    for (int i = 0; i < G; ++i) value_off[i + 1] = value_off[i] + grids[i].g->N;
    if (value_off[G] >= (int64_t(1) << 31)) {
        set_error("batch: more than 2^31 grid points in total");
        return kRange;
    }
    // the evaluate kernels address one grid's surpluses through a buffer of at most
    // 2 GiB (as a single-grid evaluate does)
    max_grid_values = 0;
    for (int i = 0; i < G; ++i) {
        if (grids[i].g->N * int64_t(sizeof(float)) > int64_t(0x7fffffff)) {
            set_error("batch: grid " + std::to_string(i) + "'s surpluses exceed 2 GiB");
            return kRange;
        }
        max_grid_values = std::max(max_grid_values, grids[i].g->N);
    }
    int maxd = 0;
    for (auto& r : grids) maxd = std::max(maxd, r.g->d);
    // Sweeps in launch order.  For each dimension pair (k, k+1): one plane-family sweep
    // over the grids whose pair tables exist, then pole-family sweeps of k and of k+1 over
    // the others (grids are independent, so only each grid's own order matters).
    hier.clear();
    auto uses_pair = [&](const Grid& g, int k) {
        return k % 2 == 0 && k + 1 < g.d && g.hier2[k / 2].ok;
    };
    auto add_pole_sweep = [&](int k) -> int {
        std::vector<int32_t> tiles, fams, blocks;
        int tv = 0, mf = 0;
        for (int i = 0; i < G; ++i) {
            const Grid& g = *grids[i].g;
            if (k >= g.d || uses_pair(g, k - k % 2)) continue;
            const HierTables& h = g.hier[k];
            const int32_t fam0 = int32_t(fams.size() / 3), blk0 = int32_t(blocks.size());
            for (size_t t = 0; t < h.tiles.size(); t += 3) {
                tiles.push_back(h.tiles[t] + fam0);
                tiles.push_back(h.tiles[t + 1]);
                tiles.push_back(h.tiles[t + 2]);
                tv = std::max(tv, int(h.tiles[t + 2]));
            }
            for (size_t f = 0; f < h.fams.size(); f += 3) {
                fams.push_back(h.fams[f]);
                fams.push_back(h.fams[f + 1]);
                fams.push_back(h.fams[f + 2] + blk0);
            }
            for (int32_t b : h.blocks) blocks.push_back(int32_t(b + value_off[i]));
            mf = std::max(mf, h.max_families_per_tile);
        }
        if (tiles.empty()) return kOk;
        HierDim hd;
        hd.num_tiles = int(tiles.size() / 3);
        hd.tile_values = tv;
        hd.max_families = mf;
        int st;
        if ((st = upload_vec(hd.tiles, tiles)) != kOk) return st;
        if ((st = upload_vec(hd.fams, fams)) != kOk) return st;
        if ((st = upload_vec(hd.blocks, blocks)) != kOk) return st;
        hier.push_back(hd);
        return kOk;
    };
    auto add_pair_sweep = [&](int k) -> int {
        std::vector<int32_t> tiles, fams, subs, blocks;
        int tv = 0, mf = 0, ms = 0, mb = 0;
        for (int i = 0; i < G; ++i) {
            const Grid& g = *grids[i].g;
            if (!uses_pair(g, k)) continue;
            const PairTables& h = g.hier2[k / 2];
            const int32_t fam0 = int32_t(fams.size() / 3), sub0 = int32_t(subs.size() / 2),
                          blk0 = int32_t(blocks.size());
            for (size_t t = 0; t < h.tiles.size(); t += kPairTileInts) {
                tiles.push_back(h.tiles[t] + fam0);
                tiles.push_back(h.tiles[t + 1]);
                tiles.push_back(h.tiles[t + 2]);
                tiles.push_back(h.tiles[t + 3] + sub0);
                tiles.push_back(h.tiles[t + 4]);
                tiles.push_back(h.tiles[t + 5] + blk0);
                tiles.push_back(h.tiles[t + 6]);
                tiles.push_back(0);
                tv = std::max(tv, int(h.tiles[t + 2]));
            }
            for (size_t f = 0; f < h.fams.size(); f += 3) {
                fams.push_back(h.fams[f]);
                fams.push_back(h.fams[f + 1]);
                fams.push_back(h.fams[f + 2] + sub0);
            }
            for (size_t q = 0; q < h.subs.size(); q += 2) {
                subs.push_back(h.subs[q]);
                subs.push_back(h.subs[q + 1] + blk0);
            }
            for (int32_t b : h.blocks) blocks.push_back(int32_t(b + value_off[i]));
            mf = std::max(mf, h.max_families_per_tile);
            ms = std::max(ms, h.max_subs_per_tile);
            mb = std::max(mb, h.max_blocks_per_tile);
        }
        if (tiles.empty()) return kOk;
        HierDim hd;
        hd.pair = true;
        hd.num_tiles = int(tiles.size() / kPairTileInts);
        hd.tile_values = tv;
        hd.max_families = mf;
        hd.max_subs = ms;
        hd.max_blocks = mb;
        int st;
        if ((st = upload_vec(hd.tiles, tiles)) != kOk) return st;
        if ((st = upload_vec(hd.fams, fams)) != kOk) return st;
        if ((st = upload_vec(hd.subs, subs)) != kOk) return st;
        if ((st = upload_vec(hd.blocks, blocks)) != kOk) return st;
        hier.push_back(hd);
        return kOk;
    };
    for (int k = 0; k < maxd; k += 2) {
        int st;
        if ((st = add_pair_sweep(k)) != kOk) return st;
        if ((st = add_pole_sweep(k)) != kOk) return st;
        if (k + 1 < maxd && (st = add_pole_sweep(k + 1)) != kOk) return st;
    }
    return kOk;
}

int Batch::set_points(const std::vector<int64_t>& counts) {
    const int G = int(grids.size());
    point_off.assign(G + 1, 0);
    coord_off.assign(G + 1, 0);
    for (int i = 0; i < G; ++i) {
        if (counts[i] < 0 || counts[i] > 0x7fffffff) {
            set_error("batch: bad point count");
            return kInvalid;
        }
        point_off[i + 1] = point_off[i] + counts[i];
        coord_off[i + 1] = coord_off[i] + counts[i] * grids[i].g->d;
    }
    // descriptors and work lists, grouped by (explicit set, dimensions, table length)
    std::vector<eval::GridDesc> desc(G);
    // key: explicit set, d, table length, FP64 and FP32 variants, and n - 1 where a
    // variant is the block-staged walk (its kernels are built per n)
    std::map<std::tuple<int, int, int, int, int, int>, std::vector<eval::WorkItem>> groups_h;
    for (int i = 0; i < G; ++i) {
        const Grid& g = *grids[i].g;
        eval::GridDesc& D = desc[i];
        D = grids[i].t->host_desc;               // walk and stage tables
        D.alpha_off = value_off[i];
        D.alpha_ptr = nullptr;
        D.coord_off = coord_off[i];
        D.out_off = point_off[i];
        // table length valid for every variant (variant 0 tables the most dimensions)
        const int lt = eval::table_levels(g, 0);
        int v64 = work_list_variant(g, grids[i].variant_f64, 8);
        int v32 = work_list_variant(g, grids[i].variant_f32, 4);
        // both precisions share the group's work items, which hold as many points as the
        // walks' workgroups: where those differ (one precision at 4 waves per SIMD, 512
        // threads) that precision runs the staged walk instead
        const int p64 = work_item_points(g, v64, 8), p32 = work_item_points(g, v32, 4);
        if (p64 > p32) v64 = effective_variant(g, kStagedVariant, 8);
        if (p32 > p64) v32 = effective_variant(g, kStagedVariant, 4);
        const int64_t per_item = std::min(p64, p32);
        const int nm1 = (v64 == kBlkVariant || v32 == kBlkVariant) ? g.n - 1 : -1;
        auto& items = groups_h[std::make_tuple(g.explicit_set ? 1 : 0, g.d, lt, v64, v32, nm1)];
        for (int64_t f = 0; f < counts[i]; f += per_item)
            items.push_back({i, int32_t(f), int32_t(std::min<int64_t>(per_item, counts[i] - f)), 0});
    }
    std::vector<eval::WorkItem> all;
    groups.clear();
    for (auto& kv : groups_h) {
        Group gr;
        gr.explicit_set = std::get<0>(kv.first) != 0;
        gr.d = std::get<1>(kv.first);
        gr.lt = std::get<2>(kv.first);
        gr.variant_f64 = std::get<3>(kv.first);
        gr.variant_f32 = std::get<4>(kv.first);
        gr.nm1 = std::get<5>(kv.first);
        gr.first = int64_t(all.size());
        gr.count = int64_t(kv.second.size());
        all.insert(all.end(), kv.second.begin(), kv.second.end());
        if (gr.count > 0) groups.push_back(gr);
    }
    int st;
    if ((st = upload_vec(d_desc, desc)) != kOk) return st;
    if ((st = upload_vec(d_work, all)) != kOk) return st;
    invalidate_graphs();
    return kOk;
}

template <class T>
int Batch::hierarchize(T* v, hipStream_t s) const {
    for (const HierDim& hd : hier) {
        if (hd.num_tiles == 0) continue;
        int st = hd.pair
            ? launch_hierarchize_pairs<T>(v, hd.tiles, hd.fams, hd.subs, hd.blocks, hd.num_tiles,
                                          hd.tile_values, hd.max_families, hd.max_subs,
                                          hd.max_blocks, s)
            : launch_hierarchize_tiles<T>(v, hd.tiles, hd.fams, hd.blocks, hd.num_tiles,
                                          hd.tile_values, hd.max_families, s);
        if (st != kOk) return st;
    }
    return kOk;
}

template <class T, class Acc>
int Batch::evaluate_group(const Group& gr, const T* alpha, const T* coords, T* out,
                          hipStream_t s) const {
    eval::Params<T> prm{};
    prm.alpha = alpha;
    prm.coords = coords;
    prm.out = out;
    prm.grids = d_desc;
    prm.work = d_work + gr.first;
    prm.perm = nullptr;
    prm.P = 0;
    const int variant = sizeof(T) == 4 ? gr.variant_f32 : gr.variant_f64;
    hipError_t e = variant == kBlkVariant
        ? eval::launch_d_blk<T, Acc>(gr.d, gr.nm1, prm, gr.count, s)
        : gr.explicit_set
        ? eval::launch_d<true, T, Acc>(gr.d, variant == kStagedVariant ? variant : 0, gr.lt,
                                       prm, gr.count, eval::kThreads, s)
        : eval::launch_d<false, T, Acc>(gr.d, variant, gr.lt, prm, gr.count, eval::kThreads, s);
    if (e != hipSuccess) {
        set_error(std::string("batch evaluate launch: ") + hipGetErrorString(e));
        return kHip;
    }
    return kOk;
}

template <class T>
int Batch::check_range() const {
    if (max_grid_values * int64_t(sizeof(T)) > int64_t(0x7fffffff)) {
        set_error("batch evaluate: a grid's surpluses exceed 2 GiB in this precision");
        return kRange;
    }
    return kOk;
}

template <class T, class Acc>
int Batch::evaluate(const T* alpha, const T* coords, T* out, hipStream_t s) const {
    if (int st = check_range<T>(); st != kOk) return st;
    for (const Group& gr : groups) {
        int st = evaluate_group<T, Acc>(gr, alpha, coords, out, s);
        if (st != kOk) return st;
    }
    return kOk;
}

// Capture hierarchize (optional) + evaluate into one graph: the sweeps in sequence on
// the capture stream, then one branch per evaluation group, joined at the end.
template <class T, class Acc>
int Batch::run_graph(T* alpha, const T* coords, T* out, bool with_hierarchize,
                     hipStream_t s) {
    if (int st = check_range<T>(); st != kOk) return st;
    std::lock_guard<std::mutex> lock(mu);
    const GraphKey key{static_cast<const void*>(alpha), static_cast<const void*>(coords),
                       static_cast<const void*>(out), with_hierarchize, int(sizeof(T))};
    auto it = graphs.find(key);
    if (it == graphs.end()) {
        hipStream_t cs = nullptr;
        hipError_t e = hipStreamCreateWithFlags(&cs, hipStreamNonBlocking);
        if (e != hipSuccess) {
            set_error(std::string("batch graph: stream: ") + hipGetErrorString(e));
            return kHip;
        }
        std::vector<hipStream_t> branch(groups.size(), nullptr);
        std::vector<hipEvent_t> done(groups.size(), nullptr);
        hipEvent_t fork = nullptr;
        int st = kOk;
        hipGraph_t graph = nullptr;
        hipGraphExec_t exec = nullptr;
        for (size_t i = 0; i < groups.size() && st == kOk; ++i) {
            if (hipStreamCreateWithFlags(&branch[i], hipStreamNonBlocking) != hipSuccess ||
                hipEventCreateWithFlags(&done[i], hipEventDisableTiming) != hipSuccess)
                st = kHip;
        }
        if (st == kOk && hipEventCreateWithFlags(&fork, hipEventDisableTiming) != hipSuccess)
            st = kHip;
        if (st == kOk && hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal) != hipSuccess)
            st = kHip;
        if (st == kOk) {
            if (with_hierarchize) st = hierarchize<T>(alpha, cs);
            if (st == kOk && hipEventRecord(fork, cs) != hipSuccess) st = kHip;
            for (size_t i = 0; i < groups.size() && st == kOk; ++i) {
                if (hipStreamWaitEvent(branch[i], fork, 0) != hipSuccess) { st = kHip; break; }
                st = evaluate_group<T, Acc>(groups[i], alpha, coords, out, branch[i]);
                if (st == kOk && hipEventRecord(done[i], branch[i]) != hipSuccess) st = kHip;
                if (st == kOk && hipStreamWaitEvent(cs, done[i], 0) != hipSuccess) st = kHip;
            }
            hipError_t ee = hipStreamEndCapture(cs, &graph);
            if (st == kOk && ee != hipSuccess) st = kHip;
        }
        if (st == kOk && hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) != hipSuccess)
            st = kHip;
        if (graph) (void)hipGraphDestroy(graph);
        for (size_t i = 0; i < groups.size(); ++i) {
            if (done[i]) (void)hipEventDestroy(done[i]);
            if (branch[i]) (void)hipStreamDestroy(branch[i]);
        }
        if (fork) (void)hipEventDestroy(fork);
        (void)hipStreamDestroy(cs);
        if (st != kOk) {
            if (sg::last_error().empty()) set_error("batch graph: capture failed");
            else set_error("batch graph: capture failed: " + sg::last_error());
            return st;
        }
        it = graphs.emplace(key, exec).first;
    }
    hipError_t e = hipGraphLaunch(it->second, s);
    if (e != hipSuccess) {
        set_error(std::string("batch graph launch: ") + hipGetErrorString(e));
        return kHip;
    }
    return kOk;
}

void Batch::invalidate_graphs() {
    for (auto& kv : graphs) (void)hipGraphExecDestroy(kv.second);
    graphs.clear();
}

void Batch::release() {
    invalidate_graphs();
    for (HierDim& hd : hier) {
        free_ptr(hd.tiles);
        free_ptr(hd.fams);
        free_ptr(hd.subs);
        free_ptr(hd.blocks);
    }
    free_ptr(d_desc);
    free_ptr(d_work);
}

int Batch::hierarchize_f64(double* v, hipStream_t s) const { return hierarchize<double>(v, s); }
int Batch::hierarchize_f32(float* v, hipStream_t s) const { return hierarchize<float>(v, s); }
int Batch::evaluate_f64(const double* a, const double* x, double* o, hipStream_t s) const {
    return evaluate<double, eval::AccF64>(a, x, o, s);
}
int Batch::evaluate_f32c(const float* a, const float* x, float* o, hipStream_t s) const {
    return evaluate<float, eval::AccKahan>(a, x, o, s);
}
int Batch::run_graph_f64(double* a, const double* x, double* o, bool h, hipStream_t s) {
    return run_graph<double, eval::AccF64>(a, x, o, h, s);
}
int Batch::run_graph_f32c(float* a, const float* x, float* o, bool h, hipStream_t s) {
    return run_graph<float, eval::AccKahan>(a, x, o, h, s);
}

}  // namespace sg
