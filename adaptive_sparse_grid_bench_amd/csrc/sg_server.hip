// Evaluation server with continuous batching (the served form).
//
// Resident grids are registered once (surpluses stay in HBM).  Clients submit requests
// (grid, host points, host output) from any thread.  One worker thread runs the loop:
//   1. retire finished launches: copy their outputs to the requesters' buffers, release
//      their share of the device budget, wake waiters;
//   2. if a launch slot is free, admit waiting requests in arrival order while their
//      staging bytes fit the device budget and the slot's capacity, pack them into one
//      launch (one evaluation group per dimension count, over a work list of
//      (request, first point, count) items), stage the points through pinned memory and
//      enqueue copy-in, kernels and copy-out on the slot's stream.
// Several slots (streams) are in flight at once, so requests that arrive while others
// run are admitted as soon as a slot and budget free up, and their copies and kernels
// overlap the launches already on the device.  A request's results do not depend on
// what it is batched with: every point runs the same walk.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/sgrid.h"
#include "sg_eval.hpp"
#include "sg_server.hpp"

namespace sg {

namespace {
constexpr int kMaxD = kMaxDims;
}

struct Server::Impl {
    struct GridEntry {
        const Grid* g;
        const DeviceTables* t;
        const void* alpha;      // device surpluses in the server's precision
        int variant;            // walk variant that runs for this grid
    };
    struct Request {
        int64_t id;
        int grid;
        const void* points;
        int64_t P;
        void* out;
        int64_t bytes;          // staging bytes it holds while admitted
        std::chrono::steady_clock::time_point t_submit, t_done;
        int status = 1;         // 1 pending, 0 done, < 0 error
    };
    struct Slot {
        hipStream_t stream = nullptr;
        hipEvent_t done = nullptr;
        void* h_coords = nullptr;       // pinned staging
        void* h_out = nullptr;
        eval::GridDesc* h_desc = nullptr;
        eval::WorkItem* h_work = nullptr;
        void* d_coords = nullptr;
        void* d_out = nullptr;
        eval::GridDesc* d_desc = nullptr;
        eval::WorkItem* d_work = nullptr;
        std::vector<int64_t> reqs;      // requests in flight in this slot
        std::vector<int64_t> out_off;
        bool busy = false;
    };

    ServerConfig cfg;
    size_t elem;                        // bytes per value (8 or 4)
    int device = -1;
    std::vector<GridEntry> grids;
    std::vector<Slot> slots;
    std::deque<int64_t> waiting;
    std::unordered_map<int64_t, Request> reqs;
    int64_t next_id = 1;
    int64_t budget_used = 0;
    int64_t launches = 0;
    int64_t points_done = 0;
    bool stop = false;
    std::string error;
    std::unordered_map<int, int> inject;   // grid -> launch groups still to fail (test hook)
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::thread worker;

    int init() {
        if (hipGetDevice(&device) != hipSuccess) return fail("hipGetDevice failed");
        slots.resize(cfg.slots);
        const size_t P = size_t(cfg.max_batch_points);
        const size_t items = (P + eval::kThreads - 1) / eval::kThreads + size_t(cfg.max_requests);
        for (Slot& s : slots) {
            if (hipStreamCreateWithFlags(&s.stream, hipStreamNonBlocking) != hipSuccess ||
                hipEventCreateWithFlags(&s.done, hipEventDisableTiming) != hipSuccess ||
                hipHostMalloc(&s.h_coords, P * kMaxD * elem, hipHostMallocDefault) != hipSuccess ||
                hipHostMalloc(&s.h_out, P * elem, hipHostMallocDefault) != hipSuccess ||
                hipHostMalloc(reinterpret_cast<void**>(&s.h_desc),
                              size_t(cfg.max_requests) * sizeof(eval::GridDesc),
                              hipHostMallocDefault) != hipSuccess ||
                hipHostMalloc(reinterpret_cast<void**>(&s.h_work), items * sizeof(eval::WorkItem),
                              hipHostMallocDefault) != hipSuccess ||
                hipMalloc(&s.d_coords, P * kMaxD * elem) != hipSuccess ||
                hipMalloc(&s.d_out, P * elem) != hipSuccess ||
                hipMalloc(reinterpret_cast<void**>(&s.d_desc),
                          size_t(cfg.max_requests) * sizeof(eval::GridDesc)) != hipSuccess ||
                hipMalloc(reinterpret_cast<void**>(&s.d_work), items * sizeof(eval::WorkItem)) !=
                    hipSuccess)
                return fail("server: staging allocation failed");
        }
        worker = std::thread([this] { loop(); });
        return kOk;
    }

    int fail(const std::string& m) {
        set_error(m);
        return kNoMem;
    }

    void shutdown() {
        {
            std::lock_guard<std::mutex> lk(mu);
            stop = true;
        }
        cv_work.notify_all();
        if (worker.joinable()) worker.join();
        for (Slot& s : slots) {
            if (s.stream) (void)hipStreamSynchronize(s.stream);
            if (s.stream) (void)hipStreamDestroy(s.stream);
            if (s.done) (void)hipEventDestroy(s.done);
            if (s.h_coords) (void)hipHostFree(s.h_coords);
            if (s.h_out) (void)hipHostFree(s.h_out);
            if (s.h_desc) (void)hipHostFree(s.h_desc);
            if (s.h_work) (void)hipHostFree(s.h_work);
            if (s.d_coords) (void)hipFree(s.d_coords);
            if (s.d_out) (void)hipFree(s.d_out);
            if (s.d_desc) (void)hipFree(s.d_desc);
            if (s.d_work) (void)hipFree(s.d_work);
        }
        slots.clear();
    }

    // staging bytes a request holds while admitted: its points and outputs
    int64_t request_bytes(int grid, int64_t P) const {
        return P * (int64_t(grids[grid].g->d) + 1) * int64_t(elem);
    }

    // --- worker side -------------------------------------------------------------

    void retire(Slot& s, std::unique_lock<std::mutex>& lk) {
        // outputs: pinned slot buffer -> requesters' buffers (outside the lock)
        std::vector<std::pair<Request*, int64_t>> todo;
        for (size_t i = 0; i < s.reqs.size(); ++i) todo.push_back({&reqs[s.reqs[i]], s.out_off[i]});
        lk.unlock();
        for (auto& [r, off] : todo)
            std::memcpy(r->out, static_cast<char*>(s.h_out) + off * elem, size_t(r->P) * elem);
        const auto now = std::chrono::steady_clock::now();
        lk.lock();
        for (auto& [r, off] : todo) {
            r->status = 0;
            r->t_done = now;
            budget_used -= r->bytes;
            points_done += r->P;
        }
        s.reqs.clear();
        s.out_off.clear();
        s.busy = false;
        cv_done.notify_all();
    }

    // pack waiting requests into slot s; returns false if nothing was admitted
    bool admit(Slot& s, std::unique_lock<std::mutex>& lk) {
        std::vector<int64_t> batch;
        int64_t pts = 0;
        while (!waiting.empty() && int(batch.size()) < cfg.max_requests) {
            Request& r = reqs[waiting.front()];
            if (pts + r.P > cfg.max_batch_points) break;
            if (budget_used + r.bytes > cfg.budget_bytes && !batch.empty()) break;
            if (budget_used + r.bytes > cfg.budget_bytes && budget_used > 0) break;
            budget_used += r.bytes;
            pts += r.P;
            batch.push_back(r.id);
            waiting.pop_front();
        }
        if (batch.empty()) return false;
        s.busy = true;
        s.reqs = batch;
        // staging and work lists, grouped by dimension count (outside the lock)
        std::vector<Request*> rs;
        for (int64_t id : batch) rs.push_back(&reqs[id]);
        // injected failures: each pending one takes the next launch group of its grid
        std::vector<char> inject_of(rs.size(), 0);
        for (size_t i = 0; i < rs.size(); ++i) {
            auto it = inject.find(rs[i]->grid);
            if (it != inject.end() && it->second > 0) inject_of[i] = 1;
        }
        for (size_t i = 0; i < rs.size(); ++i) {
            auto it = inject.find(rs[i]->grid);
            if (inject_of[i] && it != inject.end() && it->second > 0) {
                if (--it->second == 0) inject.erase(it);
                for (size_t j = i + 1; j < rs.size(); ++j)   // one group per grid and launch
                    if (rs[j]->grid == rs[i]->grid) inject_of[j] = 1;
                break;
            }
        }
        lk.unlock();
        // a grid with a pending injected failure gets a launch group of its own
        std::map<std::tuple<int, int, int, int, int, int>, std::vector<eval::WorkItem>> groups;
        std::map<std::tuple<int, int, int, int, int, int>, std::vector<size_t>> group_reqs;
        int64_t coff = 0, ooff = 0;
        std::vector<int64_t> out_off(rs.size());
        for (size_t i = 0; i < rs.size(); ++i) {
            Request& r = *rs[i];
            const GridEntry& ge = grids[r.grid];
            const Grid& g = *ge.g;
            std::memcpy(static_cast<char*>(s.h_coords) + coff * elem, r.points,
                        size_t(r.P) * g.d * elem);
            eval::GridDesc& D = s.h_desc[i];
            D = ge.t->host_desc;                 // walk and stage tables
            D.alpha_off = 0;
            D.alpha_ptr = ge.alpha;
            D.coord_off = coff;
            D.out_off = ooff;
            const auto key = std::make_tuple(g.explicit_set ? 1 : 0, g.d, eval::table_levels(g, 0),
                                             ge.variant, inject_of[i] ? r.grid : -1,
                                             ge.variant == kBlkVariant ? g.n - 1 : -1);
            auto& items = groups[key];
            group_reqs[key].push_back(i);
            const int64_t per_item = work_item_points(g, ge.variant, int(elem));
            for (int64_t f = 0; f < r.P; f += per_item)
                items.push_back({int32_t(i), int32_t(f),
                                 int32_t(std::min<int64_t>(per_item, r.P - f)), 0});
            out_off[i] = ooff;
            coff += r.P * g.d;
            ooff += r.P;
        }
        int64_t nitems = 0;
        struct LaunchGroup {
            int expl, d, lt, variant, nm1;
            bool injected;
            int64_t first, count;
            const std::vector<size_t>* reqs;
        };
        std::vector<LaunchGroup> launches_v;
        for (auto& kv : groups) {
            std::memcpy(s.h_work + nitems, kv.second.data(), kv.second.size() * sizeof(eval::WorkItem));
            launches_v.push_back({std::get<0>(kv.first), std::get<1>(kv.first),
                                  std::get<2>(kv.first), std::get<3>(kv.first),
                                  std::get<5>(kv.first), std::get<4>(kv.first) >= 0, nitems,
                                  int64_t(kv.second.size()), &group_reqs[kv.first]});
            nitems += int64_t(kv.second.size());
        }
        hipError_t e = hipMemcpyAsync(s.d_coords, s.h_coords, size_t(coff) * elem,
                                      hipMemcpyHostToDevice, s.stream);
        if (e == hipSuccess)
            e = hipMemcpyAsync(s.d_desc, s.h_desc, rs.size() * sizeof(eval::GridDesc),
                               hipMemcpyHostToDevice, s.stream);
        if (e == hipSuccess)
            e = hipMemcpyAsync(s.d_work, s.h_work, size_t(nitems) * sizeof(eval::WorkItem),
                               hipMemcpyHostToDevice, s.stream);
        // A launch group that fails to launch fails its own requests only; the others'
        // kernels are queued and complete as usual.
        std::vector<char> failed(rs.size(), 0);
        std::string launch_error;
        for (const LaunchGroup& L : launches_v) {
            if (e != hipSuccess) break;
            const hipError_t le = L.injected
                ? hipErrorLaunchFailure
                : launch_group(L.expl != 0, L.d, L.lt, L.variant, L.nm1, L.first, L.count, s);
            if (le != hipSuccess) {
                (void)hipGetLastError();               // clear the launch error
                launch_error = std::string("server launch") + (L.injected ? " (injected)" : "") +
                               ": " + hipGetErrorString(le);
                for (size_t i : *L.reqs) failed[i] = 1;
            }
        }
        if (e == hipSuccess)
            e = hipMemcpyAsync(s.h_out, s.d_out, size_t(ooff) * elem, hipMemcpyDeviceToHost,
                               s.stream);
        if (e == hipSuccess) e = hipEventRecord(s.done, s.stream);
        const auto now = std::chrono::steady_clock::now();
        lk.lock();
        ++launches;
        if (e != hipSuccess) {
            error = std::string("server launch: ") + hipGetErrorString(e);
            (void)hipStreamSynchronize(s.stream);
            for (Request* r : rs) {
                r->status = kHip;
                r->t_done = now;
                budget_used -= r->bytes;
            }
            s.reqs.clear();
            s.out_off.clear();
            s.busy = false;
            cv_done.notify_all();
            return true;
        }
        s.reqs.clear();
        s.out_off.clear();
        bool any_failed = false;
        for (size_t i = 0; i < rs.size(); ++i) {
            if (failed[i]) {
                rs[i]->status = kHip;
                rs[i]->t_done = now;
                budget_used -= rs[i]->bytes;
                any_failed = true;
            } else {
                s.reqs.push_back(rs[i]->id);
                s.out_off.push_back(out_off[i]);
            }
        }
        if (any_failed) {
            error = launch_error;
            cv_done.notify_all();
        }
        if (s.reqs.empty()) {                      // nothing ran: free the slot
            lk.unlock();
            (void)hipStreamSynchronize(s.stream);
            lk.lock();
            s.busy = false;
        }
        return true;
    }

    hipError_t launch_group(bool expl, int d, int lt, int variant, int nm1, int64_t first,
                            int64_t count, Slot& s) {
        if (elem == 8) {
            eval::Params<double> prm{};
            prm.alpha = nullptr;                       // descriptors carry the addresses
            prm.coords = static_cast<const double*>(s.d_coords);
            prm.out = static_cast<double*>(s.d_out);
            prm.grids = s.d_desc;
            prm.work = s.d_work + first;
            if (variant == kBlkVariant)
                return eval::launch_d_blk<double, eval::AccF64>(d, nm1, prm, count, s.stream);
            return expl ? eval::launch_d<true, double, eval::AccF64>(d, variant, lt, prm, count,
                                                                     eval::kThreads, s.stream)
                        : eval::launch_d<false, double, eval::AccF64>(
                              d, variant, lt, prm, count, eval::kThreads, s.stream);
        }
        eval::Params<float> prm{};
        prm.alpha = nullptr;
        prm.coords = static_cast<const float*>(s.d_coords);
        prm.out = static_cast<float*>(s.d_out);
        prm.grids = s.d_desc;
        prm.work = s.d_work + first;
        if (variant == kBlkVariant)
            return eval::launch_d_blk<float, eval::AccKahan>(d, nm1, prm, count, s.stream);
        return expl ? eval::launch_d<true, float, eval::AccKahan>(d, variant, lt, prm, count,
                                                                  eval::kThreads, s.stream)
                    : eval::launch_d<false, float, eval::AccKahan>(
                          d, variant, lt, prm, count, eval::kThreads, s.stream);
    }

    void loop() {
        std::unique_lock<std::mutex> lk(mu);
        (void)hipSetDevice(device);
        while (true) {
            bool progressed = false;
            for (Slot& s : slots) {
                if (s.busy && !s.reqs.empty() && hipEventQuery(s.done) == hipSuccess) {
                    retire(s, lk);
                    progressed = true;
                }
            }
            for (Slot& s : slots) {
                if (!s.busy && !waiting.empty()) progressed |= admit(s, lk);
            }
            bool inflight = false;
            for (Slot& s : slots) inflight |= s.busy;
            if (stop && !inflight && waiting.empty()) break;
            if (!progressed) {
                if (inflight) {
                    // poll the device briefly; new arrivals also wake us
                    cv_work.wait_for(lk, std::chrono::microseconds(50));
                } else {
                    cv_work.wait(lk, [&] { return stop || !waiting.empty(); });
                }
            }
        }
    }
};

Server::Server() : p(new Impl()) {}
Server::~Server() {
    p->shutdown();
    delete p;
}

int Server::start(const ServerConfig& cfg) {
    if (cfg.slots < 1 || cfg.slots > 16 || cfg.max_batch_points < eval::kThreads ||
        cfg.max_requests < 1 || cfg.budget_bytes <= 0 || (cfg.elem_bytes != 8 && cfg.elem_bytes != 4)) {
        set_error("server: bad configuration");
        return kInvalid;
    }
    p->cfg = cfg;
    p->elem = size_t(cfg.elem_bytes);
    return p->init();
}

int Server::add_grid(const Grid* g, const DeviceTables* t, const void* alpha, int variant,
                     int* id) {
    std::lock_guard<std::mutex> lk(p->mu);
    if (t->device != p->device) {
        set_error("server: grid lives on another device");
        return kInvalid;
    }
    // the evaluate kernels address one grid's surpluses through a buffer of at most 2 GiB
    if (g->N * int64_t(p->elem) > int64_t(0x7fffffff)) {
        set_error("server: the grid's surpluses exceed 2 GiB in this precision");
        return kRange;
    }
    p->grids.push_back({g, t, alpha, work_list_variant(*g, variant, int(p->elem))});
    *id = int(p->grids.size()) - 1;
    return kOk;
}

int Server::submit(int grid, const void* points, int64_t P, void* out, int64_t* id) {
    std::lock_guard<std::mutex> lk(p->mu);
    if (grid < 0 || grid >= int(p->grids.size()) || P < 0 || (P > 0 && (!points || !out))) {
        set_error("server: bad request");
        return kInvalid;
    }
    if (P > p->cfg.max_batch_points || p->request_bytes(grid, P) > p->cfg.budget_bytes) {
        set_error("server: request larger than a launch slot or the device budget");
        return kRange;
    }
    Impl::Request r;
    r.id = p->next_id++;
    r.grid = grid;
    r.points = points;
    r.P = P;
    r.out = out;
    r.bytes = p->request_bytes(grid, P);
    r.t_submit = std::chrono::steady_clock::now();
    if (P == 0) {
        r.status = 0;
        r.t_done = r.t_submit;
        p->reqs[r.id] = r;
        *id = r.id;
        return kOk;
    }
    p->reqs[r.id] = r;
    p->waiting.push_back(r.id);
    *id = r.id;
    p->cv_work.notify_one();
    return kOk;
}

int Server::wait(int64_t id, double* latency_s) {
    std::unique_lock<std::mutex> lk(p->mu);
    auto it = p->reqs.find(id);
    if (it == p->reqs.end()) {
        set_error("server: unknown request");
        return kInvalid;
    }
    p->cv_done.wait(lk, [&] { return it->second.status <= 0; });
    const int st = it->second.status;
    if (latency_s)
        *latency_s = std::chrono::duration<double>(it->second.t_done - it->second.t_submit).count();
    p->reqs.erase(it);
    if (st != kOk) set_error(p->error);
    return st;
}

int Server::poll(int64_t id, int* done) {
    std::lock_guard<std::mutex> lk(p->mu);
    auto it = p->reqs.find(id);
    if (it == p->reqs.end()) {
        set_error("server: unknown request");
        return kInvalid;
    }
    *done = it->second.status <= 0 ? 1 : 0;
    return kOk;
}

int Server::inject_launch_failure(int grid, int count) {
    std::lock_guard<std::mutex> lk(p->mu);
    if (grid < 0 || grid >= int(p->grids.size()) || count < 0) {
        set_error("server: bad grid or count");
        return kInvalid;
    }
    if (count == 0) p->inject.erase(grid);
    else p->inject[grid] = count;
    return kOk;
}

int Server::stats(int64_t* launches, int64_t* points, int64_t* waiting, int64_t* budget_used) {
    std::lock_guard<std::mutex> lk(p->mu);
    if (launches) *launches = p->launches;
    if (points) *points = p->points_done;
    if (waiting) *waiting = int64_t(p->waiting.size());
    if (budget_used) *budget_used = p->budget_used;
    return kOk;
}

}  // namespace sg
