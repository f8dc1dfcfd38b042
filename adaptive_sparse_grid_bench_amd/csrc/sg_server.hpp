// Evaluation server with continuous batching (sg_server.hip).
#pragma once

#include <cstdint>

#include "sg_grid.hpp"
#include "sg_kernels.hpp"

namespace sg {

struct ServerConfig {
    int elem_bytes = 8;                 // 8: FP64 walk, 4: FP32 Kahan walk
    int slots = 3;                      // launches in flight at once (one stream each)
    int64_t max_batch_points = 1 << 20; // points per launch
    int max_requests = 1024;            // requests per launch
    int64_t budget_bytes = 0;           // device bytes admitted requests may hold
};

class Server {
  public:
    Server();
    ~Server();
    int start(const ServerConfig& cfg);
    // variant: the grid's walk variant for the server's precision
    int add_grid(const Grid* g, const DeviceTables* t, const void* alpha, int variant, int* id);
    int submit(int grid, const void* points, int64_t P, void* out, int64_t* id);
    int wait(int64_t id, double* latency_s);
    int poll(int64_t id, int* done);
    int stats(int64_t* launches, int64_t* points, int64_t* waiting, int64_t* budget_used);
    // test hook: the next `count` launch groups holding grid's requests fail to launch
    int inject_launch_failure(int grid, int count);

  private:
    struct Impl;
    Impl* p;
};

}  // namespace sg
