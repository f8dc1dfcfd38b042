// sgbench: command-line driver with the reference CUDA benchmark's argv and report
// lines (adaptive_sparse_grid_bench_cuda.cu:1605-1922), running libsgrid.so.
//
//   sgbench <num. dimensions> <refinement level> <num. evals> <hier. optim.>
//           <eval. optim.> <num. evals. per warp> [f64]
//
// The optimisation selectors and evals-per-warp are accepted for compatibility (the
// reference's h_opt 0..7, e_opt 0..5) and all run the same kernels.  FP32 by default,
// as the reference; "f64" selects the FP64 path.  Grid values are the reference's test
// function at the grid points; query points are random grid points (seeded, where the
// reference seeds with time()), so the interpolant must equal the function there
// (the reference's correctness test, asg.cu:1906-1907).  Times are kernel-only; the
// transfer-inclusive times the reference reports are printed on separate lines.
#include <hip/hip_runtime_api.h>
#include <unistd.h>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <vector>

#include "../../include/sgrid.h"

namespace {

const char* kHelp =
    "Usage: sgbench <num. dimensions> <refinement level> <num. evals> <hier. optim.> "
    "<eval. optim> <num. evals. per warp> [f64]\n"
    "\t<hier. optim.> 0..7 and <eval. optim.> 0..5 are accepted for compatibility with\n"
    "\tthe reference benchmark; every value runs the same hierarchize and evaluate\n"
    "\tkernels.  f64 selects FP64 (default FP32, as the reference).\n";

#define CHECK_SG(x)                                                              \
    do {                                                                         \
        int st_ = (x);                                                           \
        if (st_ != SG_OK) {                                                      \
            std::fprintf(stderr, "%s failed: %s\n", #x, sg_last_error());        \
            return 1;                                                            \
        }                                                                        \
    } while (0)
#define CHECK_HIP(x)                                                             \
    do {                                                                         \
        hipError_t e_ = (x);                                                     \
        if (e_ != hipSuccess) {                                                  \
            std::fprintf(stderr, "%s failed: %s\n", #x, hipGetErrorString(e_));  \
            return 1;                                                            \
        }                                                                        \
    } while (0)

template <class T>
T fct(int d, const T* x) {   // asg.cu:55-62
    T res = 0;
    for (int i = 0; i < d; i++) res += x[i] * x[i] * (T(1) - x[i]) * (T(1) - x[i]);
    return res;
}

template <class T>
int run(int d, int n, int P) {
    sg_grid* g = nullptr;
    CHECK_SG(sg_grid_create(d, n, nullptr, &g));
    long long N = 0, S = 0;
    sg_grid_info(g, nullptr, nullptr, &N, &S);
    std::vector<int> lv(d), iv(d);
    std::vector<T> vals(N), x(d), coords(size_t(P) * d), out(P);
    for (long long i = 0; i < N; ++i) {
        sg_idx2gp(g, i, lv.data(), iv.data());
        for (int j = 0; j < d; ++j)
            x[j] = iv[j] * (T(1) - T(0)) / (1 << lv[j]) + (T(1) - T(0)) / (1 << (lv[j] + 1));
        vals[i] = fct(d, x.data());
    }
    std::mt19937_64 rng(42);
    for (int p = 0; p < P; ++p) {
        sg_idx2gp(g, (long long)(rng() % uint64_t(N)), lv.data(), iv.data());
        for (int j = 0; j < d; ++j)
            coords[size_t(p) * d + j] = iv[j] * (T(1) - T(0)) / (1 << lv[j]) +
                                        (T(1) - T(0)) / (1 << (lv[j] + 1));
    }
    char host[256] = {0};
    gethostname(host, sizeof(host) - 1);
    const double hi_flops = d * 3.0 * double(N) / 1e9;            // asg.cu:20-21
    const double ev_flops = P * double(S) * (11 * d + 2) / 1e9;    // asg.cu:23-24
    std::printf("==============================================\n");
    std::printf("# host name: %s\n", host);
    std::printf("# num_levels: %d, num_dims: %d, num_evals: %d\n", n, d, P);
    std::printf("# num. of gridpoints: %lld\n", N);
    std::printf("# num. of floating point ops. hierarchization: %.10lf GFlop\n", hi_flops);
    std::printf("# num. of floating point ops. evaluation: %.10lf GFlop\n", ev_flops);
    std::printf("\n");

    T *d_vals, *d_coords, *d_out;
    CHECK_HIP(hipMalloc(&d_vals, N * sizeof(T)));
    CHECK_HIP(hipMalloc(&d_coords, coords.size() * sizeof(T)));
    CHECK_HIP(hipMalloc(&d_out, size_t(P) * sizeof(T)));
    hipEvent_t e0, e1, e2, e3;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    (void)hipEventCreate(&e2);
    (void)hipEventCreate(&e3);
    auto hierarchize = [&](T* v) {
        if constexpr (sizeof(T) == 8) return sg_hierarchize_f64(g, (double*)v, nullptr);
        else return sg_hierarchize_f32(g, (float*)v, nullptr);
    };
    auto evaluate = [&]() {
        if constexpr (sizeof(T) == 8)
            return sg_evaluate_f64(g, (double*)d_vals, (double*)d_coords, P, (double*)d_out, nullptr);
        else
            return sg_evaluate_f32c(g, (float*)d_vals, (float*)d_coords, P, (float*)d_out, nullptr);
    };
    // hierarchization (a warm-up pass first, on a scratch copy)
    CHECK_HIP(hipMemcpy(d_vals, vals.data(), N * sizeof(T), hipMemcpyHostToDevice));
    CHECK_SG(hierarchize(d_vals));
    CHECK_HIP(hipDeviceSynchronize());
    (void)hipEventRecord(e0, nullptr);
    CHECK_HIP(hipMemcpy(d_vals, vals.data(), N * sizeof(T), hipMemcpyHostToDevice));
    (void)hipEventRecord(e1, nullptr);
    CHECK_SG(hierarchize(d_vals));
    (void)hipEventRecord(e2, nullptr);
    (void)hipEventSynchronize(e2);
    float t_copy = 0, t_hier = 0;
    (void)hipEventElapsedTime(&t_copy, e0, e1);
    (void)hipEventElapsedTime(&t_hier, e1, e2);
    std::printf("exec. time hierarchization: %.10lf (ms)\n", double(t_hier));
    std::printf("GFLOPS rate hierarchization: %.10lf\n", hi_flops / t_hier * 1000.0);
    std::printf("exec. time hierarchization incl. host-to-device copy: %.10lf (ms)\n",
                double(t_copy + t_hier));
    std::printf("\n");
    // evaluation
    CHECK_SG(evaluate());                                   // warm-up
    CHECK_HIP(hipDeviceSynchronize());
    (void)hipEventRecord(e0, nullptr);
    CHECK_HIP(hipMemcpy(d_coords, coords.data(), coords.size() * sizeof(T), hipMemcpyHostToDevice));
    (void)hipEventRecord(e1, nullptr);
    CHECK_SG(evaluate());
    (void)hipEventRecord(e2, nullptr);
    CHECK_HIP(hipMemcpy(out.data(), d_out, size_t(P) * sizeof(T), hipMemcpyDeviceToHost));
    (void)hipEventRecord(e3, nullptr);
    (void)hipEventSynchronize(e3);
    float t_in = 0, t_eval = 0, t_outc = 0;
    (void)hipEventElapsedTime(&t_in, e0, e1);
    (void)hipEventElapsedTime(&t_eval, e1, e2);
    (void)hipEventElapsedTime(&t_outc, e2, e3);
    std::printf("exec. time evaluation: %.10lf (ms)\n", double(t_eval));
    std::printf("GFLOPS rate evaluation: %.10lf\n", ev_flops / t_eval * 1000.0);
    std::printf("exec. time evaluation incl. transfers: %.10lf (ms)\n",
                double(t_in + t_eval + t_outc));
    std::printf("evaluated points per second: %.6e\n", P / (t_eval * 1e-3));
    std::printf("\n");
    // the reference's correctness test (asg.cu:1906-1907)
    for (int p = 0; p < P; ++p) {
        const T f = fct(d, &coords[size_t(p) * d]);
        if (!(std::fabs(double(out[p]) - double(f)) < 0.000001)) {
            std::printf("Correctness test FAILED at point %d: %.9g vs %.9g\n", p,
                        double(out[p]), double(f));
            return 1;
        }
    }
    std::printf("Correctness test passed\n");
    std::printf("==============================================\n");
    (void)hipFree(d_vals);
    (void)hipFree(d_coords);
    (void)hipFree(d_out);
    sg_grid_destroy(g);
    return 0;
}

}  // namespace

int main(int argc, char** argv) {
    if (argc != 7 && argc != 8) {
        std::printf("%s", kHelp);
        return -1;
    }
    const int d = std::atoi(argv[1]), n = std::atoi(argv[2]), P = std::atoi(argv[3]);
    const int h_opt = std::atoi(argv[4]), e_opt = std::atoi(argv[5]), w = std::atoi(argv[6]);
    const bool f64 = argc == 8 && std::strcmp(argv[7], "f64") == 0;
    if (d < 1 || n < 1 || P < 1 || h_opt < 0 || h_opt > 7 || e_opt < 0 || e_opt > 5 || w < 1 ||
        (argc == 8 && !f64)) {
        std::printf("%s", kHelp);
        return -1;
    }
    return f64 ? run<double>(d, n, P) : run<float>(d, n, P);
}
