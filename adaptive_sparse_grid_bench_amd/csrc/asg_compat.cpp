// libasg_compat.so: the reference benchmark's global-state C interface
// (include/asg_compat.h) on top of libsgrid.so.  Host buffers stay the caller's; each
// call stages them through device buffers kept between calls.
#include "../../include/asg_compat.h"

#include <hip/hip_runtime_api.h>

#include <cstdio>
#include <cstdlib>
#include <vector>

#include "../../include/sgrid.h"

int num_levels, num_dims, num_evals, num_subspaces;
int num_grid_points;
float* sg1d;
int** par_lili;
int* limits;
int** dp_mat;
int* lidx_vect;

namespace {

sg_grid* g_grid = nullptr;

struct DevBuf {
    void* p = nullptr;
    size_t cap = 0;
    void* get(size_t bytes) {
        if (bytes <= cap) return p;
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
        if (hipMalloc(&p, bytes) != hipSuccess) return nullptr;
        cap = bytes;
        return p;
    }
};
DevBuf d_vals, d_coords, d_out;

int report(const char* what, const char* msg) {
    std::fprintf(stderr, "asg_compat: %s: %s\n", what, msg);
    return -1;
}

int check_grid(const char* what) {
    if (!g_grid) return report(what, "init() has not been called");
    return 0;
}

// the reference's test function, FP32 in its operation order (asg.c:87-94)
float fct(const float* coords) {
    float res = 0;
    for (int i = 0; i < num_dims; i++)
        res += coords[i] * coords[i] * (1.0f - coords[i]) * (1.0f - coords[i]);
    return res;
}

int hierarchize_gpu(const char* what) {
    if (check_grid(what)) return -1;
    const size_t bytes = size_t(num_grid_points) * sizeof(float);
    float* d = static_cast<float*>(d_vals.get(bytes));
    if (!d) return report(what, "device allocation failed");
    if (hipMemcpy(d, sg1d, bytes, hipMemcpyHostToDevice) != hipSuccess)
        return report(what, "copy to device failed");
    if (sg_hierarchize_f32(g_grid, d, nullptr) != SG_OK) return report(what, sg_last_error());
    if (hipMemcpy(sg1d, d, bytes, hipMemcpyDeviceToHost) != hipSuccess)
        return report(what, "copy to host failed");
    return 0;
}

int evaluate_gpu(const char* what, float* alpha, float* coord_mat, float* out, int P) {
    if (check_grid(what)) return -1;
    if (P <= 0) return 0;
    const size_t abytes = size_t(num_grid_points) * sizeof(float);
    const size_t cbytes = size_t(P) * num_dims * sizeof(float);
    const size_t obytes = size_t(P) * sizeof(float);
    float* da = static_cast<float*>(d_vals.get(abytes));
    float* dc = static_cast<float*>(d_coords.get(cbytes));
    float* dout = static_cast<float*>(d_out.get(obytes));
    if (!da || !dc || !dout) return report(what, "device allocation failed");
    if (hipMemcpy(da, alpha, abytes, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(dc, coord_mat, cbytes, hipMemcpyHostToDevice) != hipSuccess)
        return report(what, "copy to device failed");
    if (sg_evaluate_ref_f32(g_grid, da, dc, P, dout, nullptr) != SG_OK)
        return report(what, sg_last_error());
    if (hipMemcpy(out, dout, obytes, hipMemcpyDeviceToHost) != hipSuccess)
        return report(what, "copy to host failed");
    return 0;
}

// FP64 twins: host values in, host values out, through the same device buffers
int values_f64_gpu(const char* what, double* values, bool sample, bool hierarchize) {
    if (check_grid(what)) return -1;
    const size_t bytes = size_t(num_grid_points) * sizeof(double);
    double* d = static_cast<double*>(d_vals.get(bytes));
    if (!d) return report(what, "device allocation failed");
    if (sample) {
        if (sg_sample_f64(g_grid, SG_FN_REFERENCE, nullptr, 0, d, nullptr) != SG_OK)
            return report(what, sg_last_error());
    } else if (hipMemcpy(d, values, bytes, hipMemcpyHostToDevice) != hipSuccess) {
        return report(what, "copy to device failed");
    }
    if (hierarchize && sg_hierarchize_f64(g_grid, d, nullptr) != SG_OK)
        return report(what, sg_last_error());
    if (hipMemcpy(values, d, bytes, hipMemcpyDeviceToHost) != hipSuccess)
        return report(what, "copy to host failed");
    return 0;
}

int evaluate_f64_gpu(const char* what, double* alpha, double* coord_mat, double* out, int P,
                     bool walk) {
    if (check_grid(what)) return -1;
    if (P <= 0) return 0;
    const size_t abytes = size_t(num_grid_points) * sizeof(double);
    const size_t cbytes = size_t(P) * num_dims * sizeof(double);
    const size_t obytes = size_t(P) * sizeof(double);
    double* da = static_cast<double*>(d_vals.get(abytes));
    double* dc = static_cast<double*>(d_coords.get(cbytes));
    double* dout = static_cast<double*>(d_out.get(obytes));
    if (!da || !dc || !dout) return report(what, "device allocation failed");
    if (hipMemcpy(da, alpha, abytes, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(dc, coord_mat, cbytes, hipMemcpyHostToDevice) != hipSuccess)
        return report(what, "copy to device failed");
    const int st = walk ? sg_evaluate_f64(g_grid, da, dc, P, dout, nullptr)
                        : sg_evaluate_ref_f64(g_grid, da, dc, P, dout, nullptr);
    if (st != SG_OK) return report(what, sg_last_error());
    if (hipMemcpy(out, dout, obytes, hipMemcpyDeviceToHost) != hipSuccess)
        return report(what, "copy to host failed");
    return 0;
}

}  // namespace

extern "C" {

// asg.c:101-184: tables, grid values sampled from fct, parent look-up table
void init(void) {
    if (g_grid) sg_grid_destroy(g_grid);
    g_grid = nullptr;
    if (sg_grid_create(num_dims, num_levels, limits, &g_grid) != SG_OK) {
        report("init", sg_last_error());
        return;
    }
    long long N = 0, S = 0;
    sg_grid_info(g_grid, nullptr, nullptr, &N, &S);
    num_grid_points = int(N);
    num_subspaces = int(S);
    std::vector<long long> dp(size_t(num_dims) * num_levels), lidx(num_levels + 1);
    sg_grid_tables(g_grid, dp.data(), lidx.data());
    dp_mat = static_cast<int**>(std::malloc(num_dims * sizeof(int*)));
    for (int i = 0; i < num_dims; i++) {
        dp_mat[i] = static_cast<int*>(std::malloc(num_levels * sizeof(int)));
        for (int s = 0; s < num_levels; s++) dp_mat[i][s] = int(dp[size_t(i) * num_levels + s]);
    }
    lidx_vect = static_cast<int*>(std::malloc(num_levels * sizeof(int)));
    for (int s = 0; s < num_levels; s++) lidx_vect[s] = int(lidx[s]);

    sg1d = static_cast<float*>(std::malloc(size_t(num_grid_points) * sizeof(float)));
    std::vector<int> lv(num_dims), iv(num_dims);
    std::vector<float> coords(num_dims);
    for (int i = 0; i < num_grid_points; i++) {
        sg_idx2gp(g_grid, i, lv.data(), iv.data());
        for (int j = 0; j < num_dims; j++)
            coords[j] = iv[j] * (1.0f - 0.0f) / (1 << lv[j]) + (1.0f - 0.0f) / (1 << (lv[j] + 1));
        sg1d[i] = fct(coords.data());
    }

    par_lili = static_cast<int**>(std::malloc(num_levels * sizeof(int*)));
    for (int i = 0; i < num_levels; i++) {
        par_lili[i] = static_cast<int*>(std::malloc((size_t(1) << i) * 4 * sizeof(int)));
        for (int j = 0; j < (1 << i); j++) {
            int* e = par_lili[i] + (j << 2);
            if (j == 0) {
                e[0] = -1;
                e[1] = 0;
            } else {
                get_left_li(i, j, &e[0], &e[1]);
            }
            if (j == (1 << i) - 1) {
                e[2] = -1;
                e[3] = 1;
            } else {
                get_right_li(i, j, &e[2], &e[3]);
            }
        }
    }
}

int gp2idx(int* levels, int* indices) {
    long long o = -1;
    if (check_grid("gp2idx") || sg_gp2idx(g_grid, levels, indices, &o) != SG_OK) return -1;
    return int(o);
}

int idx2gp(int index, int* levels, int* indices) {
    if (check_grid("idx2gp")) return -1;
    return sg_idx2gp(g_grid, index, levels, indices) == SG_OK ? 0 : -1;
}

int idx2l(int index, int sum, int* levels) {
    if (check_grid("idx2l")) return -1;
    return sg_idx2l(g_grid, index, sum, levels) == SG_OK ? 0 : -1;
}

int l2idx(int* levels) {
    long long r = -1;
    if (check_grid("l2idx") || sg_l2idx(g_grid, levels, &r) != SG_OK) return -1;
    return int(r);
}

// asg.c:828-854 (callers never pass an index without the parent in question)
int get_left_li(int l, int i, int* ll, int* li) {
    *ll = l - 1;
    *li = i;
    while (*li != 0 && (*li & 1) == 0) {
        *li >>= 1;
        (*ll)--;
    }
    *li >>= 1;
    return 0;
}

int get_right_li(int l, int i, int* rl, int* ri) {
    *rl = l - 1;
    *ri = i + 1;
    while (*ri != 0 && (*ri & 1) == 0) {
        *ri >>= 1;
        (*rl)--;
    }
    *ri >>= 1;
    return 0;
}

int hierarchize0(void) { return hierarchize_gpu("hierarchize0"); }
int hierarchize1(void) { return hierarchize_gpu("hierarchize1"); }
int hierarchize2(void) { return hierarchize_gpu("hierarchize2"); }
int hierarchize3(void) { return hierarchize_gpu("hierarchize3"); }
int hierarchize4(void) { return hierarchize_gpu("hierarchize4"); }
int hierarchize5(void) { return hierarchize_gpu("hierarchize5"); }
int hierarchize6(void) { return hierarchize_gpu("hierarchize6"); }
int hierarchize7(void) { return hierarchize_gpu("hierarchize7"); }
int hierarchize8(void) { return hierarchize_gpu("hierarchize8"); }
int omp_hierarchize3(void) { return hierarchize_gpu("omp_hierarchize3"); }

int evaluate0(float* a, float* c, float* o, int P) { return evaluate_gpu("evaluate0", a, c, o, P); }
int evaluate1(float* a, float* c, float* o, int P) { return evaluate_gpu("evaluate1", a, c, o, P); }
int evaluate2(float* a, float* c, float* o, int P) { return evaluate_gpu("evaluate2", a, c, o, P); }
int evaluate3(float* a, float* c, float* o, int P) { return evaluate_gpu("evaluate3", a, c, o, P); }
int evaluate4(float* a, float* c, float* o, int P) { return evaluate_gpu("evaluate4", a, c, o, P); }
int evaluate5(float* a, float* c, float* o, int P) { return evaluate_gpu("evaluate5", a, c, o, P); }
int omp_evaluate2(float* a, float* c, float* o, int P) {
    return evaluate_gpu("omp_evaluate2", a, c, o, P);
}
int omp_evaluate5(float* a, float* c, float* o, int P) {
    return evaluate_gpu("omp_evaluate5", a, c, o, P);
}

int init_values_f64(double* values) { return values_f64_gpu("init_values_f64", values, true, false); }
int hierarchize_f64(double* values) { return values_f64_gpu("hierarchize_f64", values, false, true); }
int evaluate_f64(double* a, double* c, double* o, int P) {
    return evaluate_f64_gpu("evaluate_f64", a, c, o, P, false);
}
int evaluate_walk_f64(double* a, double* c, double* o, int P) {
    return evaluate_f64_gpu("evaluate_walk_f64", a, c, o, P, true);
}

}  // extern "C"
