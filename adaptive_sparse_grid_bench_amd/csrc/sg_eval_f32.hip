// FP32 instantiations of the evaluation walk (sg_eval.hpp): FP32 surpluses, points and
// hat products with Kahan-compensated accumulation across subspaces.
#include "sg_eval.hpp"

namespace sg {

int launch_evaluate_f32c(const Grid& g, const DeviceTables& t, const float* alpha,
                         const float* coords, int64_t P, float* out, const EvalTuning& tu,
                         const uint32_t* perm, hipStream_t s) {
    return eval::evaluate_launch<float, eval::AccKahan>(g, t, alpha, coords, P, out, tu, perm, s);
}

int launch_stage_probe_f32(const Grid& g, const DeviceTables& t, int variant, const float* alpha,
                          int64_t P, uint32_t seed, float* out, hipStream_t s) {
    return eval::launch_stage_probe<float>(g, t, variant, alpha, P, seed, out, s);
}

}  // namespace sg
