// FP64 instantiations of the evaluation walk (sg_eval.hpp).
#include "sg_eval.hpp"

namespace sg {

int launch_evaluate_f64(const Grid& g, const DeviceTables& t, const double* alpha,
                        const double* coords, int64_t P, double* out, const EvalTuning& tu,
                        const uint32_t* perm, hipStream_t s) {
    return eval::evaluate_launch<double, eval::AccF64>(g, t, alpha, coords, P, out, tu, perm, s);
}

int launch_stage_probe_f64(const Grid& g, const DeviceTables& t, int variant, const double* alpha,
                          int64_t P, uint32_t seed, double* out, hipStream_t s) {
    return eval::launch_stage_probe<double>(g, t, variant, alpha, P, seed, out, s);
}

}  // namespace sg

#include "../../include/sgrid.h"
static_assert(sg::eval::kNumVariants == SG_EVAL_VARIANTS, "variant count out of sync");
