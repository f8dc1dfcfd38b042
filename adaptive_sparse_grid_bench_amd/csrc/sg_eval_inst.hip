// Instantiations of the evaluation walk for one precision (SG_PREC = 64 or 32) and one
// range of dimensions (SG_DMIN..SG_DMAX); the Makefile compiles this file once per
// range (and once per dimension for the block-staged kernels) so the kernels build in
// parallel.
#define SG_EVAL_INSTANTIATE 1
#include "sg_eval.hpp"

namespace sg {
namespace eval {

#if SG_PREC == 64
using TT = double;
using AA = AccF64;
#else
using TT = float;
using AA = AccKahan;
#endif

// SG_BLK_ONLY: this object holds the block-staged kernels of the range (launch_blk),
// otherwise the other walks (launch_variant), so the heavy kernels build in parallel.
#ifdef SG_BLK_ONLY
#define SG_INST(DD) \
    template hipError_t launch_blk<DD, TT, AA>(int, const Params<TT>&, int64_t, hipStream_t); \
    template hipError_t launch_blk2<DD, TT, AA>(int, const Params<TT>&, int64_t, hipStream_t);
#else
#define SG_INST(DD)                                                                        \
    template hipError_t launch_variant<DD, false, TT, AA>(int, int, const Params<TT>&,     \
                                                          int64_t, int, hipStream_t);      \
    template hipError_t launch_variant<DD, true, TT, AA>(int, int, const Params<TT>&,      \
                                                         int64_t, int, hipStream_t);
#endif
#define SG_IN_RANGE(DD) (SG_DMIN <= DD && DD <= SG_DMAX)

#if SG_IN_RANGE(1)
SG_INST(1)
#endif
#if SG_IN_RANGE(2)
SG_INST(2)
#endif
#if SG_IN_RANGE(3)
SG_INST(3)
#endif
#if SG_IN_RANGE(4)
SG_INST(4)
#endif
#if SG_IN_RANGE(5)
SG_INST(5)
#endif
#if SG_IN_RANGE(6)
SG_INST(6)
#endif
#if SG_IN_RANGE(7)
SG_INST(7)
#endif
#if SG_IN_RANGE(8)
SG_INST(8)
#endif
#if SG_IN_RANGE(9)
SG_INST(9)
#endif
#if SG_IN_RANGE(10)
SG_INST(10)
#endif
#if SG_IN_RANGE(11)
SG_INST(11)
#endif
#if SG_IN_RANGE(12)
SG_INST(12)
#endif
#if SG_IN_RANGE(13)
SG_INST(13)
#endif
#if SG_IN_RANGE(14)
SG_INST(14)
#endif
#if SG_IN_RANGE(15)
SG_INST(15)
#endif
#if SG_IN_RANGE(16)
SG_INST(16)
#endif
#if SG_IN_RANGE(17)
SG_INST(17)
#endif
#if SG_IN_RANGE(18)
SG_INST(18)
#endif
#if SG_IN_RANGE(19)
SG_INST(19)
#endif
#if SG_IN_RANGE(20)
SG_INST(20)
#endif

}  // namespace eval
}  // namespace sg
