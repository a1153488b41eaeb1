// reduce_dyn_float.hip — generic-interpreter reduction kernels, one accumulator family.
#include "reduce_kernels.hpp"

namespace xg {

bool dyn_reduce_float(int ak, int op, RedKernels* k) {
#define XG_D(AK, OP) if (ak == AK && op == OP) { *k = dyn_kernels_for<AK, OP>(); return true; }
    XG_D(AK_F32C, RO_ADD)
    XG_D(AK_F64C, RO_ADD)
    XG_D(AK_F32, RO_MUL)
    XG_D(AK_F32, RO_MAX)
    XG_D(AK_F32, RO_MIN)
    XG_D(AK_F64, RO_MUL)
    XG_D(AK_F64, RO_MAX)
    XG_D(AK_F64, RO_MIN)
    XG_D(AK_F32, RO_MAXNUM)
    XG_D(AK_F32, RO_MINNUM)
    XG_D(AK_F64, RO_MAXNUM)
    XG_D(AK_F64, RO_MINNUM)
#undef XG_D
    return false;
}

}  // namespace xg
