// reduce_dyn_int64.hip — generic-interpreter reduction kernels, one accumulator family.
#include "reduce_kernels.hpp"

namespace xg {

bool dyn_reduce_int64(int ak, int op, RedKernels* k) {
#define XG_D(AK, OP) if (ak == AK && op == OP) { *k = dyn_kernels_for<AK, OP>(); return true; }
    XG_D(AK_I64, RO_ADD)
    XG_D(AK_I64, RO_MUL)
    XG_D(AK_I64, RO_MAX)
    XG_D(AK_I64, RO_MIN)
    XG_D(AK_U64, RO_ADD)
    XG_D(AK_U64, RO_MUL)
    XG_D(AK_U64, RO_MAX)
    XG_D(AK_U64, RO_MIN)
#undef XG_D
    return false;
}

}  // namespace xg
