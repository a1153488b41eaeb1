// reduce_combine.hip — split-combine kernel instantiations.
#include "reduce_kernels.hpp"

namespace xg {

KFn combine_kernel(int ak, int op, bool wide) {
#define XG_C(AK, OP) \
    if (ak == AK && op == OP) return wide ? reduce_combine_wide_kernel<AK, OP> : reduce_combine_kernel<AK, OP>;
    XG_C(AK_F32C, RO_ADD)
    XG_C(AK_F64C, RO_ADD)
    XG_C(AK_F32, RO_MUL)
    XG_C(AK_F32, RO_MAX)
    XG_C(AK_F32, RO_MIN)
    XG_C(AK_F64, RO_MUL)
    XG_C(AK_F64, RO_MAX)
    XG_C(AK_F64, RO_MIN)
    XG_C(AK_F32, RO_MAXNUM)
    XG_C(AK_F32, RO_MINNUM)
    XG_C(AK_F64, RO_MAXNUM)
    XG_C(AK_F64, RO_MINNUM)
    XG_C(AK_I32, RO_ADD)
    XG_C(AK_I32, RO_MUL)
    XG_C(AK_I32, RO_MAX)
    XG_C(AK_I32, RO_MIN)
    XG_C(AK_U32, RO_ADD)
    XG_C(AK_U32, RO_MUL)
    XG_C(AK_U32, RO_MAX)
    XG_C(AK_U32, RO_MIN)
    XG_C(AK_I64, RO_ADD)
    XG_C(AK_I64, RO_MUL)
    XG_C(AK_I64, RO_MAX)
    XG_C(AK_I64, RO_MIN)
    XG_C(AK_U64, RO_ADD)
    XG_C(AK_U64, RO_MUL)
    XG_C(AK_U64, RO_MAX)
    XG_C(AK_U64, RO_MIN)
#undef XG_C
    return nullptr;
}

}  // namespace xg
