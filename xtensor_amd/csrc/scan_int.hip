// scan_int.hip — scan kernel instantiations for one accumulator family.
#include "scan_kernels.hpp"

namespace xg {

bool scan_kernels_int(int ak, int op, ScanKernels* k) {
#define XG_S(AK, OP) if (ak == AK && op == OP) { *k = scan_kernels<AK, OP>(); return true; }
    XG_S(AK_I32, RO_ADD)
    XG_S(AK_I32, RO_MUL)
    XG_S(AK_U32, RO_ADD)
    XG_S(AK_U32, RO_MUL)
    XG_S(AK_I64, RO_ADD)
    XG_S(AK_I64, RO_MUL)
    XG_S(AK_U64, RO_ADD)
    XG_S(AK_U64, RO_MUL)
#undef XG_S
    return false;
}

}  // namespace xg
