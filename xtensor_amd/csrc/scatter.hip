// scatter.hip — assignment into an index view: target[offset_i] = value_i.
//
// xindex_view's assignment (views/xindex_view.hpp:786-797, through
// xview_semantic) writes e element by element into the positions the indices
// name.  Here the values and the offsets are first evaluated into contiguous
// temporaries (so an expression that reads the target sees it unchanged, as
// a temporary-based assignment does), then one kernel stores value i at
// offset i.  Offsets outside the target's extent are skipped; where two
// indices name the same element, which of their values lands is unspecified.
#include <hip/hip_runtime.h>

#include "ew.hpp"
#include "planner.hpp"
#include "xg_api_internal.hpp"

namespace xg {

namespace {

template <int ESZ>
__global__ __launch_bounds__(256) void scatter_kernel(char* base, const char* vals, const int64_t* offs, uint64_t n,
                                                      int64_t lo, int64_t hi) {
    using T = std::conditional_t<ESZ == 1, uint8_t, std::conditional_t<ESZ == 2, uint16_t,
                                 std::conditional_t<ESZ == 4, uint32_t, uint64_t>>>;
    const uint64_t stride = uint64_t(gridDim.x) * 256u;
    for (uint64_t i = uint64_t(blockIdx.x) * 256u + threadIdx.x; i < n; i += stride) {
        const int64_t o = offs[i];
        if (o >= lo && o <= hi) reinterpret_cast<T*>(base)[o] = reinterpret_cast<const T*>(vals)[i];
    }
}

// out (contiguous [n] of dtype dt) = e, broadcast to [n]
void eval_into(const xg_expr* e, void* data, int dt, int64_t n, hipStream_t st) {
    xg_tensor t{};
    t.data = data;
    t.dtype = dt;
    t.ndim = 1;
    t.shape[0] = n;
    t.strides[0] = 1;
    int V = 4;
    EwPlan p = plan_assign_best(e, &t, &V);
    if (p.units) check_hip(launch_ew(p, V, device_cus(), st, nullptr), "elementwise launch");
}

}  // namespace

}  // namespace xg

using namespace xg;

extern "C" xg_status xg_assign_indexed(const xg_expr* values, const xg_expr* offsets, const xg_tensor* target,
                                       xg_stream_t stream) {
    return guard([&] {
        if (!target) raise(XG_INVALID_ARGUMENT, "null target");
        if (target->ndim < 0 || target->ndim > XG_MAX_RANK || target->dtype < 0 || target->dtype >= XG_NUM_DTYPES)
            raise(XG_INVALID_ARGUMENT, "malformed target tensor");
        hipStream_t st = reinterpret_cast<hipStream_t>(stream);
        ExprInfo oi = analyse(offsets);
        if (oi.shape.size() != 1) raise(XG_INVALID_ARGUMENT, "offsets must be one-dimensional");
        if (oi.dtype == XG_BOOL || is_float(oi.dtype)) raise(XG_INVALID_ARGUMENT, "offsets must be integers");
        const int64_t n = oi.shape[0];
        ExprInfo vi = analyse(values);
        if (!broadcastable(vi.shape, oi.shape))
            raise(XG_BROADCAST_ERROR, "cannot broadcast " + shape_str(vi.shape) + " into " + shape_str(oi.shape));
        if (n == 0) return;
        const int esz = dtype_size(target->dtype);
        const size_t vbytes = (size_t(n) * size_t(esz) + 255) & ~size_t(255);
        char* ws = static_cast<char*>(workspace(vbytes + size_t(n) * 8, st));
        eval_into(values, ws, target->dtype, n, st);
        eval_into(offsets, ws + vbytes, XG_INT64, n, st);
        LeafView lv{};
        set_gather_leaf(*target, &lv);
        char* base = const_cast<char*>(lv.ptr);
        const int64_t* offs = reinterpret_cast<const int64_t*>(ws + vbytes);
        const uint64_t blocks = std::min<uint64_t>((uint64_t(n) + 255) / 256, 1u << 20);
        switch (esz) {
        case 1: hipLaunchKernelGGL(scatter_kernel<1>, dim3(uint32_t(blocks)), dim3(256), 0, st, base, ws, offs, uint64_t(n), lv.stride[0], lv.stride[1]); break;
        case 2: hipLaunchKernelGGL(scatter_kernel<2>, dim3(uint32_t(blocks)), dim3(256), 0, st, base, ws, offs, uint64_t(n), lv.stride[0], lv.stride[1]); break;
        case 4: hipLaunchKernelGGL(scatter_kernel<4>, dim3(uint32_t(blocks)), dim3(256), 0, st, base, ws, offs, uint64_t(n), lv.stride[0], lv.stride[1]); break;
        default: hipLaunchKernelGGL(scatter_kernel<8>, dim3(uint32_t(blocks)), dim3(256), 0, st, base, ws, offs, uint64_t(n), lv.stride[0], lv.stride[1]); break;
        }
        check_hip(hipGetLastError(), "scatter launch");
        note_kernel("scatter");
    });
}
