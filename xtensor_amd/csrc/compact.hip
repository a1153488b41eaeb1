// compact.hip — xg_argwhere: the multi-indices of the nonzero elements of an
// expression, listed on the device (stream compaction).
//
// xt::argwhere<L> (core/xoperation.hpp:893-930) walks the expression in
// traversal order L and pushes the index of every element that converts to
// true.  Here: flags = (bool)e as int64, evaluated in that order into a
// contiguous temporary; an inclusive scan of the flags (the scan kernels,
// in place) gives every true element its slot k = prefix - 1; one kernel
// writes the element's multi-index into row k of the output.  For
// column-major traversal the expression is read through its reversed axes,
// so that its row-major order is the column-major order of the original.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "ew.hpp"
#include "planner.hpp"
#include "scan.hpp"
#include "xg_api_internal.hpp"

namespace xg {

namespace {

struct Shape8 {
    int64_t v[XG_MAX_RANK];
};

// Row k = prefix[i] - 1 of `out` receives the index of every position i
// whose flag is set (prefix[i] > prefix[i - 1]); rows past `rows` are
// dropped.  The index unravels i over `shape` (row-major); `reverse` writes
// its components in reverse order (column-major traversal).
__global__ __launch_bounds__(256) void argwhere_kernel(const int64_t* prefix, uint64_t n, int nd, Shape8 shape,
                                                       int reverse, char* out, int64_t rows, int64_t s_row,
                                                       int64_t s_col) {
    const uint64_t stride = uint64_t(gridDim.x) * 256u;
    for (uint64_t i = uint64_t(blockIdx.x) * 256u + threadIdx.x; i < n; i += stride) {
        const int64_t prev = i ? prefix[i - 1] : 0;
        const int64_t k = prev;
        if (prefix[i] == prev || k >= rows) continue;
        uint64_t q = i;
        int64_t* dst = reinterpret_cast<int64_t*>(out) + k * s_row;
        for (int d = nd - 1; d >= 0; --d) {
            const uint64_t ext = uint64_t(shape.v[d]);
            const uint64_t r = q % ext;
            q /= ext;
            dst[int64_t(reverse ? nd - 1 - d : d) * s_col] = int64_t(r);
        }
    }
}

}  // namespace

}  // namespace xg

using namespace xg;

extern "C" xg_status xg_argwhere(const xg_expr* cond, int32_t column_major, const xg_tensor* out,
                                 xg_stream_t stream) {
    return guard([&] {
        if (!cond || !cond->nodes || cond->n_nodes < 1) raise(XG_INVALID_ARGUMENT, "empty expression");
        if (!out || out->dtype != XG_INT64 || out->ndim != 2) raise(XG_INVALID_ARGUMENT, "output must be int64 [count, ndim]");
        hipStream_t st = reinterpret_cast<hipStream_t>(stream);
        const ExprInfo info = analyse(cond);
        const int nd = int(info.shape.size());
        if (nd < 1) raise(XG_INVALID_ARGUMENT, "argwhere of a 0-d expression");
        if (out->shape[1] != nd) raise(XG_BROADCAST_ERROR, "output rows must have one column per dimension");
        const int64_t n = numel(info.shape);
        if (n == 0 || out->shape[0] == 0) return;
        // the traversal shape, and the expression read through reversed axes
        // for column-major order (leaves padded to the full rank first, so
        // that broadcasting stays right-aligned)
        Shape trav = info.shape;
        std::vector<xg_tensor> leaves(cond->leaves, cond->leaves + cond->n_leaves);
        if (column_major) {
            std::reverse(trav.begin(), trav.end());
            std::vector<bool> gathered(size_t(cond->n_leaves), false);  // addressed by value, not position
            for (int i = 0; i < cond->n_nodes; ++i)
                if (cond->nodes[i].op == XG_OP_GATHER && cond->nodes[i].index >= 0 && cond->nodes[i].index < cond->n_leaves)
                    gathered[size_t(cond->nodes[i].index)] = true;
            for (int l = 0; l < cond->n_leaves; ++l) {
                if (gathered[size_t(l)]) continue;
                xg_tensor& t = leaves[size_t(l)];
                const int pad = nd - t.ndim;
                if (pad < 0) raise(XG_BROADCAST_ERROR, "leaf of higher rank than the expression");
                xg_tensor r = t;
                r.ndim = nd;
                for (int k = 0; k < nd; ++k) {
                    const int src = nd - 1 - k - pad;  // the leaf axis that lands on reversed axis k
                    r.shape[k] = src >= 0 ? t.shape[src] : 1;
                    r.strides[k] = src >= 0 ? t.strides[src] : 0;
                }
                t = r;
            }
        }
        // flags = int64((bool)cond): the tree, then two casts as new roots
        std::vector<xg_node> nodes(cond->nodes, cond->nodes + cond->n_nodes);
        xg_node to_bool{};
        to_bool.op = XG_OP_CAST;
        to_bool.nargs = 1;
        to_bool.args[0] = cond->n_nodes - 1;
        to_bool.index = XG_BOOL;
        nodes.push_back(to_bool);
        xg_node to_i64 = to_bool;
        to_i64.args[0] = int32_t(nodes.size()) - 1;
        to_i64.index = XG_INT64;
        nodes.push_back(to_i64);
        xg_expr fe = *cond;
        fe.n_nodes = int32_t(nodes.size());
        fe.nodes = nodes.data();
        fe.leaves = leaves.data();
        // the flags in traversal order, scanned in place as one row
        const size_t flag_bytes = (size_t(n) * 8 + 255) & ~size_t(255);
        xg_tensor flat{};
        flat.dtype = XG_INT64;
        flat.ndim = 1;
        flat.shape[0] = n;
        flat.strides[0] = 1;
        flat.data = reinterpret_cast<void*>(uintptr_t(1) << 40);  // sizing: an aligned placeholder
        xg_node leaf{};
        leaf.op = XG_OP_LEAF;
        xg_expr se{};
        se.n_nodes = 1;
        se.n_leaves = 1;
        se.nodes = &leaf;
        se.leaves = &flat;
        xg_scan_desc sd{};
        sd.op = XG_SCAN_CUMSUM;
        sd.acc_dtype = -1;
        sd.axis = 0;
        const int cus = device_cus();
        ScanPlan sp = plan_scan(&se, &sd, &flat, 4, cus);
        const size_t wsb = scan_workspace_bytes(sp);
        char* block = static_cast<char*>(workspace(flag_bytes + wsb, st));
        flat.data = block;
        xg_tensor grid = flat;  // the flags laid out row-major in the traversal shape
        grid.ndim = nd;
        for (int64_t k = nd - 1, stv = 1; k >= 0; --k) {
            grid.shape[k] = trav[size_t(k)];
            grid.strides[k] = stv;
            stv *= std::max<int64_t>(trav[size_t(k)], 1);
        }
        int V = 4;
        EwPlan ep = plan_assign_best(&fe, &grid, &V);
        check_hip(launch_ew(ep, V, cus, st, nullptr), "flags launch");
        sp = plan_scan(&se, &sd, &flat, 4, cus);
        if (scan_workspace_bytes(sp) > wsb) raise(XG_INVALID_ARGUMENT, "internal: scan workspace grew after sizing");
        check_hip(launch_scan(sp, 4, wsb ? block + flag_bytes : nullptr, st, nullptr), "scan launch");
        Shape8 sh{};
        for (int k = 0; k < nd; ++k) sh.v[k] = trav[size_t(k)];
        const int osz = 8;
        char* obase = static_cast<char*>(out->data) + out->offset * osz;
        const uint64_t blocks = std::min<uint64_t>((uint64_t(n) + 255) / 256, 1u << 20);
        hipLaunchKernelGGL(argwhere_kernel, dim3(uint32_t(blocks)), dim3(256), 0, st,
                           reinterpret_cast<const int64_t*>(block), uint64_t(n), nd, sh, column_major ? 1 : 0,
                           obase, out->shape[0], out->strides[0], out->shape[1] == 1 ? 0 : out->strides[1]);
        check_hip(hipGetLastError(), "argwhere launch");
        note_kernel("argwhere");
    });
}
