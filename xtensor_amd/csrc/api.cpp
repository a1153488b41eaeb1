// api.cpp — the extern "C" entry points declared in include/xg.h.
//
// Each entry point validates the descriptor, plans on the host and enqueues
// kernels on the caller's stream.  Errors become status codes with the
// meaning of the exceptions xtensor throws at the same point
// (broadcast_error, utils/xexception.hpp:63; std::runtime_error for axes,
// reducers/xreducer.hpp:1426-1440 and reducers/xaccumulator.hpp:231-234).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

#include <map>

#include "ew.hpp"
#include "jit.hpp"
#include "planner.hpp"
#include "reduce.hpp"
#include "scan.hpp"
#include "xg_api_internal.hpp"

namespace xg {

namespace {
thread_local std::string t_err;
thread_local std::string t_last_kernel;

struct DeviceInfo {
    int cus = 0;
    bool ok = false;
};

std::mutex g_dev_mu;
std::vector<DeviceInfo> g_devs;
}  // namespace

void set_error(const std::string& m) { t_err = m; }

int device_cus() {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) raise(XG_DEVICE_ERROR, "hipGetDevice failed");
    std::lock_guard<std::mutex> lk(g_dev_mu);
    if (int(g_devs.size()) <= dev) g_devs.resize(size_t(dev) + 1);
    DeviceInfo& d = g_devs[size_t(dev)];
    if (!d.ok) {
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, dev) != hipSuccess) raise(XG_DEVICE_ERROR, "hipGetDeviceProperties failed");
        d.cus = prop.multiProcessorCount;
        d.ok = true;
    }
    return d.cus;
}

void note_kernel(const char* name) { t_last_kernel = name ? name : ""; }

void check_hip(hipError_t e, const char* what) {
    if (e != hipSuccess) raise(XG_DEVICE_ERROR, std::string(what) + ": " + hipGetErrorString(e));
}

namespace {
// Workspace (partial results of split reductions, scan temporaries): one
// growing buffer per (device, stream), so work on one stream reuses it in
// stream order and different streams never share one.
struct WsKey {
    int dev;
    void* stream;
    bool operator<(const WsKey& o) const { return dev != o.dev ? dev < o.dev : stream < o.stream; }
};
struct WsBuf {
    void* ptr = nullptr;
    size_t bytes = 0;
};
std::mutex g_ws_mu;
std::map<WsKey, WsBuf> g_ws;
}  // namespace

void* workspace(size_t bytes, hipStream_t stream) {
    if (bytes == 0) return nullptr;
    int dev = 0;
    check_hip(hipGetDevice(&dev), "hipGetDevice");
    std::lock_guard<std::mutex> lk(g_ws_mu);
    WsBuf& b = g_ws[WsKey{dev, reinterpret_cast<void*>(stream)}];
    if (b.bytes < bytes) {
        if (b.ptr) {
            // earlier work on this stream may still read the old buffer
            check_hip(hipStreamSynchronize(stream), "hipStreamSynchronize");
            check_hip(hipFree(b.ptr), "hipFree");
            b.ptr = nullptr;
            b.bytes = 0;
        }
        size_t n = std::max<size_t>(bytes, size_t(1) << 20);
        if (hipMalloc(&b.ptr, n) != hipSuccess) raise(XG_OUT_OF_MEMORY, "workspace allocation failed");
        b.bytes = n;
    }
    return b.ptr;
}

}  // namespace xg

using namespace xg;

// plan with the compiled-kernel vector width; without a matching compiled
// program, re-plan for the generic interpreter (V = 1, registers in LDS)
EwPlan xg::plan_assign_best(const xg_expr* e, const xg_tensor* out, int* V) {
    *V = choose_vector_width(e, out);
    EwPlan p = plan_assign(e, out, *V);
    if (!ew_find_spec(p, *V) && !ew_jit_ok(p, *V)) {  // the generic interpreter, one element per item
        *V = 1;
        p = plan_assign(e, out, 1);
    }
    return p;
}

// plan with the compiled-kernel vector width; without a matching compiled
// program, re-plan for the generic kernels (V = 4)
RedPlan xg::plan_reduce_best(const xg_expr* e, const xg_reduce_desc* r, const xg_tensor* out, int* V) {
    const int cus = device_cus();
    *V = choose_vector_width(e, nullptr);
    RedPlan p = plan_reduce(e, r, out, *V, cus, nullptr, nullptr);
    if (*V != 4 && !find_rspec(p, *V) && !reduce_jit_ok(p, *V)) {
        *V = 4;
        p = plan_reduce(e, r, out, *V, cus, nullptr, nullptr);
    }
    return p;
}

namespace {

// xt::variance / xt::stddev (core/xmath.hpp:2031-2113): the inner
// mean<T>(e, axes) is evaluated first into a kept-dims temporary m, then
// mean<T>(square(e - m), axes, ddof) -- two passes over e, as in xtensor.
// m and the split partials share the stream's workspace.
void reduce_moment(const xg_expr* e, const xg_reduce_desc* r, const xg_tensor* out, hipStream_t st) {
    if (!e || !e->nodes || e->n_nodes < 1) raise(XG_INVALID_ARGUMENT, "empty expression");
    xg_reduce_desc r1 = *r;
    r1.op = XG_RED_MEAN;
    r1.keep_dims = 1;
    r1.has_initial = 0;
    r1.ddof = 0;
    int mdt = 0;
    Shape ms;
    reduce_result(e, &r1, &mdt, &ms);
    const int64_t nm = numel(ms);
    const size_t mbytes = (size_t(std::max<int64_t>(nm, 1)) * size_t(dtype_size(mdt)) + 255) & ~size_t(255);
    xg_tensor m{};
    m.dtype = mdt;
    m.ndim = int32_t(ms.size());
    int64_t stv = 1;
    for (int k = m.ndim - 1; k >= 0; --k) {
        m.shape[k] = ms[size_t(k)];
        m.strides[k] = stv;
        stv *= std::max<int64_t>(ms[size_t(k)], 1);
    }
    // e - m: the tree of e, m as one more leaf, and the subtraction as root
    std::vector<xg_node> nodes(e->nodes, e->nodes + e->n_nodes);
    std::vector<xg_tensor> leaves(e->leaves, e->leaves + e->n_leaves);
    leaves.push_back(m);
    xg_node ln{};
    ln.op = XG_OP_LEAF;
    ln.index = e->n_leaves;
    nodes.push_back(ln);
    xg_node sub{};
    sub.op = XG_OP_MINUS;
    sub.nargs = 2;
    sub.args[0] = e->n_nodes - 1;
    sub.args[1] = e->n_nodes;
    nodes.push_back(sub);
    xg_expr e2 = *e;
    e2.n_nodes = int32_t(nodes.size());
    e2.n_leaves = int32_t(leaves.size());
    e2.nodes = nodes.data();
    e2.leaves = leaves.data();
    // plan with a placeholder address to size the workspace, then for real
    auto plan_both = [&](char* base, RedPlan* p1, int* V1, RedPlan* p2, int* V2) {
        m.data = base;
        leaves.back().data = base;
        *p1 = plan_reduce_best(e, &r1, &m, V1);
        *p2 = plan_reduce_best(&e2, r, out, V2);
    };
    RedPlan p1, p2;
    int V1 = 4, V2 = 4;
    plan_both(reinterpret_cast<char*>(uintptr_t(1) << 40), &p1, &V1, &p2, &V2);
    if (p2.n_out == 0) {
        note_kernel("empty");
        return;
    }
    const size_t wbytes = std::max(reduce_workspace_bytes(p1), reduce_workspace_bytes(p2));
    char* base = static_cast<char*>(workspace(mbytes + wbytes, st));
    plan_both(base, &p1, &V1, &p2, &V2);
    const char* used = nullptr;
    if (p1.n_out) {
        p1.partials = wbytes ? base + mbytes : nullptr;
        check_hip(launch_reduce(p1, V1, st, &used), "variance mean launch");
    }
    p2.partials = wbytes ? base + mbytes : nullptr;
    check_hip(launch_reduce(p2, V2, st, &used), "variance launch");
    note_kernel(used);
}

// xt::minmax (core/xmath.hpp:2195-2230): out[0] = the std::min fold,
// out[1] = the std::max fold over every element, from max() / lowest().
void reduce_minmax(const xg_expr* e, const xg_reduce_desc* r, const xg_tensor* out, hipStream_t st) {
    int dt = 0;
    Shape rs;
    reduce_result(e, r, &dt, &rs);
    if (!out || out->ndim != 1 || out->shape[0] != 2)
        raise(XG_BROADCAST_ERROR, "minmax output must have shape (2)");
    const char* used = nullptr;
    for (int k = 0; k < 2; ++k) {
        xg_reduce_desc rk{};
        rk.op = k == 0 ? kRedMinNum : kRedMaxNum;
        rk.acc_dtype = -1;
        rk.n_axes = -1;
        xg_tensor ok = *out;
        ok.ndim = 0;
        ok.offset = out->offset + int64_t(k) * out->strides[0];
        int V = 4;
        RedPlan p = plan_reduce_best(e, &rk, &ok, &V);
        p.partials = workspace(reduce_workspace_bytes(p), st);
        check_hip(launch_reduce(p, V, st, &used), "minmax launch");
    }
    note_kernel(used);
}

}  // namespace

extern "C" {

const char* xg_version(void) { return "xtensor_amd 0.1.0 (gfx950)"; }

const char* xg_last_error_message(void) { return t_err.c_str(); }

// Name of the kernel family the calling thread's last successful entry point
// used ("dynamic", "spec:<name>", "reduce_rows", ...).  Diagnostic only.
const char* xg_last_kernel(void) { return t_last_kernel.c_str(); }

xg_status xg_init(void) {
    return guard([&] { device_cus(); });
}

xg_status xg_expr_info(const xg_expr* e, int32_t* dtype, int32_t* ndim, int64_t* shape) {
    return guard([&] {
        if (!dtype || !ndim || !shape) raise(XG_INVALID_ARGUMENT, "null output pointer");
        ExprInfo info = analyse(e);
        *dtype = info.dtype;
        *ndim = int32_t(info.shape.size());
        for (size_t k = 0; k < info.shape.size(); ++k) shape[k] = info.shape[k];
    });
}

namespace {
// An assignment whose offsets need 64 bits, cut along the output's first
// dimension into pieces whose offsets fit 32 bits, so that every piece runs
// a compiled (or run-time compiled) kernel instead of the generic
// interpreter.  Each piece slices the output and every leaf that spans that
// dimension; leaves broadcast along it (missing or of extent 1) stay whole.
// Returns false, planning nothing, when the expression reads through a
// gather (its source is addressed by value, not by position) or one index
// of the first dimension alone already needs 64-bit offsets.
using Piece = std::pair<EwPlan, int>;  // a plan and its vector width
bool plan_in_pieces(const xg_expr* e, const xg_tensor* out, std::vector<Piece>* pieces) {
    if (out->ndim < 1 || out->shape[0] < 2) return false;
    for (int i = 0; i < e->n_nodes; ++i)
        if (e->nodes[i].op == XG_OP_GATHER) return false;
    auto spans_dim0 = [&](const xg_tensor& t) {
        const int d = t.ndim - out->ndim;  // the leaf's axis aligned with the output's first
        return d >= 0 && t.shape[d] != 1 ? d : -1;
    };
    // element offsets of one index along the first dimension, and its step
    int64_t rest = 0, step0 = 0;
    auto account = [&](const xg_tensor& t, int d0) {
        int64_t r = 0;
        for (int k = d0 + 1; k < t.ndim; ++k) r += (std::max<int64_t>(t.shape[k], 1) - 1) * std::llabs(t.strides[k]);
        rest = std::max(rest, r);
        step0 = std::max<int64_t>(step0, std::llabs(t.strides[d0]));
    };
    account(*out, 0);
    for (int l = 0; l < e->n_leaves; ++l) {
        const int d = spans_dim0(e->leaves[l]);
        if (d >= 0) account(e->leaves[l], d);
    }
    constexpr int64_t kLim = (int64_t(1) << 31) - 1;
    if (rest >= kLim) return false;
    int64_t rows = step0 > 0 ? (kLim - rest) / step0 + 1 : out->shape[0];
    rows = std::min<int64_t>(rows, int64_t(1) << 30);  // and the inner extent stays below 2^31
    if (rows >= out->shape[0]) return false;
    std::vector<xg_tensor> leaves(e->leaves, e->leaves + e->n_leaves);
    xg_expr pe = *e;
    pe.leaves = leaves.data();
    for (int64_t s0 = 0; s0 < out->shape[0]; s0 += rows) {
        const int64_t len = std::min(rows, out->shape[0] - s0);
        xg_tensor o = *out;
        o.offset += s0 * out->strides[0];
        o.shape[0] = len;
        for (int l = 0; l < e->n_leaves; ++l) {
            leaves[size_t(l)] = e->leaves[l];
            const int d = spans_dim0(e->leaves[l]);
            if (d >= 0) {
                leaves[size_t(l)].offset += s0 * e->leaves[l].strides[d];
                leaves[size_t(l)].shape[d] = len;
            }
        }
        int V = 4;
        EwPlan p = plan_assign_best(&pe, &o, &V);
        if (p.units) pieces->emplace_back(p, V);
    }
    return true;
}
}  // namespace

xg_status xg_assign(const xg_expr* e, const xg_tensor* out, xg_stream_t stream) {
    return guard([&] {
        int V = 4;
        EwPlan p{};
        bool whole = true;  // false: the iteration space does not fit one plan's 32-bit unit indexing
        Error refused{};
        try {
            p = plan_assign_best(e, out, &V);
        } catch (const Error& err) {
            if (err.code != XG_UNSUPPORTED) throw;
            whole = false;
            refused = err;
        }
        const char* used = nullptr;
        std::vector<Piece> pieces;
        if ((!whole || p.index64) && plan_in_pieces(e, out, &pieces)) {
            for (const Piece& pc : pieces)
                check_hip(launch_ew(pc.first, pc.second, device_cus(), reinterpret_cast<hipStream_t>(stream), &used),
                          "elementwise launch");
            note_kernel(used ? used : "empty");
            return;
        }
        if (!whole) throw refused;
        if (p.units == 0) {
            note_kernel("empty");
            return;
        }
        check_hip(launch_ew(p, V, device_cus(), reinterpret_cast<hipStream_t>(stream), &used), "elementwise launch");
        note_kernel(used);
    });
}

xg_status xg_reduce_info(const xg_expr* e, const xg_reduce_desc* r, int32_t* dtype, int32_t* ndim, int64_t* shape) {
    return guard([&] {
        if (!dtype || !ndim || !shape) raise(XG_INVALID_ARGUMENT, "null output pointer");
        int dt = 0;
        Shape rs;
        reduce_result(e, r, &dt, &rs);
        *dtype = dt;
        *ndim = int32_t(rs.size());
        for (size_t k = 0; k < rs.size(); ++k) shape[k] = rs[k];
    });
}

namespace {
// sum / prod / amax / amin over a reduced dimension of 2^31 elements or more
// (beyond 32-bit indexing): that dimension is cut into pieces of 2^30, each
// piece reduces into its own slice of a temporary of out's shape (in the
// reduction's result type), and a second reduction of the same kind over
// the pieces gives out (with xt::initial applied there, once).  Returns false, doing nothing, for other
// reductions, gathers, or a second dimension that is itself too long.
bool reduce_in_pieces(const xg_expr* e, const xg_reduce_desc* r, const xg_tensor* out, hipStream_t st) {
    if (!r || !out || (r->op != XG_RED_SUM && r->op != XG_RED_PROD && r->op != XG_RED_AMAX && r->op != XG_RED_AMIN))
        return false;
    for (int i = 0; i < e->n_nodes; ++i)
        if (e->nodes[i].op == XG_OP_GATHER) return false;
    const ExprInfo info = analyse(e);
    const int R = int(info.shape.size());
    constexpr int64_t kBig = int64_t(1) << 31;
    std::vector<bool> red(size_t(R), r->n_axes < 0);
    for (int i = 0; i < r->n_axes; ++i) {
        const int a = r->axes[i] < 0 ? r->axes[i] + R : r->axes[i];
        if (a < 0 || a >= R) return false;
        red[size_t(a)] = true;
    }
    int d = -1;
    for (int k = 0; k < R; ++k) {
        if (info.shape[size_t(k)] < kBig) continue;
        if (!red[size_t(k)] || d >= 0) return false;
        d = k;
    }
    if (d < 0) return false;
    int64_t out_elems = 1;
    for (int k = 0; k < out->ndim; ++k) out_elems *= out->shape[k];
    if (out_elems <= 0) return false;
    const int64_t n = info.shape[size_t(d)], piece = int64_t(1) << 30, k_pieces = (n + piece - 1) / piece;
    // the pieces keep the reduction's own result type (out's may be narrower: cast on the final store)
    int rdt = 0;
    Shape rshape;
    reduce_result(e, r, &rdt, &rshape);
    const int esz = dtype_size(rdt);
    void* tmp = nullptr;
    if (hipMalloc(&tmp, size_t(k_pieces * out_elems * esz)) != hipSuccess)
        raise(XG_OUT_OF_MEMORY, "temporary for a reduction in pieces");
    try {
        std::vector<xg_tensor> leaves(e->leaves, e->leaves + e->n_leaves);
        xg_expr pe = *e;
        pe.leaves = leaves.data();
        xg_reduce_desc pr = *r;
        pr.has_initial = 0;
        xg_tensor o = *out;  // one slice of the temporary, contiguous, of out's shape
        o.data = tmp;
        o.dtype = rdt;
        for (int64_t k = o.ndim - 1, stride = 1; k >= 0; --k) {
            o.strides[k] = stride;
            stride *= o.shape[k];
        }
        const char* used = nullptr;
        for (int64_t i = 0; i < k_pieces; ++i) {
            const int64_t s0 = i * piece, len = std::min(piece, n - s0);
            for (int l = 0; l < e->n_leaves; ++l) {
                const xg_tensor& t = e->leaves[l];
                leaves[size_t(l)] = t;
                const int dl = d - (R - t.ndim);  // the leaf's axis aligned with d
                if (dl >= 0 && t.shape[dl] != 1) {
                    leaves[size_t(l)].offset += s0 * t.strides[dl];
                    leaves[size_t(l)].shape[dl] = len;
                }
            }
            o.offset = i * out_elems;
            int V = 4;
            RedPlan p = plan_reduce_best(&pe, &pr, &o, &V);
            p.partials = workspace(reduce_workspace_bytes(p), st);
            check_hip(launch_reduce(p, V, st, &used), "reduction launch");
        }
        // the pieces' results as [pieces, out shape...], reduced over axis 0 into out
        xg_tensor parts{};
        parts.data = tmp;
        parts.dtype = rdt;
        parts.ndim = out->ndim + 1;
        parts.shape[0] = k_pieces;
        parts.strides[0] = out_elems;
        for (int k = 0; k < out->ndim; ++k) {
            parts.shape[k + 1] = o.shape[k];
            parts.strides[k + 1] = o.strides[k];
        }
        xg_node leaf{};
        leaf.op = XG_OP_LEAF;
        xg_expr me{};
        me.n_nodes = 1;
        me.n_leaves = 1;
        me.nodes = &leaf;
        me.leaves = &parts;
        xg_reduce_desc mr = *r;
        mr.n_axes = 1;
        mr.axes[0] = 0;
        mr.keep_dims = 0;
        int V = 4;
        RedPlan p = plan_reduce_best(&me, &mr, out, &V);
        p.partials = workspace(reduce_workspace_bytes(p), st);
        check_hip(launch_reduce(p, V, st, &used), "reduction launch");
        note_kernel(used);
        check_hip(hipStreamSynchronize(st), "hipStreamSynchronize");
    } catch (...) {
        (void)hipStreamSynchronize(st);
        (void)hipFree(tmp);
        throw;
    }
    check_hip(hipFree(tmp), "hipFree");
    return true;
}
}  // namespace

xg_status xg_reduce(const xg_expr* e, const xg_reduce_desc* r, const xg_tensor* out, xg_stream_t stream) {
    return guard([&] {
        hipStream_t st = reinterpret_cast<hipStream_t>(stream);
        if (r && (r->op == XG_RED_VARIANCE || r->op == XG_RED_STDDEV)) {
            reduce_moment(e, r, out, st);
            return;
        }
        if (r && r->op == XG_RED_MINMAX) {
            reduce_minmax(e, r, out, st);
            return;
        }
        int V = 4;
        RedPlan p{};
        try {
            p = plan_reduce_best(e, r, out, &V);
        } catch (const Error& err) {
            if (err.code != XG_UNSUPPORTED || !reduce_in_pieces(e, r, out, st)) throw;
            return;
        }
        if (p.n_out == 0) {
            note_kernel("empty");
            return;
        }
        p.partials = workspace(reduce_workspace_bytes(p), st);
        const char* used = nullptr;
        check_hip(launch_reduce(p, V, st, &used), "reduction launch");
        note_kernel(used);
    });
}

xg_status xg_scan_info(const xg_expr* e, const xg_scan_desc* s, int32_t* dtype, int32_t* ndim, int64_t* shape) {
    return guard([&] {
        if (!dtype || !ndim || !shape) raise(XG_INVALID_ARGUMENT, "null output pointer");
        int dt = 0;
        Shape rs;
        scan_result(e, s, &dt, &rs);
        *dtype = dt;
        *ndim = int32_t(rs.size());
        for (size_t k = 0; k < rs.size(); ++k) shape[k] = rs[k];
    });
}

xg_status xg_scan(const xg_expr* e, const xg_scan_desc* s, const xg_tensor* out, xg_stream_t stream) {
    return guard([&] {
        if (!s || !out) raise(XG_INVALID_ARGUMENT, "null argument");
        hipStream_t st = reinterpret_cast<hipStream_t>(stream);
        int dt = 0;
        Shape rs;
        scan_result(e, s, &dt, &rs);
        const int V = 4;
        if (s->axis != XG_AXIS_NONE) {
            ScanPlan p = plan_scan(e, s, out, V);
            const char* used = nullptr;
            check_hip(launch_scan(p, V, st, &used), "scan launch");
            note_kernel(used);
            return;
        }
        // flattened scan: evaluate e row-major into a contiguous temporary of
        // the result type, then scan it as one row
        if (out->ndim != 1 || out->shape[0] != rs[0]) raise(XG_BROADCAST_ERROR, "flattened scan output must be 1-d");
        ExprInfo info = analyse(e);
        int64_t n = numel(info.shape);
        if (n == 0) return;
        xg_tensor tmp{};
        tmp.dtype = dt;
        tmp.ndim = int32_t(info.shape.size());
        int64_t stv = 1;
        for (int k = tmp.ndim - 1; k >= 0; --k) {
            tmp.shape[k] = info.shape[size_t(k)];
            tmp.strides[k] = stv;
            stv *= info.shape[size_t(k)];
        }
        tmp.data = workspace(size_t(n) * dtype_size(dt), st);
        check_hip(hipSuccess, "");
        {
            int Vw = 4;
            EwPlan ep = plan_assign_best(e, &tmp, &Vw);
            check_hip(launch_ew(ep, Vw, device_cus(), st, nullptr), "scan copy launch");
        }
        // 1-d view of the temporary as the only leaf
        xg_tensor flat{};
        flat.data = tmp.data;
        flat.dtype = dt;
        flat.ndim = 1;
        flat.shape[0] = n;
        flat.strides[0] = 1;
        xg_node node{};
        node.op = XG_OP_LEAF;
        xg_expr fe{};
        fe.n_nodes = 1;
        fe.n_leaves = 1;
        fe.nodes = &node;
        fe.leaves = &flat;
        xg_scan_desc s1 = *s;
        s1.axis = 0;
        ScanPlan p = plan_scan(&fe, &s1, out, V);
        const char* used = nullptr;
        check_hip(launch_scan(p, V, st, &used), "scan launch");
        note_kernel(used);
    });
}

// ---------------------------------------------------------------- plans
struct xg_plan_s {
    int kind = 0;  // 0 assign, 1 reduce, 2 scan
    int V = 4;
    xg::EwPlan ew;
    std::vector<std::pair<xg::EwPlan, int>> pieces;  // assign past 32-bit offsets: one plan per piece (ew unused)
    xg::RedPlan red;
    xg::ScanPlan scan;
    size_t ws = 0;
    std::string kernel;
};

xg_status xg_plan_assign(const xg_expr* e, const xg_tensor* out, xg_plan_t* plan) {
    return guard([&] {
        if (!plan) raise(XG_INVALID_ARGUMENT, "null plan pointer");
        auto* p = new xg_plan_s();
        p->kind = 0;
        try {
            bool whole = true;
            Error refused{};
            try {
                p->ew = plan_assign_best(e, out, &p->V);
            } catch (const Error& err) {
                if (err.code != XG_UNSUPPORTED) throw;
                whole = false;
                refused = err;
            }
            if ((!whole || p->ew.index64) && plan_in_pieces(e, out, &p->pieces)) p->ew.units = 0;
            else if (!whole) throw refused;
        } catch (...) {
            delete p;
            throw;
        }
        *plan = p;
    });
}

xg_status xg_plan_reduce(const xg_expr* e, const xg_reduce_desc* r, const xg_tensor* out, xg_plan_t* plan) {
    return guard([&] {
        if (!plan) raise(XG_INVALID_ARGUMENT, "null plan pointer");
        auto* p = new xg_plan_s();
        try {
            p->kind = 1;
            if (r && (r->op == XG_RED_VARIANCE || r->op == XG_RED_STDDEV || r->op == XG_RED_MINMAX))
                raise(XG_UNSUPPORTED, "variance, stddev and minmax run several passes; use xg_reduce");
            p->red = plan_reduce_best(e, r, out, &p->V);
            p->ws = reduce_workspace_bytes(p->red);
        } catch (...) {
            delete p;
            throw;
        }
        *plan = p;
    });
}

xg_status xg_plan_scan(const xg_expr* e, const xg_scan_desc* s, const xg_tensor* out, xg_plan_t* plan) {
    return guard([&] {
        if (!plan || !s) raise(XG_INVALID_ARGUMENT, "null argument");
        if (s->axis == XG_AXIS_NONE) raise(XG_UNSUPPORTED, "flattened scans are not planned; use xg_scan");
        auto* p = new xg_plan_s();
        try {
            p->kind = 2;
            p->V = 4;
            p->scan = plan_scan(e, s, out, p->V);
        } catch (...) {
            delete p;
            throw;
        }
        *plan = p;
    });
}

size_t xg_plan_workspace_size(xg_plan_t plan) { return plan ? plan->ws : 0; }

xg_status xg_plan_execute(xg_plan_t p, void* ws, xg_stream_t stream) {
    return guard([&] {
        if (!p) raise(XG_INVALID_ARGUMENT, "null plan");
        hipStream_t st = reinterpret_cast<hipStream_t>(stream);
        const char* used = nullptr;
        if (p->kind == 0) {
            if (p->ew.units) check_hip(launch_ew(p->ew, p->V, device_cus(), st, &used), "elementwise launch");
            for (const auto& pc : p->pieces) check_hip(launch_ew(pc.first, pc.second, device_cus(), st, &used), "elementwise launch");
        } else if (p->kind == 1) {
            if (p->red.n_out) {
                p->red.partials = p->ws ? (ws ? ws : workspace(p->ws, st)) : nullptr;
                check_hip(launch_reduce(p->red, p->V, st, &used), "reduction launch");
            }
        } else {
            check_hip(launch_scan(p->scan, p->V, st, &used), "scan launch");
        }
        note_kernel(used);
    });
}

xg_status xg_plan_execute_many(const xg_plan_t* plans, int64_t n, xg_stream_t stream) {
    return guard([&] {
        if (n < 0 || (n > 0 && !plans)) raise(XG_INVALID_ARGUMENT, "bad plan list");
        hipStream_t st = reinterpret_cast<hipStream_t>(stream);
        std::vector<const EwPlan*> ew;
        std::vector<int> V;
        for (int64_t i = 0; i < n; ++i) {
            if (!plans[i]) raise(XG_INVALID_ARGUMENT, "null plan in list");
            if (plans[i]->kind == 0) {
                ew.push_back(&plans[i]->ew);
                V.push_back(plans[i]->V);
                for (const auto& pc : plans[i]->pieces) {
                    ew.push_back(&pc.first);
                    V.push_back(pc.second);
                }
            }
        }
        launch_ew_many(ew.data(), V.data(), ew.size(), st);
        for (int64_t i = 0; i < n; ++i) {
            xg_plan_s* p = plans[i];
            const char* used = nullptr;
            if (p->kind == 1 && p->red.n_out) {
                p->red.partials = p->ws ? workspace(p->ws, st) : nullptr;
                check_hip(launch_reduce(p->red, p->V, st, &used), "reduction launch");
            } else if (p->kind == 2) {
                check_hip(launch_scan(p->scan, p->V, st, &used), "scan launch");
            }
        }
        note_kernel(ew.size() > 1 ? "multi" : "");
    });
}

void xg_plan_destroy(xg_plan_t plan) { delete plan; }

// Testing aid: 1 routes every assignment through the dynamic interpreter.
void xg_debug_force_dynamic(int on) { ew_set_force_dynamic(on); }

// Measurement aid: 1 or 4 work units per thread in compiled elementwise
// kernels, 0 = chosen by size.
void xg_debug_units_per_thread(int upt) { ew_set_units_per_thread(upt); }

// Measurement aid: non-temporal loads of large streaming leaves and
// non-temporal stores of large outputs (both on by default), for plans made
// after the call.
void xg_debug_nontemporal(int loads, int stores) { set_nontemporal(loads, stores); }

// Run-time compiled specialisations: 0 off, 1 for launches of at least
// XG_JIT_MIN_ELEMENTS elements (default), 2 for every compact plan.
void xg_debug_jit(int mode) { jit_set_mode(mode); }

// Measurement aid: waves per CU the reduction work split aims at (0: default).
void xg_debug_reduce_waves(int per_cu) { set_reduce_waves_per_cu(per_cu); }

// Counters of the run-time compiler: compiles, disk-cache hits, in-process
// hits, failures (out[0..3]); returns the last failure message.
const char* xg_debug_jit_stats(int64_t* out) {
    JitStats s = jit_stats();
    if (out) {
        out[0] = s.compiles;
        out[1] = s.disk_hits;
        out[2] = s.memory_hits;
        out[3] = s.failures;
    }
    return jit_last_error();
}

// Name of the compiled specialisation an assignment would use, "" if none.
xg_status xg_debug_assign_kernel(const xg_expr* e, const xg_tensor* out, char* name, int32_t cap) {
    return guard([&] {
        int V = 4;
        EwPlan p = plan_assign_best(e, out, &V);
        const char* s = ew_find_spec(p, V);
        std::string n = s ? s : "";
        if (cap > 0) {
            std::strncpy(name, n.c_str(), size_t(cap - 1));
            name[cap - 1] = 0;
        }
    });
}

// Testing aid (host only, no device access): how an assignment past 32-bit
// offsets is cut.  *n_pieces = the number of pieces (0: not cut), *all_compact
// = 1 if every piece fits 32-bit offsets, *max_units = the largest piece's
// work units.  Planned with compiled and run-time compiled programs off.
xg_status xg_debug_assign_pieces(const xg_expr* e, const xg_tensor* out, int64_t* n_pieces, int32_t* all_compact,
                                 int64_t* max_units) {
    return guard([&] {
        if (!n_pieces || !all_compact || !max_units) raise(XG_INVALID_ARGUMENT, "null output pointer");
        ew_set_force_dynamic(1);
        std::vector<Piece> pieces;
        bool ok = false;
        try {
            ok = plan_in_pieces(e, out, &pieces);
        } catch (...) {
            ew_set_force_dynamic(0);
            throw;
        }
        ew_set_force_dynamic(0);
        *n_pieces = ok ? int64_t(pieces.size()) : 0;
        *all_compact = 1;
        *max_units = 0;
        for (const Piece& pc : pieces) {
            if (!pc.first.compact) *all_compact = 0;
            *max_units = std::max<int64_t>(*max_units, int64_t(pc.first.units));
        }
    });
}

}  // extern "C"
