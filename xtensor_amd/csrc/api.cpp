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

#include "assign_reduce_plan.hpp"
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

namespace {
std::map<WsKey, uint32_t*> g_counters;
}  // namespace

// Arrival counters of the column reductions that merge their own splits:
// one zeroed block of kReduceCounters words per (device, stream), cleared
// once when allocated (every launch leaves them zero again).
uint32_t* reduce_counters(hipStream_t stream) {
    int dev = 0;
    check_hip(hipGetDevice(&dev), "hipGetDevice");
    std::lock_guard<std::mutex> lk(g_ws_mu);
    uint32_t*& c = g_counters[WsKey{dev, reinterpret_cast<void*>(stream)}];
    if (!c) {
        void* m = nullptr;
        if (hipMalloc(&m, size_t(kReduceCounters) * sizeof(uint32_t)) != hipSuccess)
            raise(XG_OUT_OF_MEMORY, "counter allocation failed");
        check_hip(hipMemsetAsync(m, 0, size_t(kReduceCounters) * sizeof(uint32_t), stream), "hipMemsetAsync");
        c = static_cast<uint32_t*>(m);
    }
    return c;
}

// base address of a sizing pass's block: 256-byte aligned like a real one,
// so that plans made against it choose the same vector accesses
constexpr uintptr_t kSizingBase = uintptr_t(1) << 40;

char* Arena::take(size_t bytes) {
    if (bytes == 0) return nullptr;
    const size_t off = (used + 255) & ~size_t(255);
    used = off + bytes;
    if (measure) return reinterpret_cast<char*>(kSizingBase) + off;
    if (used > cap) raise(XG_INVALID_ARGUMENT, "internal: a call took more device memory than its sizing pass");
    return base + off;
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
    // rows per piece: offsets stay below 2^31, and so does every merged
    // dimension, since a piece holds at most max(2^30, one row) elements
    // (contiguous rows merge into one dimension of the piece's element count)
    int64_t row_elems = 1;
    for (int k = 1; k < out->ndim; ++k) row_elems *= std::max<int64_t>(out->shape[k], 1);
    if (row_elems >= kLim) return false;
    int64_t rows = step0 > 0 ? (kLim - rest) / step0 + 1 : out->shape[0];
    rows = std::min<int64_t>(rows, std::max<int64_t>(1, (int64_t(1) << 30) / row_elems));
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
// out = e: one plan, or pieces along the first dimension when the offsets
// need 64 bits or the iteration space does not fit one plan.
void assign_launch(const xg_expr* e, const xg_tensor* out, hipStream_t st, const char** used) {
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
    std::vector<Piece> pieces;
    if ((!whole || p.index64) && plan_in_pieces(e, out, &pieces)) {
        for (const Piece& pc : pieces)
            check_hip(launch_ew(pc.first, pc.second, device_cus(), st, used), "elementwise launch");
        return;
    }
    if (!whole) throw refused;
    if (p.units == 0) return;
    check_hip(launch_ew(p, V, device_cus(), st, used), "elementwise launch");
}

// Launch a planned reduction with its split partials from the arena.
void run_reduce(RedPlan& p, int V, hipStream_t st, Arena& a, const char** used) {
    p.partials = a.take(reduce_workspace_bytes(p));
    p.counters = a.counters;
    if (!a.measure && p.n_out) check_hip(launch_reduce(p, V, st, used), "reduction launch");
}

// A reduction whose reduced extent does not fit 32-bit indexing (a reduced
// dimension of 2^31 elements or more, or reduced dimensions that merge into
// one that long), cut into pieces.  The innermost reduced axis at which the
// product of the reduced extents from there inward passes 2^30 is cut into
// chunks, and every reduced axis outside it is taken one index at a time,
// so no piece reduces more than 2^30 elements per output.  Each piece folds
// into its own slice of a temporary of the output's shape, in the folded
// value type (a mean's sum, a norm_l2's sum of squares, a variance's sum of
// squared deviations), and a second reduction over the pieces merges them
// and finishes the result once: xt::initial, the divisor of the whole
// reduction for mean / variance / stddev, the square root of norm_l2 and
// stddev, the store conversion.
struct ReducePieces {
    bool ok = false;  // false: no reduced extent is too long, or a gather
    int piece_op = 0, merge_op = 0, merge_acc = -1;
    int R = 0, j = -1, d = -1;  // ax[j] = d is the axis cut into chunks
    std::vector<int> ax;
    Shape shape;
    int64_t len_max = 0, chunks = 0, k_pieces = 0, out_elems = 0;
    xg_reduce_desc pr{};           // the pieces' reduction
    xg_tensor o{};                 // a piece's slice of the temporary
    std::vector<xg_tensor> leaves;
    xg_expr pe{};                  // the expression with the sliced leaves
    const xg_expr* e = nullptr;

    ReducePieces(const xg_expr* e_, const xg_reduce_desc* r, const xg_tensor* out, int vdt) : e(e_) {
        if (!r || !out) return;
        piece_op = merge_op = r->op;
        merge_acc = r->acc_dtype;
        switch (r->op) {
        case XG_RED_SUM: case XG_RED_PROD: case XG_RED_AMAX: case XG_RED_AMIN: break;
        case kRedMinNum: case kRedMaxNum: merge_acc = -1; break;
        case XG_RED_MEAN: piece_op = XG_RED_SUM; merge_op = XG_RED_SUM; break;
        case XG_RED_VARIANCE: case XG_RED_STDDEV: piece_op = XG_RED_VARIANCE; merge_op = XG_RED_SUM; break;
        case XG_RED_NORM_SQ: case XG_RED_NORM_L1: case XG_RED_NORM_L0: merge_op = XG_RED_SUM; merge_acc = -1; break;
        case XG_RED_NORM_L2: piece_op = XG_RED_NORM_SQ; merge_op = XG_RED_SUM; merge_acc = -1; break;
        case XG_RED_NORM_LINF: merge_op = XG_RED_AMAX; merge_acc = -1; break;
        default: return;
        }
        for (int i = 0; i < e->n_nodes; ++i)
            if (e->nodes[i].op == XG_OP_GATHER) return;
        shape = analyse(e).shape;
        R = int(shape.size());
        if (r->n_axes < 0) {  // the reduced axes, ascending (plan_reduce has checked them)
            for (int k = 0; k < R; ++k) ax.push_back(k);
        } else {
            for (int i = 0; i < r->n_axes; ++i) ax.push_back(r->axes[i] < 0 ? r->axes[i] + R : r->axes[i]);
        }
        constexpr int64_t kPiece = int64_t(1) << 30;
        constexpr int64_t kMaxPieces = 4096;
        int64_t inner = 1;  // product of the reduced extents inside ax[j]
        for (int i = int(ax.size()) - 1; i >= 0; --i) {
            const int64_t n = shape[size_t(ax[size_t(i)])];
            if (n > kPiece / inner) {
                j = i;
                break;
            }
            inner *= n;
        }
        if (j < 0) return;
        d = ax[size_t(j)];
        len_max = std::max<int64_t>(1, kPiece / inner);
        chunks = (shape[size_t(d)] + len_max - 1) / len_max;
        k_pieces = chunks;
        for (int i = 0; i < j && k_pieces <= kMaxPieces; ++i) k_pieces *= shape[size_t(ax[size_t(i)])];
        if (k_pieces > kMaxPieces) raise(XG_UNSUPPORTED, "reduction too large: it would need more than 4096 pieces");
        out_elems = 1;
        for (int k = 0; k < out->ndim; ++k) out_elems *= out->shape[k];
        if (out_elems <= 0) return;
        pr = *r;
        pr.op = piece_op;
        pr.has_initial = 0;
        pr.ddof = 0;
        o = *out;
        o.dtype = vdt;
        for (int64_t k = o.ndim - 1, stride = 1; k >= 0; --k) {
            o.strides[k] = stride;
            stride *= o.shape[k];
        }
        leaves.assign(e->leaves, e->leaves + e->n_leaves);
        pe = *e;
        pe.leaves = leaves.data();
        ok = true;
    }

    // slice the leaves and the temporary for piece q
    void select(int64_t q, char* tmp) {
        int64_t rest = q / chunks;
        const int64_t s0 = (q % chunks) * len_max, len = std::min(len_max, shape[size_t(d)] - s0);
        std::vector<int64_t> idx(size_t(j), 0);  // indices of the reduced axes outside ax[j]
        for (int i = j - 1; i >= 0; --i) {
            const int64_t ext = shape[size_t(ax[size_t(i)])];
            idx[size_t(i)] = rest % ext;
            rest /= ext;
        }
        for (int l = 0; l < e->n_leaves; ++l) {
            const xg_tensor& t = e->leaves[l];
            xg_tensor& pl = leaves[size_t(l)];
            pl = t;
            auto cut = [&](int axis, int64_t start, int64_t extent) {
                const int dl = axis - (R - t.ndim);  // the leaf's axis aligned with `axis`
                if (dl >= 0 && t.shape[dl] != 1) {
                    pl.offset += start * t.strides[dl];
                    pl.shape[dl] = extent;
                }
            };
            for (int i = 0; i < j; ++i) cut(ax[size_t(i)], idx[size_t(i)], 1);
            cut(d, s0, len);
        }
        o.data = tmp;
        o.offset = q * out_elems;
    }
};

// Evaluate a reduction in pieces (see ReducePieces).  Returns false, doing
// nothing, when no cut applies.
bool reduce_in_pieces(const xg_expr* e, const xg_reduce_desc* r, const xg_tensor* out, hipStream_t st, Arena& a,
                      const char** used) {
    if (!r || r->op == XG_RED_MINMAX) return false;
    const RedFinish fin = reduce_finish(e, r);
    const int vdt = fin.vdt;
    ReducePieces P(e, r, out, vdt);
    if (!P.ok) return false;
    char* tmp = a.take(size_t(P.k_pieces) * size_t(P.out_elems) * size_t(dtype_size(vdt)));
    auto plan_piece = [&](int64_t q, int* V) {
        P.select(q, tmp);
        RedPlan p = plan_reduce_best(&P.pe, &P.pr, &P.o, V);
        if (p.vdt != vdt) raise(XG_UNSUPPORTED, "reduction in pieces: value types of the pieces differ");
        p.fin = FZ_NONE;  // the folded value only; the merge finishes
        p.rdt = p.vdt;
        return p;
    };
    // the merge: the pieces as [pieces, out shape...] reduced over axis 0
    xg_tensor parts{};
    parts.data = tmp;
    parts.dtype = vdt;
    parts.ndim = out->ndim + 1;
    parts.shape[0] = P.k_pieces;
    parts.strides[0] = P.out_elems;
    for (int k = 0; k < out->ndim; ++k) {
        parts.shape[k + 1] = P.o.shape[k];
        parts.strides[k + 1] = P.o.strides[k];
    }
    xg_node leaf{};
    leaf.op = XG_OP_LEAF;
    xg_expr me{};
    me.n_nodes = 1;
    me.n_leaves = 1;
    me.nodes = &leaf;
    me.leaves = &parts;
    xg_reduce_desc mr{};
    mr.op = P.merge_op;
    mr.acc_dtype = P.merge_acc;
    mr.n_axes = 1;
    mr.axes[0] = 0;
    mr.has_initial = r->has_initial;
    mr.initial = r->initial;
    int Vm = 4;
    RedPlan pm = plan_reduce_best(&me, &mr, out, &Vm);
    if (pm.vdt != vdt) raise(XG_UNSUPPORTED, "reduction in pieces: the merge's value type differs");
    pm.fin = fin.fin;
    pm.mdt = fin.mdt;
    pm.div_bits = fin.div_bits;
    pm.rdt = fin.rdt;
    // split partials: one region that every piece and the merge reuse in stream order
    size_t pbytes = reduce_workspace_bytes(pm);
    int V = 4;
    for (int64_t q = 0; q < P.k_pieces; ++q) pbytes = std::max(pbytes, reduce_workspace_bytes(plan_piece(q, &V)));
    char* part = a.take(pbytes);
    if (a.measure) return true;
    for (int64_t q = 0; q < P.k_pieces; ++q) {
        RedPlan p = plan_piece(q, &V);
        p.partials = reduce_workspace_bytes(p) ? part : nullptr;
        p.counters = a.counters;
        check_hip(launch_reduce(p, V, st, used), "reduction launch");
    }
    pm.partials = reduce_workspace_bytes(pm) ? part : nullptr;
    pm.counters = a.counters;
    check_hip(launch_reduce(pm, Vm, st, used), "reduction launch");
    return true;
}

// A reduction plan_reduce handles (every op but minmax; variance / stddev of
// the deviation expression), in pieces when its reduced extent is too long.
void reduce_plain(const xg_expr* e, const xg_reduce_desc* r, const xg_tensor* out, hipStream_t st, Arena& a,
                  const char** used) {
    int V = 4;
    RedPlan p{};
    try {
        p = plan_reduce_best(e, r, out, &V);
    } catch (const Error& err) {
        if (err.code != XG_UNSUPPORTED || !reduce_in_pieces(e, r, out, st, a, used)) throw;
        return;
    }
    if (p.n_out == 0) return;
    run_reduce(p, V, st, a, used);
}

// xt::variance / xt::stddev (core/xmath.hpp:2031-2113): the inner
// mean<T>(e, axes) is evaluated first into a kept-dims temporary m, then
// mean<T>(square(e - m), axes, ddof) -- two passes over e, as in xtensor.
void reduce_moment(const xg_expr* e, const xg_reduce_desc* r, const xg_tensor* out, hipStream_t st, Arena& a,
                   const char** used) {
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
    xg_tensor m{};
    m.dtype = mdt;
    m.ndim = int32_t(ms.size());
    int64_t stv = 1;
    for (int k = m.ndim - 1; k >= 0; --k) {
        m.shape[k] = ms[size_t(k)];
        m.strides[k] = stv;
        stv *= std::max<int64_t>(ms[size_t(k)], 1);
    }
    m.data = a.take(size_t(std::max<int64_t>(nm, 1)) * size_t(dtype_size(mdt)));
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
    reduce_plain(e, &r1, &m, st, a, used);
    reduce_plain(&e2, r, out, st, a, used);
}

// xt::minmax (core/xmath.hpp:2195-2230): out[0] = the std::min fold,
// out[1] = the std::max fold over every element, from max() / lowest().
void reduce_minmax(const xg_expr* e, const xg_reduce_desc* r, const xg_tensor* out, hipStream_t st, Arena& a,
                   const char** used) {
    int dt = 0;
    Shape rs;
    reduce_result(e, r, &dt, &rs);
    if (!out || out->ndim != 1 || out->shape[0] != 2)
        raise(XG_BROADCAST_ERROR, "minmax output must have shape (2)");
    for (int k = 0; k < 2; ++k) {
        xg_reduce_desc rk{};
        rk.op = k == 0 ? kRedMinNum : kRedMaxNum;
        rk.acc_dtype = -1;
        rk.n_axes = -1;
        xg_tensor ok = *out;
        ok.ndim = 0;
        ok.offset = out->offset + int64_t(k) * out->strides[0];
        reduce_plain(e, &rk, &ok, st, a, used);
    }
}

void reduce_any(const xg_expr* e, const xg_reduce_desc* r, const xg_tensor* out, hipStream_t st, Arena& a,
                const char** used) {
    if (r && (r->op == XG_RED_VARIANCE || r->op == XG_RED_STDDEV)) reduce_moment(e, r, out, st, a, used);
    else if (r && r->op == XG_RED_MINMAX) reduce_minmax(e, r, out, st, a, used);
    else reduce_plain(e, r, out, st, a, used);
}

char* stream_workspace(size_t bytes, hipStream_t st, void*) { return static_cast<char*>(workspace(bytes, st)); }

// out = e with reductions of its rows (axis 1) and / or columns (axis 0) in
// the same pass (assign_reduce.hpp), planned: the kernel argument, the
// kernel compiled at run time, and the column merge.
struct ArcPlan {
    AssignRC arg{};
    const JitKernel* k = nullptr;
    uint32_t bands = 0;
    bool cols = false;
    RedPlan comb{};     // the column merge (a split combine over the bands)
    size_t ws = 0;      // workspace bytes: the column partials
};

// Plan the single pass; false, planning nothing, when it does not apply: a
// 2-D row-major contiguous output of at least 2 x 2, a compact plan whose
// value type is the output's and whose rows are whole work units (a plan that
// merged the output to one dimension needs C % V == 0), rows of at most 512
// units (8 a lane; fewer for elements under 4 bytes), reductions of the plain stored values (no mean, variance or square
// root, no conversion of an element before it is folded) over exactly that
// axis, and the kernel compiled at run time.
bool plan_assign_rc(const xg_expr* e, const xg_tensor* out, const xg_reduce_desc* rr, const xg_tensor* ro,
                    const xg_reduce_desc* cr, const xg_tensor* co, ArcPlan* ap) {
    if (out->ndim != 2 || out->strides[1] != 1 || out->strides[0] != out->shape[1]) return false;
    const int64_t R = out->shape[0], C = out->shape[1];
    if (R < 2 || C < 2 || R >= (int64_t(1) << 31) || C >= (int64_t(1) << 31)) return false;
    if (!jit_wanted(uint64_t(R) * uint64_t(C))) return false;
    int V = 4;
    EwPlan p = plan_assign_best(e, out, &V);
    if (!p.compact || p.index64 || p.units == 0 || p.prog.result_dtype != out->dtype) return false;
    uint32_t upr = 0;
    if (p.rank == 2 && p.shape[0] == R && p.shape[1] == C) upr = p.nchunks;
    else if (p.rank == 1 && C % V == 0) upr = uint32_t(C / V);
    else return false;
    // work units per lane: a wave holds a row (at most 8 units a lane, and
    // the waves' column accumulators, 64 * UL * V of 16 bytes four times, in
    // 128 KiB of LDS)
    const uint32_t need = (upr + 63) / 64;
    const int UL = need <= 1 ? 1 : (need <= 2 ? 2 : (need <= 4 ? 4 : 8));
    if (need > 8 || UL * V > 32) return false;
    // the reductions, planned over the output as a leaf (what xg_reduce of
    // the stored values would run)
    xg_node leaf{};
    leaf.op = XG_OP_LEAF;
    xg_tensor ol = *out;
    xg_expr le{};
    le.n_nodes = 1;
    le.n_leaves = 1;
    le.nodes = &leaf;
    le.leaves = &ol;
    const int cus = device_cus();
    auto plain = [&](const xg_reduce_desc* d, int axis, RedPlan* q, const xg_tensor* o) {
        if (d->n_axes != 1 || (d->axes[0] != axis && d->axes[0] != axis - 2)) return false;
        if (d->op == XG_RED_MEAN || d->op == XG_RED_VARIANCE || d->op == XG_RED_STDDEV || d->op == XG_RED_MINMAX ||
            d->op == XG_RED_NORM_L2)
            return false;
        *q = plan_reduce(&le, d, o, 4, cus, nullptr, nullptr);
        return q->fin == FZ_NONE && q->vdt == out->dtype && q->prog.nops == 1 && q->prog.ops[0].op == MO_LOAD &&
               q->n_out == uint32_t(axis == 1 ? R : C);
    };
    RedPlan rp{}, cp{};
    if (rr && !plain(rr, 1, &rp, ro)) return false;
    if (cr && !plain(cr, 0, &cp, co)) return false;
    const int rak = rr ? rp.acc : -1, rop = rr ? rp.op : 0, cak = cr ? cp.acc : -1, cop = cr ? cp.op : 0;
    const JitKernel* k = jit_assign_rc(p, V, rak, rop, cak, cop, UL);
    if (!k) return false;
    // bands of rows (a workgroup of eight waves each): about two per CU,
    // column partials at most a quarter of the output's bytes
    const uint64_t esz = uint64_t(dtype_size(out->dtype));
    const uint64_t max_bands = std::max<uint64_t>(1, uint64_t(R) * esz / 64);
    uint64_t rpb = std::max<uint64_t>((uint64_t(R) + 2 * cus - 1) / (2 * cus), (uint64_t(R) + max_bands - 1) / max_bands);
    rpb = std::max<uint64_t>(rpb, 1);
    ap->bands = uint32_t((uint64_t(R) + rpb - 1) / rpb);
    ap->k = k;
    ap->arg.plan = p;
    FusedRC& f = ap->arg.f;
    f.rows = uint32_t(R);
    f.cols = uint32_t(C);
    f.upr = upr;
    f.rows_per_band = uint32_t(rpb);
    if (rr) {
        f.row_out = static_cast<char*>(ro->data) + ro->offset * dtype_size(ro->dtype);
        f.row_stride = ro->ndim >= 1 ? ro->strides[0] : 0;
        f.row_dtype = ro->dtype;
        f.row_vdt = rp.vdt;
        f.row_rdt = rp.rdt;
        f.row_has_initial = rp.has_initial;
        f.row_identity = rp.identity;
        f.row_initial = rp.initial;
    }
    ap->cols = cr != nullptr;
    if (cr) {
        f.col_identity = cp.identity;
        ap->ws = size_t(ap->bands) * size_t(C) * sizeof(Store);
        RedPlan& q = ap->comb;
        q = cp;
        q.splits = int(ap->bands);
        q.div_splits = make_fastdiv(ap->bands);
        q.split_major = 1;
        q.counters = nullptr;
    }
    return true;
}

// Enqueue a planned single pass; `ws` holds ap.ws bytes.
void launch_assign_rc(ArcPlan& ap, void* ws, hipStream_t st) {
    ap.arg.f.col_partials = static_cast<Store*>(ws);
    check_hip(jit_launch(ap.k, ap.bands, 512, &ap.arg, sizeof(AssignRC), st), "fused assignment launch");
    if (ap.cols) {
        ap.comb.partials = ws;
        check_hip(launch_reduce_combine(ap.comb, st), "column combine launch");
    }
}

}  // namespace

namespace {
// xt::evaluation_strategy::immediate (reducers/xreducer.hpp:944-952):
// reduce_immediate(f, eval(e), ...).  An expression is evaluated into a
// row-major temporary of its value type first; a lone container is reduced
// as it is (xt::eval returns it unchanged).
void reduce_top(const xg_expr* e, const xg_reduce_desc* r, const xg_tensor* out, hipStream_t st, Arena& a,
                const char** used) {
    if (!r) raise(XG_INVALID_ARGUMENT, "null reduce descriptor");
    if (r->flags & ~XG_REDUCE_IMMEDIATE) raise(XG_INVALID_ARGUMENT, "unknown reduction flags");
    const bool lone_leaf = e && e->n_nodes == 1 && e->nodes && e->nodes[0].op == XG_OP_LEAF;
    if (!(r->flags & XG_REDUCE_IMMEDIATE) || lone_leaf) {
        reduce_any(e, r, out, st, a, used);
        return;
    }
    const ExprInfo info = analyse(e);
    xg_tensor t{};
    t.dtype = info.dtype;
    t.ndim = int32_t(info.shape.size());
    int64_t stv = 1;
    for (int k = t.ndim - 1; k >= 0; --k) {
        t.shape[k] = info.shape[size_t(k)];
        t.strides[k] = stv;
        stv *= std::max<int64_t>(info.shape[size_t(k)], 1);
    }
    const int64_t n = numel(info.shape);
    t.data = a.take(size_t(std::max<int64_t>(n, 0)) * size_t(dtype_size(t.dtype)));
    if (!a.measure && n > 0) assign_launch(e, &t, st, used);
    xg_node leaf{};
    leaf.op = XG_OP_LEAF;
    xg_expr te{};
    te.n_nodes = 1;
    te.n_leaves = 1;
    te.nodes = &leaf;
    te.leaves = &t;
    reduce_any(&te, r, out, st, a, used);
}
}  // namespace

void xg::reduce_call(const xg_expr* e, const xg_reduce_desc* r, const xg_tensor* out, hipStream_t st, BlockSource src,
                     void* ctx, bool counters) {
    const char* used = nullptr;
    Arena sizing;
    reduce_top(e, r, out, st, sizing, &used);
    Arena a;
    a.measure = false;
    a.cap = sizing.used;
    a.base = sizing.used ? src(sizing.used, st, ctx) : nullptr;
    a.counters = counters ? reduce_counters(st) : nullptr;
    used = nullptr;
    reduce_top(e, r, out, st, a, &used);
    note_kernel(used ? used : "empty");
}

extern "C" {

const char* xg_version(void) { return "xtensor_amd 0.1.0 (gfx950)"; }

const char* xg_last_error_message(void) { return t_err.c_str(); }

// Name of the kernel family the calling thread's last successful entry point
// used ("dynamic", "spec:<name>", "reduce_rows", ...).  Diagnostic only.
const char* xg_last_kernel(void) { return t_last_kernel.c_str(); }

xg_status xg_init(void) {
    return guard([&] {
        device_cus();
        scratch_setup();
    });
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

xg_status xg_assign(const xg_expr* e, const xg_tensor* out, xg_stream_t stream) {
    return guard([&] {
        const char* used = nullptr;
        assign_launch(e, out, reinterpret_cast<hipStream_t>(stream), &used);
        note_kernel(used ? used : "empty");
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

xg_status xg_reduce(const xg_expr* e, const xg_reduce_desc* r, const xg_tensor* out, xg_stream_t stream) {
    return guard([&] { reduce_call(e, r, out, reinterpret_cast<hipStream_t>(stream), stream_workspace, nullptr, true); });
}

xg_status xg_assign_reduce2d(const xg_expr* e, const xg_tensor* out, const xg_reduce_desc* row_red,
                             const xg_tensor* row_out, const xg_reduce_desc* col_red, const xg_tensor* col_out,
                             xg_stream_t stream) {
    return guard([&] {
        if (!e || !out) raise(XG_INVALID_ARGUMENT, "null argument");
        if (!row_red != !row_out || !col_red != !col_out)
            raise(XG_INVALID_ARGUMENT, "a reduction needs both its descriptor and its output");
        hipStream_t st = reinterpret_cast<hipStream_t>(stream);
        const char* used = nullptr;
        ArcPlan ap;
        if (plan_assign_rc(e, out, row_red, row_out, col_red, col_out, &ap)) {
            launch_assign_rc(ap, ap.ws ? workspace(ap.ws, st) : nullptr, st);
            note_kernel("assign_rc");
            return;
        }
        // the same as three calls: the assignment, then each reduction of
        // the stored output
        assign_launch(e, out, st, &used);
        xg_node leaf{};
        leaf.op = XG_OP_LEAF;
        xg_tensor ol = *out;
        xg_expr le{};
        le.n_nodes = 1;
        le.n_leaves = 1;
        le.nodes = &leaf;
        le.leaves = &ol;
        if (row_red) reduce_call(&le, row_red, row_out, st, stream_workspace, nullptr, true);
        if (col_red) reduce_call(&le, col_red, col_out, st, stream_workspace, nullptr, true);
        note_kernel("assign_then_reduce");
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

// A scan along an axis of 2^31 elements or more, cut along that axis into
// pieces of 2^30: each piece is scanned into its slice of the output, and
// from the second piece on the previous piece's last slice is folded in,
// out_q = out_q + last_{q-1} (cumprod: *), one elementwise launch with that
// slice broadcast along the axis, in the output's type (for integers the
// same residues as one sequential scan; for floats one more rounding per
// element past the first piece).
struct ScanPieces {
    const xg_expr* e;
    const xg_scan_desc* s;
    const xg_tensor* out;
    int R = 0, d = -1;
    int64_t L = 0;
    std::vector<xg_tensor> leaves;
    xg_expr pe{};
    static constexpr int64_t kPiece = int64_t(1) << 30;

    // false: no piece applies (the axis is shorter, or a gather)
    bool init(const xg_expr* e_, const xg_scan_desc* s_, const xg_tensor* out_) {
        e = e_;
        s = s_;
        out = out_;
        for (int i = 0; i < e->n_nodes; ++i)
            if (e->nodes[i].op == XG_OP_GATHER) return false;
        const Shape shape = analyse(e).shape;
        R = int(shape.size());
        d = s->axis < 0 ? s->axis + R : s->axis;
        if (d < 0 || d >= R || out->ndim != R) return false;
        L = shape[size_t(d)];
        if (L < (int64_t(1) << 31)) return false;
        leaves.assign(e->leaves, e->leaves + e->n_leaves);
        pe = *e;
        pe.leaves = leaves.data();
        return true;
    }
    // the piece starting at s0: leaves and output sliced along d
    xg_tensor select(int64_t s0) {
        const int64_t len = std::min(kPiece, L - s0);
        for (int l = 0; l < e->n_leaves; ++l) {
            const xg_tensor& t = e->leaves[l];
            xg_tensor& pl = leaves[size_t(l)];
            pl = t;
            const int dl = d - (R - t.ndim);  // the leaf's axis aligned with d
            if (dl >= 0 && t.shape[dl] != 1) {
                pl.offset += s0 * t.strides[dl];
                pl.shape[dl] = len;
            }
        }
        xg_tensor oq = *out;
        oq.offset += s0 * out->strides[d];
        oq.shape[d] = len;
        return oq;
    }
    // the largest scan workspace of a piece (plans only, no launch)
    size_t workspace_bytes(int cus) {
        size_t m = 0;
        for (int64_t s0 = 0; s0 < L; s0 += kPiece) {
            xg_tensor oq = select(s0);
            m = std::max(m, scan_workspace_bytes(plan_scan(&pe, s, &oq, 4, cus)));
        }
        return m;
    }
    void run(char* ws, size_t ws_bytes, hipStream_t st, const char** used) {
        const int cus = device_cus();
        for (int64_t s0 = 0; s0 < L; s0 += kPiece) {
            xg_tensor oq = select(s0);
            ScanPlan p = plan_scan(&pe, s, &oq, 4, cus);
            if (scan_workspace_bytes(p) > ws_bytes) raise(XG_INVALID_ARGUMENT, "internal: scan workspace grew after sizing");
            check_hip(launch_scan(p, 4, ws, st, used), "scan launch");
            if (s0 == 0) continue;
            xg_tensor prev = *out;  // the previous piece's last slice, extent 1 along d
            prev.offset += (s0 - 1) * out->strides[d];
            prev.shape[d] = 1;
            xg_tensor cl[2] = {oq, prev};
            xg_node nodes[3]{};
            nodes[0].op = XG_OP_LEAF;
            nodes[0].index = 0;
            nodes[1].op = XG_OP_LEAF;
            nodes[1].index = 1;
            nodes[2].op = s->op == XG_SCAN_CUMPROD ? XG_OP_MULTIPLIES : XG_OP_PLUS;
            nodes[2].nargs = 2;
            nodes[2].args[0] = 0;
            nodes[2].args[1] = 1;
            xg_expr ce{};
            ce.n_nodes = 3;
            ce.n_leaves = 2;
            ce.nodes = nodes;
            ce.leaves = cl;
            assign_launch(&ce, &oq, st, nullptr);
        }
    }
};

// A scan that would run the generic kernels (the column kernel, a thread per
// V columns walking a strided axis, or the rows kernel interpreting the
// program per element: an expression, or a leaf of another type or layout)
// over many elements: evaluate e row-major into a contiguous temporary of the
// result type first, as accumulator_impl copies before it scans
// (reducers/xaccumulator.hpp:240-297), and scan the temporary with the tiled,
// direct or long-row kernels.  False (nothing launched) when that does not apply; with
// `launch` false, only whether it applies (planning, no device work).
constexpr int64_t kScanEvalFirstMin = int64_t(1) << 20;  // elements
bool scan_eval_first(const xg_expr* e, const xg_scan_desc* s, const xg_tensor* out, const ScanPlan& p,
                     hipStream_t st, bool launch = true) {
    if (p.tiled != ST_NONE || p.lookback || (p.layout == SL_ROWS && scan_direct_ok(p))) return false;
    ExprInfo info = analyse(e);
    const int64_t n = numel(info.shape);
    if (n < kScanEvalFirstMin) return false;
    xg_tensor tmp{};
    tmp.dtype = p.rdt;
    tmp.ndim = int32_t(info.shape.size());
    int64_t stv = 1;
    for (int k = tmp.ndim - 1; k >= 0; --k) {
        tmp.shape[k] = info.shape[size_t(k)];
        tmp.strides[k] = stv;
        stv *= info.shape[size_t(k)];
    }
    tmp.data = reinterpret_cast<void*>(kSizingBase);  // sizing pass: an aligned placeholder
    xg_node node{};
    node.op = XG_OP_LEAF;
    xg_expr te{};
    te.n_nodes = 1;
    te.n_leaves = 1;
    te.nodes = &node;
    te.leaves = &tmp;
    const int V = 4, cus = device_cus();
    ScanPlan q = plan_scan(&te, s, out, V, cus);
    const bool fast = q.tiled != ST_NONE || q.lookback || (q.layout == SL_ROWS && scan_direct_ok(q));
    if (!fast || q.rdt != p.rdt || q.acc != p.acc) return false;
    if (!launch) return true;
    const size_t tmp_bytes = (size_t(n) * size_t(dtype_size(p.rdt)) + 255) & ~size_t(255);
    const size_t wsb = scan_workspace_bytes(q);
    char* block = static_cast<char*>(workspace(tmp_bytes + wsb, st));
    tmp.data = block;
    int Vw = 4;
    EwPlan ep = plan_assign_best(e, &tmp, &Vw);
    check_hip(launch_ew(ep, Vw, cus, st, nullptr), "scan copy launch");
    q = plan_scan(&te, s, out, V, cus);
    if (scan_workspace_bytes(q) > wsb) raise(XG_INVALID_ARGUMENT, "internal: scan workspace grew after sizing");
    const char* used = nullptr;
    check_hip(launch_scan(q, V, wsb ? block + tmp_bytes : nullptr, st, &used), "scan launch");
    note_kernel(used);
    return true;
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
            ScanPlan p;
            const char* used = nullptr;
            try {
                p = plan_scan(e, s, out, V, device_cus());
            } catch (const Error& err) {
                ScanPieces sp;
                if (err.code != XG_UNSUPPORTED || !sp.init(e, s, out)) throw;
                const size_t wsb = sp.workspace_bytes(device_cus());
                sp.run(wsb ? static_cast<char*>(workspace(wsb, st)) : nullptr, wsb, st, &used);
                note_kernel(used);
                return;
            }
            if (scan_eval_first(e, s, out, p, st)) return;
            check_hip(launch_scan(p, V, workspace(scan_workspace_bytes(p), st), st, &used), "scan launch");
            note_kernel(used);
            return;
        }
        // flattened scan (accumulator_impl without an axis): a lone leaf that
        // is contiguous in row-major order is scanned where it lies; any other
        // expression is evaluated row-major into a contiguous temporary of the
        // result type first, as xtensor copies it before scanning
        if (out->ndim != 1 || out->shape[0] != rs[0]) raise(XG_BROADCAST_ERROR, "flattened scan output must be 1-d");
        ExprInfo info = analyse(e);
        int64_t n = numel(info.shape);
        if (n == 0) return;
        const xg_tensor* lone = (e->n_nodes == 1 && e->nodes[0].op == XG_OP_LEAF && e->leaves[e->nodes[0].index].dtype == dt)
                                    ? &e->leaves[e->nodes[0].index]
                                    : nullptr;
        if (lone) {
            int64_t expect = 1;
            for (int k = lone->ndim - 1; k >= 0 && lone; --k) {
                if (lone->shape[k] != 1 && lone->strides[k] != expect) lone = nullptr;
                else expect *= lone->shape[k];
            }
        }
        xg_tensor flat{};
        flat.ndim = 1;
        flat.shape[0] = n;
        flat.strides[0] = 1;
        size_t tmp_bytes = 0;
        if (lone) {
            flat.data = lone->data;
            flat.dtype = lone->dtype;
            flat.offset = lone->offset;
        } else {
            flat.dtype = dt;
            tmp_bytes = (size_t(n) * size_t(dtype_size(dt)) + 255) & ~size_t(255);
            flat.data = reinterpret_cast<void*>(kSizingBase);  // sizing pass: an aligned placeholder
        }
        xg_node node{};
        node.op = XG_OP_LEAF;
        xg_expr fe{};
        fe.n_nodes = 1;
        fe.n_leaves = 1;
        fe.nodes = &node;
        fe.leaves = &flat;
        xg_scan_desc s1 = *s;
        s1.axis = 0;
        const int cus = device_cus();
        if (n >= (int64_t(1) << 31)) {  // past 32-bit indexing: the flat row in pieces
            ScanPieces sp;
            if (!sp.init(&fe, &s1, out)) raise(XG_UNSUPPORTED, "flattened scan too large");
            const size_t wsb = sp.workspace_bytes(cus);  // sized on the placeholder address
            char* block = tmp_bytes + wsb ? static_cast<char*>(workspace(tmp_bytes + wsb, st)) : nullptr;
            if (!lone) {
                xg_tensor tmp{};
                tmp.data = block;
                tmp.dtype = dt;
                tmp.ndim = int32_t(info.shape.size());
                for (int64_t k = tmp.ndim - 1, stv = 1; k >= 0; --k) {
                    tmp.shape[k] = info.shape[size_t(k)];
                    tmp.strides[k] = stv;
                    stv *= info.shape[size_t(k)];
                }
                assign_launch(e, &tmp, st, nullptr);
                flat.data = block;
            }
            const char* used = nullptr;
            sp.run(wsb ? block + tmp_bytes : nullptr, wsb, st, &used);
            note_kernel(used);
            return;
        }
        ScanPlan p = plan_scan(&fe, &s1, out, V, cus);
        const size_t wsb = scan_workspace_bytes(p);
        char* block = tmp_bytes + wsb ? static_cast<char*>(workspace(tmp_bytes + wsb, st)) : nullptr;
        if (!lone) {
            xg_tensor tmp{};
            tmp.data = block;
            tmp.dtype = dt;
            tmp.ndim = int32_t(info.shape.size());
            int64_t stv = 1;
            for (int k = tmp.ndim - 1; k >= 0; --k) {
                tmp.shape[k] = info.shape[size_t(k)];
                tmp.strides[k] = stv;
                stv *= info.shape[size_t(k)];
            }
            int Vw = 4;
            EwPlan ep = plan_assign_best(e, &tmp, &Vw);
            check_hip(launch_ew(ep, Vw, cus, st, nullptr), "scan copy launch");
            flat.data = block;
            p = plan_scan(&fe, &s1, out, V, cus);
            if (scan_workspace_bytes(p) > wsb) raise(XG_INVALID_ARGUMENT, "internal: scan workspace grew after sizing");
        }
        const char* used = nullptr;
        check_hip(launch_scan(p, V, wsb ? block + tmp_bytes : nullptr, st, &used), "scan launch");
        note_kernel(used);
    });
}

// ---------------------------------------------------------------- plans
struct xg_plan_s {
    // 0 assign, 1 reduce, 2 scan, 3 assign with row / column reductions;
    // 4 / 5: a reduction / scan the one-launch plans refuse (past 32-bit
    // indexing, several passes, evaluated first), kept as its descriptors and
    // run as xg_reduce / xg_scan at each execution
    int kind = 0;
    int V = 4;
    ArcPlan arc;
    std::vector<xg_node> d_nodes;
    std::vector<xg_tensor> d_leaves;
    std::vector<xg_scalar> d_consts;
    xg_tensor d_out{};
    xg_reduce_desc d_red{};
    xg_scan_desc d_scan{};
    void keep_call(const xg_expr* e, const xg_tensor* out) {
        d_nodes.assign(e->nodes, e->nodes + e->n_nodes);
        d_leaves.assign(e->leaves, e->leaves + e->n_leaves);
        if (e->n_consts > 0) d_consts.assign(e->consts, e->consts + e->n_consts);
        d_out = *out;
    }
    xg_expr call_expr() const {
        xg_expr e{};
        e.n_nodes = int32_t(d_nodes.size());
        e.n_leaves = int32_t(d_leaves.size());
        e.n_consts = int32_t(d_consts.size());
        e.nodes = d_nodes.data();
        e.leaves = d_leaves.data();
        e.consts = d_consts.data();
        return e;
    }
    void run_call(hipStream_t st) const {
        const xg_expr e = call_expr();
        const xg_status r = kind == 4 ? xg_reduce(&e, &d_red, &d_out, reinterpret_cast<xg_stream_t>(st))
                                      : xg_scan(&e, &d_scan, &d_out, reinterpret_cast<xg_stream_t>(st));
        if (r != XG_OK) raise(r, xg_last_error_message());
    }
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
            bool deferred = r && (r->op == XG_RED_VARIANCE || r->op == XG_RED_STDDEV || r->op == XG_RED_MINMAX ||
                                  r->flags != 0);
            if (!deferred) {
                try {
                    p->red = plan_reduce_best(e, r, out, &p->V);
                    p->ws = reduce_workspace_bytes(p->red);
                } catch (const Error& err) {
                    if (err.code != XG_UNSUPPORTED) throw;
                    deferred = true;  // past 32-bit indexing: xg_reduce cuts it into pieces
                }
            }
            if (deferred) {
                if (!r || !e || !out) raise(XG_INVALID_ARGUMENT, "null argument");
                int dt = 0;
                Shape rs;
                reduce_result(e, r, &dt, &rs);  // checks the axes and types now
                p->kind = 4;
                p->keep_call(e, out);
                p->d_red = *r;
                p->ws = 0;
            }
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
            try {
                p->scan = plan_scan(e, s, out, p->V, device_cus());
                p->ws = scan_workspace_bytes(p->scan);
                if (scan_eval_first(e, s, out, p->scan, nullptr, false)) {
                    p->kind = 5;  // evaluated first at each execution (xg_scan)
                    p->keep_call(e, out);
                    p->d_scan = *s;
                    p->ws = 0;
                }
            } catch (const Error& err) {
                if (err.code != XG_UNSUPPORTED) throw;
                p->kind = 5;  // past 32-bit indexing: xg_scan scans it in pieces
                p->keep_call(e, out);
                p->d_scan = *s;
                p->ws = 0;
            }
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
                p->red.counters = ws ? nullptr : reduce_counters(st);
                check_hip(launch_reduce(p->red, p->V, st, &used), "reduction launch");
            }
        } else if (p->kind == 3) {
            launch_assign_rc(p->arc, p->ws ? (ws ? ws : workspace(p->ws, st)) : nullptr, st);
            used = "assign_rc";
        } else if (p->kind >= 4) {
            p->run_call(st);
            used = p->kind == 4 ? "reduce_call" : "scan_call";
        } else {
            void* w = p->ws ? (ws ? ws : workspace(p->ws, st)) : nullptr;
            check_hip(launch_scan(p->scan, p->V, w, st, &used), "scan launch");
        }
        note_kernel(used);
    });
}

xg_status xg_plan_assign_reduce2d(const xg_expr* e, const xg_tensor* out, const xg_reduce_desc* row_red,
                                  const xg_tensor* row_out, const xg_reduce_desc* col_red, const xg_tensor* col_out,
                                  xg_plan_t* plan) {
    return guard([&] {
        if (!plan || !e || !out) raise(XG_INVALID_ARGUMENT, "null argument");
        if (!row_red != !row_out || !col_red != !col_out)
            raise(XG_INVALID_ARGUMENT, "a reduction needs both its descriptor and its output");
        auto* p = new xg_plan_s();
        p->kind = 3;
        try {
            if (!plan_assign_rc(e, out, row_red, row_out, col_red, col_out, &p->arc))
                raise(XG_UNSUPPORTED, "the single pass does not apply to this assignment (plan the three calls)");
        } catch (...) {
            delete p;
            throw;
        }
        p->ws = p->arc.ws;
        *plan = p;
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
                p->red.counters = reduce_counters(st);
                check_hip(launch_reduce(p->red, p->V, st, &used), "reduction launch");
            } else if (p->kind == 3) {
                launch_assign_rc(p->arc, p->ws ? workspace(p->ws, st) : nullptr, st);
            } else if (p->kind >= 4) {
                p->run_call(st);
            } else if (p->kind == 2) {
                check_hip(launch_scan(p->scan, p->V, p->ws ? workspace(p->ws, st) : nullptr, st, &used),
                          "scan launch");
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
// Measurement aid: 0 turns off the granule-fill reads of partly written output granules.
void xg_debug_fill_granules(int on) { set_fill_granules(on); }

// Run-time compiled specialisations: 0 off, 1 for launches of at least
// XG_JIT_MIN_ELEMENTS elements (default), 2 for every compact plan.
void xg_debug_jit(int mode) { jit_set_mode(mode); }

// Measurement aid: waves per CU the reduction work split aims at (0: default).
void xg_debug_reduce_waves(int per_cu) { set_reduce_waves_per_cu(per_cu); }
// Measurement aid: column reductions' workgroups per CU and partials budget multiplier.
void xg_debug_cols_split(int blocks_per_cu, int partial_mult) { set_cols_split(blocks_per_cu, partial_mult); }

// Measurement aid: workgroups per CU the tiled scans aim at (0: default).
void xg_debug_scan_blocks_per_cu(int per_cu) { set_scan_blocks_per_cu(per_cu); }
// Measurement aid: 0 keeps strided-axis scans on the three-launch tiled form.
void xg_debug_scan_single_pass(int on) { set_scan_single_pass(on); }
// Measurement aid: 1 runs few long rows as the one-pass look-back scan (default 0: three launches).
void xg_debug_scan_flat_single_pass(int on) { set_scan_flat_single_pass(on); }

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
        // this thread only, restored afterwards: other threads keep planning as they were
        const int prev = ew_set_thread_force_dynamic(1);
        std::vector<Piece> pieces;
        bool ok = false;
        try {
            ok = plan_in_pieces(e, out, &pieces);
        } catch (...) {
            ew_set_thread_force_dynamic(prev);
            throw;
        }
        ew_set_thread_force_dynamic(prev);
        *n_pieces = ok ? int64_t(pieces.size()) : 0;
        *all_compact = 1;
        *max_units = 0;
        for (const Piece& pc : pieces) {
            if (!pc.first.compact) *all_compact = 0;
            *max_units = std::max<int64_t>(*max_units, int64_t(pc.first.units));
        }
    });
}

// Testing aid (host only, no device access): how a reduction whose reduced
// extent is too long for one plan is cut.  *n_pieces = the number of pieces
// (0: one plan fits, not cut), *axis = the axis cut into chunks, *piece_len =
// its chunk length, *all_fit = 1 if every piece and the merge plan without
// needing 64-bit dimensions.  Planned with the generic kernels (no run-time
// compiles) and an assumed 256 compute units.
xg_status xg_debug_reduce_pieces(const xg_expr* e, const xg_reduce_desc* r, const xg_tensor* out, int64_t* n_pieces,
                                 int32_t* axis, int64_t* piece_len, int32_t* all_fit) {
    return guard([&] {
        if (!n_pieces || !axis || !piece_len || !all_fit) raise(XG_INVALID_ARGUMENT, "null output pointer");
        *n_pieces = 0;
        *axis = -1;
        *piece_len = 0;
        *all_fit = 1;
        constexpr int kCus = 256;
        try {
            plan_reduce(e, r, out, 4, kCus, nullptr, nullptr);
            return;
        } catch (const Error& err) {
            if (err.code != XG_UNSUPPORTED) throw;
        }
        const RedFinish fin = reduce_finish(e, r);
        ReducePieces P(e, r, out, fin.vdt);
        if (!P.ok) {
            *all_fit = 0;
            return;
        }
        *n_pieces = P.k_pieces;
        *axis = P.d;
        *piece_len = P.len_max;
        char* tmp = reinterpret_cast<char*>(kSizingBase);
        for (int64_t q = 0; q < P.k_pieces; ++q) {
            P.select(q, tmp);
            try {
                plan_reduce(&P.pe, &P.pr, &P.o, 4, kCus, nullptr, nullptr);
            } catch (const Error& err) {
                if (err.code != XG_UNSUPPORTED) throw;
                *all_fit = 0;
            }
        }
    });
}

}  // extern "C"
