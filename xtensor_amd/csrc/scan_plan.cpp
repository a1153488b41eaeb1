// scan_plan.cpp — host planning of cumsum / cumprod along an axis.
//
// xtensor's accumulator (reducers/xaccumulator.hpp:215-297) copies the
// expression into the result type R = decltype(f(init_type, value_type))
// and runs res[k] = f(res[k-1], res[k]) along the axis.  The plan keeps the
// same result type and element conversion; the order of the float additions
// becomes a tree (with compensation for sums).
#include <algorithm>
#include <cstring>

#include "scan.hpp"

namespace xg {

namespace {

int scan_rdt(int op, int acc_dtype, int edt) {
    int it = acc_dtype >= 0 ? acc_dtype : edt;
    if (it < 0 || it >= XG_NUM_DTYPES || it == XG_BFLOAT16) raise(XG_INVALID_ARGUMENT, "bad accumulator dtype");
    int at[2] = {it, edt};
    return result_dtype(op == XG_SCAN_CUMSUM ? XG_OP_PLUS : XG_OP_MULTIPLIES, 2, at, 0);
}

struct SDim {
    int64_t n;
    int64_t out;
    int64_t leaf[kMaxLeaves];
};

std::vector<SDim> merge(const std::vector<SDim>& d, int nl) {
    std::vector<SDim> m;
    for (const auto& x : d) {
        if (!m.empty()) {
            SDim& o = m.back();
            bool ok = o.out == x.out * x.n;
            for (int l = 0; l < nl && ok; ++l) ok = o.leaf[l] == x.leaf[l] * x.n;
            if (ok) {
                o.n *= x.n;
                o.out = x.out;
                for (int l = 0; l < nl; ++l) o.leaf[l] = x.leaf[l];
                continue;
            }
        }
        m.push_back(x);
    }
    return m;
}

}  // namespace

void scan_result(const xg_expr* e, const xg_scan_desc* s, int* dtype, Shape* shape) {
    if (!s) raise(XG_INVALID_ARGUMENT, "null scan descriptor");
    if (s->op != XG_SCAN_CUMSUM && s->op != XG_SCAN_CUMPROD) raise(XG_INVALID_ARGUMENT, "unknown scan op");
    ExprInfo info = analyse(e);
    *dtype = scan_rdt(s->op, s->acc_dtype, info.dtype);
    int nd = int(info.shape.size());
    if (s->axis == XG_AXIS_NONE) {
        *shape = Shape{numel(info.shape)};
        return;
    }
    int a = s->axis < 0 ? s->axis + nd : s->axis;
    if (a < 0 || a >= nd) raise(XG_BAD_AXES, "Axis larger than expression dimension in accumulator.");
    *shape = info.shape;
}

ScanPlan plan_scan(const xg_expr* e, const xg_scan_desc* s, const xg_tensor* out, int V) {
    if (!out || out->ndim < 0 || out->ndim > XG_MAX_RANK || out->dtype < 0 || out->dtype >= XG_NUM_DTYPES)
        raise(XG_INVALID_ARGUMENT, "malformed output tensor");
    ExprInfo info = analyse(e);
    int nd = int(info.shape.size());
    int axis = s->axis < 0 ? s->axis + nd : s->axis;
    if (axis < 0 || axis >= nd) raise(XG_BAD_AXES, "Axis larger than expression dimension in accumulator.");
    Shape os(out->shape, out->shape + out->ndim);
    if (os != info.shape) raise(XG_BROADCAST_ERROR, "scan output has shape " + shape_str(os) + ", expected " + shape_str(info.shape));
    ScanPlan p{};
    p.rdt = scan_rdt(s->op, s->acc_dtype, info.dtype);
    p.op = s->op == XG_SCAN_CUMSUM ? RO_ADD : RO_MUL;
    p.out_dtype = out->dtype;
    int r = p.rdt;
    if (r == XG_FLOAT32) p.acc = p.op == RO_ADD ? AK_F32C : AK_F32;
    else if (r == XG_FLOAT64) p.acc = p.op == RO_ADD ? AK_F64C : AK_F64;
    else if (r == XG_INT64) p.acc = AK_I64;
    else if (r == XG_UINT64) p.acc = AK_U64;
    else if (r == XG_UINT32) p.acc = AK_U32;
    else p.acc = AK_I32;
    p.prog = lower(e, info, r);
    uint64_t one = 0;
    if (p.op == RO_MUL) {
        if (r == XG_FLOAT32) { float f = 1.f; uint32_t b; std::memcpy(&b, &f, 4); one = b; }
        else if (r == XG_FLOAT64) { double d = 1.0; std::memcpy(&one, &d, 8); }
        else one = 1;
    }
    p.identity = one;
    if (numel(info.shape) == 0) { p.rows = 0; return p; }

    auto ls = leaf_strides(e, info.shape);
    int nl = p.prog.nleaves;
    SDim ad{};
    std::vector<SDim> before, after;
    for (int k = 0; k < nd; ++k) {
        SDim d{};
        d.n = info.shape[size_t(k)];
        d.out = out->shape[k] == 1 ? 0 : out->strides[k];
        for (int pl = 0; pl < nl; ++pl) d.leaf[pl] = p.prog.leaf_gather[pl] ? 0 : ls[size_t(p.prog.leaf_src[pl])][size_t(k)];
        if (k == axis) ad = d;
        else if (d.n != 1) (k < axis ? before : after).push_back(d);
    }
    if (ad.out < 0) raise(XG_UNSUPPORTED, "scan into a reversed output view");
    bool rows = ad.out == 1 || ad.n == 1 || after.empty();
    if (!rows) {
        after = merge(after, nl);
        if (after.back().out != 1) rows = true;  // no contiguous column dim: scan row by row
    }
    std::vector<SDim> order;
    if (rows) {
        p.layout = SL_ROWS;
        std::vector<SDim> outer = before;
        outer.insert(outer.end(), after.begin(), after.end());
        order = merge(outer, nl);
        order.push_back(ad);
    } else {
        p.layout = SL_COLS;
        std::vector<SDim> outer = merge(before, nl);
        for (size_t k = 0; k + 1 < after.size(); ++k) outer.push_back(after[k]);
        order = outer;
        order.push_back(ad);
        order.push_back(after.back());
    }
    p.rank = int(order.size());
    if (p.rank > kMaxRank) raise(XG_UNSUPPORTED, "scan rank too large");
    int64_t max_off = 0;
    for (int k = 0; k < p.rank; ++k) {
        if (order[size_t(k)].n >= (int64_t(1) << 31)) raise(XG_UNSUPPORTED, "dimension too large for 32-bit indexing");
        p.shape[k] = order[size_t(k)].n;
        p.div[k] = make_fastdiv(uint32_t(order[size_t(k)].n));
        p.out_stride[k] = order[size_t(k)].out;
        max_off += (order[size_t(k)].n - 1) * std::abs(order[size_t(k)].out);
    }
    int osz = dtype_size(out->dtype);
    p.out_ptr = static_cast<char*>(out->data) + out->offset * osz;
    uint64_t outer = 1;
    if (p.layout == SL_ROWS) {
        for (int k = 0; k + 1 < p.rank; ++k) outer *= uint64_t(p.shape[k]);
        p.len = uint32_t(p.shape[p.rank - 1]);
        p.rows = uint32_t(outer);
    } else {
        for (int k = 0; k + 2 < p.rank; ++k) outer *= uint64_t(p.shape[k]);
        p.len = uint32_t(p.shape[p.rank - 2]);
        p.inner = uint32_t(p.shape[p.rank - 1]);
        p.nchunks_in = (p.inner + uint32_t(V) - 1) / uint32_t(V);
        p.div_nchunks_in = make_fastdiv(p.nchunks_in);
        p.rows = uint32_t(outer * p.nchunks_in);
    }
    if (outer >= (uint64_t(1) << 31)) raise(XG_UNSUPPORTED, "scan too large for 32-bit indexing");
    // out vector stores along the last dim when contiguous and aligned
    bool oal = p.out_stride[p.rank - 1] == 1 && reinterpret_cast<uintptr_t>(p.out_ptr) % uintptr_t(osz * V) == 0;
    for (int k = 0; k + 1 < p.rank; ++k)
        if (p.out_stride[k] % V != 0) oal = false;
    p.out_mode = oal ? LM_CONTIG : LM_STRIDED;
    {
        int64_t total = 1;
        for (int k = 0; k < p.rank; ++k) total *= std::max<int64_t>(p.shape[k], 0);
        if (total * osz >= kNtMinBytes) p.out_mode |= LM_NT;
    }
    for (int pl = 0; pl < nl; ++pl) {
        LeafView& lv = p.leaves[pl];
        const xg_tensor& t = e->leaves[p.prog.leaf_src[pl]];
        if (p.prog.leaf_gather[pl]) {
            set_gather_leaf(t, &lv);
            continue;
        }
        int esz = dtype_size(t.dtype);
        lv.ptr = static_cast<const char*>(t.data) + t.offset * esz;
        lv.dtype = t.dtype;
        bool al = reinterpret_cast<uintptr_t>(lv.ptr) % uintptr_t(esz * V) == 0;
        int64_t lmax = 0;
        for (int k = 0; k < p.rank; ++k) {
            lv.stride[k] = order[size_t(k)].leaf[pl];
            lmax += (order[size_t(k)].n - 1) * std::abs(lv.stride[k]);
            if (k + 1 < p.rank && lv.stride[k] % V != 0) al = false;
        }
        max_off = std::max(max_off, lmax);
        int64_t is = lv.stride[p.rank - 1];
        lv.mode = is == 0 ? LM_BCAST : (is == 1 && al) ? LM_CONTIG : LM_STRIDED;
        if (lv.mode != LM_BCAST && streams(lv.stride, p.shape, p.rank, esz)) lv.mode |= LM_NT;
    }
    p.index64 = max_off >= (int64_t(1) << 31) ? 1 : 0;
    return p;
}

}  // namespace xg
