// reduce_plan.cpp — host planning of axis reductions.
//
// Result types follow xtensor 0.27.1: sum/prod/amax/amin have value type
// decltype(F(init_type, value_type)) with init_type = T or the expression's
// value type (core/xmath.hpp:367-423, reducers/xreducer.hpp:924-930); mean
// is sum<T> / double(e.size() / s.size()) (core/xmath.hpp:1825-1874); norms
// accumulate r + f(v) in the big-promoted type (reducers/xnorm.hpp:397-426,
// 493-526).  Axes are normalised and checked like xreducer's constructor
// (reducers/xreducer.hpp:1426-1440).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>

#include "reduce.hpp"

namespace xg {

int g_reduce_waves_per_cu = 16;
void set_reduce_waves_per_cu(int w) { g_reduce_waves_per_cu = w > 0 ? w : 16; }
// column fast kernel: 512-thread workgroups per CU, and the split partials'
// share of the input bytes, g_cols_partial_mult / 16
int g_cols_blocks_per_cu = 1;
int g_cols_partial_mult = 1;
void set_cols_split(int blocks_per_cu, int partial_mult) {
    g_cols_blocks_per_cu = blocks_per_cu > 0 ? blocks_per_cu : 1;
    g_cols_partial_mult = partial_mult > 0 ? partial_mult : 1;
}

namespace {

int big_promote(int dt) {  // restatement of xtl::big_promote_type_t for arithmetic types
    if (dt == XG_FLOAT32 || dt == XG_FLOAT64) return XG_FLOAT64;
    if (is_signed_int(dt)) return XG_INT64;
    return XG_UINT64;
}

int bin(int op, int a, int b) {
    int at[2] = {a, b};
    return result_dtype(op, 2, at, 0);
}

struct RedTypes {
    int vdt;   // value type of the accumulation
    int rdt;   // result type
    int map;   // per-element term
    int op;    // RedOp
    int fin;   // Finalize
    int init_dt;
    int div_dt;  // FZ_MEAN*: the divisor is static_cast to this type (T, or double)
    int mdt;     // FZ_MEAN*: type of value / divisor
};

// Types of one reduction pass.  For XG_RED_VARIANCE / XG_RED_STDDEV, `edt`
// is the type of the deviation e - mean(e) whose squares are averaged.
RedTypes red_types(int op, int acc_dtype, int edt) {
    RedTypes t{};
    const bool templ = op <= XG_RED_MEAN || op == XG_RED_VARIANCE || op == XG_RED_STDDEV;
    int it = (acc_dtype >= 0 && templ) ? acc_dtype : edt;
    if (it == XG_BFLOAT16 || it < 0 || it >= XG_NUM_DTYPES) raise(XG_INVALID_ARGUMENT, "bad accumulator dtype");
    t.init_dt = it;
    t.map = MAP_NONE;
    t.fin = FZ_NONE;
    // mean_division / mean_noaxis: value_type = T, or double without T (core/xmath.hpp:1827-1873)
    t.div_dt = acc_dtype >= 0 ? acc_dtype : XG_FLOAT64;
    switch (op) {
    case XG_RED_SUM: t.vdt = bin(XG_OP_PLUS, it, edt); t.op = RO_ADD; break;
    case XG_RED_PROD: t.vdt = bin(XG_OP_MULTIPLIES, it, edt); t.op = RO_MUL; break;
    case XG_RED_AMAX: t.vdt = bin(XG_OP_MAXIMUM, it, edt); t.op = RO_MAX; break;
    case XG_RED_AMIN: t.vdt = bin(XG_OP_MINIMUM, it, edt); t.op = RO_MIN; break;
    case XG_RED_MEAN: t.vdt = bin(XG_OP_PLUS, it, edt); t.op = RO_ADD; t.fin = FZ_MEAN; break;
    case XG_RED_VARIANCE: case XG_RED_STDDEV: {
        // detail::mean<T>(square(e - m), axes, ddof): sum<T> of the squares / T-or-double(n)
        t.map = MAP_SQUARE;
        int q = mapped_dtype(MAP_SQUARE, edt);
        int it2 = acc_dtype >= 0 ? acc_dtype : q;
        t.init_dt = it2;
        t.vdt = bin(XG_OP_PLUS, it2, q);
        t.op = RO_ADD;
        t.fin = op == XG_RED_VARIANCE ? FZ_MEAN : FZ_MEAN_SQRT;
        break;
    }
    case kRedMinNum: case kRedMaxNum:  // the two folds of xt::minmax, in the value type
        t.vdt = edt;
        t.init_dt = edt;
        if (is_float(edt)) t.op = op == kRedMinNum ? RO_MINNUM : RO_MAXNUM;
        else t.op = op == kRedMinNum ? RO_MIN : RO_MAX;
        break;
    case XG_RED_NORM_SQ: case XG_RED_NORM_L2:
        t.map = MAP_SQUARE;
        t.vdt = bin(XG_OP_PLUS, big_promote(edt), mapped_dtype(MAP_SQUARE, edt));
        t.op = RO_ADD;
        if (op == XG_RED_NORM_L2) t.fin = FZ_SQRT;
        break;
    case XG_RED_NORM_L1:
        t.map = MAP_ABS;
        t.vdt = bin(XG_OP_PLUS, big_promote(edt), mapped_dtype(MAP_ABS, edt));
        t.op = RO_ADD;
        break;
    case XG_RED_NORM_LINF: t.map = MAP_ABS; t.vdt = mapped_dtype(MAP_ABS, edt); t.op = RO_MAX; break;
    case XG_RED_NORM_L0: t.map = MAP_NONZERO; t.vdt = XG_UINT64; t.op = RO_ADD; break;
    default: raise(XG_INVALID_ARGUMENT, "unknown reduction " + std::to_string(op));
    }
    t.rdt = t.vdt;
    if (t.fin == FZ_MEAN || t.fin == FZ_MEAN_SQRT) {
        if (t.div_dt == XG_BFLOAT16) raise(XG_INVALID_ARGUMENT, "bad accumulator dtype");
        t.mdt = bin(XG_OP_DIVIDES, t.vdt, t.div_dt);
        t.rdt = t.mdt;
    }
    if (t.fin == FZ_SQRT || t.fin == FZ_MEAN_SQRT) {
        int at[1] = {t.rdt};
        t.rdt = result_dtype(XG_OP_SQRT, 1, at, 0);
    }
    return t;
}

// the type of e - mean<T>(e) for an element type edt (variance / stddev)
int deviation_dtype(int acc_dtype, int edt) {
    RedTypes m = red_types(XG_RED_MEAN, acc_dtype, edt);
    return bin(XG_OP_MINUS, edt, m.rdt);
}

std::vector<int> norm_axes(const xg_reduce_desc* r, int ndim) {
    std::vector<int> ax;
    if (r->n_axes < 0) {
        for (int k = 0; k < ndim; ++k) ax.push_back(k);
        return ax;
    }
    if (r->n_axes > XG_MAX_RANK) raise(XG_BAD_AXES, "too many reduction axes");
    for (int k = 0; k < r->n_axes; ++k) {
        int a = r->axes[k];
        if (a < 0) a += ndim;
        if (a < 0) raise(XG_BAD_AXES, "Axis " + std::to_string(r->axes[k]) + " out of bounds for reduction.");
        ax.push_back(a);
    }
    if (!std::is_sorted(ax.begin(), ax.end())) raise(XG_BAD_AXES, "Reducing axes should be sorted.");
    if (std::adjacent_find(ax.begin(), ax.end()) != ax.end())
        raise(XG_BAD_AXES, "Reducing axes should not contain duplicates.");
    if (!ax.empty() && ax.back() > ndim - 1)
        raise(XG_BAD_AXES, "Axis " + std::to_string(ax.back()) + " out of bounds for reduction.");
    return ax;
}

int acc_kind(const RedTypes& t, int edt, int xop) {
    int v = t.vdt;
    bool elem_f32 = edt == XG_FLOAT32;
    if (v == XG_FLOAT32 || v == XG_FLOAT64) {
        if (t.op == RO_ADD) {
            if (v == XG_FLOAT32) return AK_F32C;
            // float elements of the norms and of mean accumulate in float with compensation
            bool norm = xop == XG_RED_NORM_SQ || xop == XG_RED_NORM_L2 || xop == XG_RED_NORM_L1;
            return (elem_f32 && norm) ? AK_F32C : AK_F64C;
        }
        return v == XG_FLOAT32 ? AK_F32 : AK_F64;
    }
    switch (v) {
    case XG_INT64: return AK_I64;
    case XG_UINT64: return AK_U64;
    case XG_UINT32: return AK_U32;
    case XG_INT32: return AK_I32;
    default: return is_unsigned_int(v) ? AK_U32 : AK_I32;  // small types held in 32 bits
    }
}

int acc_input_dtype(int ak) {
    switch (ak) {
    case AK_F32C: case AK_F32: return XG_FLOAT32;
    case AK_F64C: case AK_F64: return XG_FLOAT64;
    case AK_I32: return XG_INT32;
    case AK_U32: return XG_UINT32;
    case AK_I64: return XG_INT64;
    default: return XG_UINT64;
    }
}

// bits of a value (given as double / int64 / uint64 by kind of dt) in register form of `dt`
uint64_t scalar_bits(const xg_scalar& s, int dt) {
    // static_cast<dt>(value of s.dtype)
    double f = 0;
    int64_t i = 0;
    uint64_t u = 0;
    bool is_f = s.dtype == XG_FLOAT32 || s.dtype == XG_FLOAT64;
    bool is_u = is_unsigned_int(s.dtype) && s.dtype != XG_BOOL;
    if (is_f) f = s.dtype == XG_FLOAT32 ? double(float(s.v.f)) : s.v.f;
    else if (is_u) u = s.v.u;
    else i = s.v.i;
    auto as_f64 = [&]() { return is_f ? f : is_u ? double(u) : double(i); };
    auto as_i64 = [&]() { return is_f ? int64_t(f) : is_u ? int64_t(u) : i; };
    uint64_t bits = 0;
    switch (dt) {
    case XG_FLOAT32: { float x = is_f ? float(f) : is_u ? float(u) : float(i); uint32_t b; std::memcpy(&b, &x, 4); bits = b; break; }
    case XG_FLOAT64: { double x = as_f64(); std::memcpy(&bits, &x, 8); break; }
    case XG_BOOL: bits = is_f ? (f != 0) : (as_i64() != 0); break;
    case XG_INT8: bits = uint32_t(int32_t(int8_t(as_i64()))); break;
    case XG_UINT8: bits = uint8_t(as_i64()); break;
    case XG_INT16: bits = uint32_t(int32_t(int16_t(as_i64()))); break;
    case XG_UINT16: bits = uint16_t(as_i64()); break;
    case XG_INT32: bits = uint32_t(int32_t(as_i64())); break;
    case XG_UINT32: bits = uint32_t(is_u ? u : uint64_t(as_i64())); break;
    case XG_INT64: bits = uint64_t(as_i64()); break;
    case XG_UINT64: bits = is_f ? uint64_t(f) : is_u ? u : uint64_t(i); break;
    default: raise(XG_INVALID_ARGUMENT, "bad dtype");
    }
    return bits;
}

xg_scalar limit_scalar(int dt, bool lowest) {
    xg_scalar s{};
    s.dtype = dt;
    switch (dt) {
    case XG_FLOAT32: s.v.f = lowest ? -double(std::numeric_limits<float>::max()) : double(std::numeric_limits<float>::max()); break;
    case XG_FLOAT64: s.v.f = lowest ? std::numeric_limits<double>::lowest() : std::numeric_limits<double>::max(); break;
    case XG_BOOL: s.v.i = lowest ? 0 : 1; break;
    case XG_INT8: s.v.i = lowest ? INT8_MIN : INT8_MAX; break;
    case XG_UINT8: s.v.u = lowest ? 0 : UINT8_MAX; break;
    case XG_INT16: s.v.i = lowest ? INT16_MIN : INT16_MAX; break;
    case XG_UINT16: s.v.u = lowest ? 0 : UINT16_MAX; break;
    case XG_INT32: s.v.i = lowest ? INT32_MIN : INT32_MAX; break;
    case XG_UINT32: s.v.u = lowest ? 0 : UINT32_MAX; break;
    case XG_INT64: s.v.i = lowest ? INT64_MIN : INT64_MAX; break;
    case XG_UINT64: s.v.u = lowest ? 0 : UINT64_MAX; break;
    default: raise(XG_INVALID_ARGUMENT, "bad dtype");
    }
    return s;
}

// the value held in register bits `vb` of type dt, as a scalar
xg_scalar scalar_of_bits(uint64_t vb, int dt) {
    xg_scalar sv{};
    sv.dtype = dt;
    if (dt == XG_FLOAT32) { float x; uint32_t b = uint32_t(vb); std::memcpy(&x, &b, 4); sv.v.f = x; }
    else if (dt == XG_FLOAT64) std::memcpy(&sv.v.f, &vb, 8);
    else if (is_unsigned_int(dt) && dt != XG_BOOL) sv.v.u = (dt == XG_UINT64) ? vb : uint64_t(uint32_t(vb));
    else sv.v.i = (dt == XG_INT64) ? int64_t(vb) : int64_t(int32_t(uint32_t(vb)));
    return sv;
}

// value -> V -> accumulator representation (identity / initial)
uint64_t to_acc_bits(const xg_scalar& s, int vdt, int ak) {
    return scalar_bits(scalar_of_bits(scalar_bits(s, vdt), vdt), acc_input_dtype(ak));
}

struct RDim {
    int64_t n;
    bool red;
    int64_t out;  // output stride (kept dims)
    int64_t leaf[kMaxLeaves];
};

bool contiguous(const RDim& outer, const RDim& inner, int nleaves) {
    if (outer.red != inner.red) return false;
    if (!outer.red && outer.out != inner.out * inner.n) return false;
    for (int l = 0; l < nleaves; ++l)
        if (outer.leaf[l] != inner.leaf[l] * inner.n) return false;
    return true;
}

std::vector<RDim> merge_dims(const std::vector<RDim>& d, int nleaves) {
    std::vector<RDim> m;
    for (const auto& x : d) {
        if (!m.empty() && contiguous(m.back(), x, nleaves)) {
            RDim& o = m.back();
            o.n *= x.n;
            o.out = x.out;
            for (int l = 0; l < nleaves; ++l) o.leaf[l] = x.leaf[l];
            continue;
        }
        m.push_back(x);
    }
    return m;
}

// static_cast<T-or-double>((size - ddof) / outputs), in size_t arithmetic
// (detail::mean / mean_division, core/xmath.hpp:1827-1850), then converted
// to the type of the division: the divisor in register form of t.mdt
uint64_t mean_divisor_bits(const RedTypes& t, int64_t total, int64_t n_out, int64_t ddof) {
    if (ddof < 0 || ddof > total) raise(XG_INVALID_ARGUMENT, "ddof out of range");
    xg_scalar q{};
    q.dtype = XG_UINT64;
    q.v.u = n_out != 0 ? uint64_t(total - ddof) / uint64_t(n_out) : 0;
    return scalar_bits(scalar_of_bits(scalar_bits(q, t.div_dt), t.div_dt), t.mdt);
}

uint32_t pow2_ceil(uint64_t x) {
    uint32_t p = 1;
    while (p < x && p < (1u << 30)) p <<= 1;
    return p;
}

}  // namespace

void reduce_result(const xg_expr* e, const xg_reduce_desc* r, int* dtype, Shape* shape) {
    if (!r) raise(XG_INVALID_ARGUMENT, "null reduce descriptor");
    ExprInfo info = analyse(e);
    if (r->op == XG_RED_MINMAX) {  // std::array<value_type, 2> over every element
        if (r->n_axes >= 0 || r->keep_dims || r->has_initial)
            raise(XG_INVALID_ARGUMENT, "minmax reduces over all axes and takes no options");
        *dtype = info.dtype;
        *shape = Shape{2};
        return;
    }
    std::vector<int> ax = norm_axes(r, int(info.shape.size()));
    int edt = info.dtype;
    if (r->op == XG_RED_VARIANCE || r->op == XG_RED_STDDEV) edt = deviation_dtype(r->acc_dtype, edt);
    RedTypes t = red_types(r->op, r->acc_dtype, edt);
    Shape rs;
    for (int k = 0; k < int(info.shape.size()); ++k) {
        bool red = std::find(ax.begin(), ax.end(), k) != ax.end();
        if (!red) rs.push_back(info.shape[size_t(k)]);
        else if (r->keep_dims) rs.push_back(1);
    }
    *dtype = t.rdt;
    *shape = rs;
}

RedFinish reduce_finish(const xg_expr* e, const xg_reduce_desc* r) {
    if (!r) raise(XG_INVALID_ARGUMENT, "null reduce descriptor");
    if (r->op == XG_RED_MINMAX) raise(XG_INVALID_ARGUMENT, "minmax is planned as two folds");
    ExprInfo info = analyse(e);
    const int nd = int(info.shape.size());
    const std::vector<int> ax = norm_axes(r, nd);
    const RedTypes t = red_types(r->op, r->acc_dtype, info.dtype);
    int64_t n_out = 1;
    for (int k = 0; k < nd; ++k)
        if (std::find(ax.begin(), ax.end(), k) == ax.end()) n_out *= info.shape[size_t(k)];
    RedFinish f{};
    f.vdt = t.vdt;
    f.rdt = t.rdt;
    f.fin = t.fin;
    f.mdt = t.mdt;
    if (t.fin == FZ_MEAN || t.fin == FZ_MEAN_SQRT) {
        const bool moment = r->op == XG_RED_VARIANCE || r->op == XG_RED_STDDEV;
        f.div_bits = mean_divisor_bits(t, numel(info.shape), n_out, moment ? r->ddof : 0);
    }
    return f;
}

RedPlan plan_reduce(const xg_expr* e, const xg_reduce_desc* r, const xg_tensor* out, int V, int cus,
                    int* result_dtype_out, Shape* result_shape_out) {
    if (!r) raise(XG_INVALID_ARGUMENT, "null reduce descriptor");
    if (!out || out->ndim < 0 || out->ndim > XG_MAX_RANK || out->dtype < 0 || out->dtype >= XG_NUM_DTYPES)
        raise(XG_INVALID_ARGUMENT, "malformed output tensor");
    ExprInfo info = analyse(e);
    const Shape& in = info.shape;
    int nd = int(in.size());
    if (r->op == XG_RED_MINMAX) raise(XG_INVALID_ARGUMENT, "minmax is planned as two folds");
    const bool moment = r->op == XG_RED_VARIANCE || r->op == XG_RED_STDDEV;
    if (r->ddof != 0 && !moment) raise(XG_INVALID_ARGUMENT, "ddof applies to variance and stddev only");
    std::vector<int> ax = norm_axes(r, nd);
    RedTypes t = red_types(r->op, r->acc_dtype, info.dtype);
    Shape rs;
    std::vector<int> out_dim_of(size_t(nd), -1);
    for (int k = 0; k < nd; ++k) {
        bool red = std::find(ax.begin(), ax.end(), k) != ax.end();
        if (!red) { out_dim_of[size_t(k)] = int(rs.size()); rs.push_back(in[size_t(k)]); }
        else if (r->keep_dims) rs.push_back(1);
    }
    if (result_dtype_out) *result_dtype_out = t.rdt;
    if (result_shape_out) *result_shape_out = rs;
    Shape os(out->shape, out->shape + out->ndim);
    if (os != rs) raise(XG_BROADCAST_ERROR, "reduction output has shape " + shape_str(os) + ", expected " + shape_str(rs));

    RedPlan p{};
    p.op = t.op;
    p.fin = t.fin;
    p.vdt = t.vdt;
    p.rdt = t.rdt;
    p.out_dtype = out->dtype;
    p.acc = acc_kind(t, info.dtype, r->op);
    int acc_dt = acc_input_dtype(p.acc);
    p.prog = lower(e, info, acc_dt, t.map);
    // identity = xtensor's init value (0, 1, lowest(), max()) in the value type
    xg_scalar init{};
    if (t.op == RO_ADD || t.op == RO_MUL) {
        init.dtype = XG_INT32;
        init.v.i = t.op == RO_MUL ? 1 : 0;
    } else {
        init = limit_scalar(t.init_dt, t.op == RO_MAX || t.op == RO_MAXNUM);
    }
    p.identity = to_acc_bits(init, t.vdt, p.acc);
    p.has_initial = r->has_initial ? 1 : 0;
    if (r->has_initial) {
        if (r->initial.dtype < 0 || r->initial.dtype >= XG_NUM_DTYPES || r->initial.dtype == XG_BFLOAT16)
            raise(XG_INVALID_ARGUMENT, "bad initial value dtype");
        p.initial = to_acc_bits(r->initial, t.vdt, p.acc);
    }
    int64_t n_out = numel(rs);
    int64_t total = numel(in);
    if (t.fin == FZ_MEAN || t.fin == FZ_MEAN_SQRT) {
        p.mdt = t.mdt;
        p.div_bits = mean_divisor_bits(t, total, n_out, moment ? r->ddof : 0);
    }
    if (n_out == 0) {
        p.n_out = 0;
        return p;
    }
    if (n_out >= (int64_t(1) << 31)) raise(XG_UNSUPPORTED, "too many reduction outputs");

    // dims of the iteration space (drop extent-1 dims; keep empty ones visible)
    auto ls = leaf_strides(e, in);
    int nl = p.prog.nleaves;
    std::vector<RDim> dims;
    int64_t out_off = out->offset;
    std::vector<int64_t> leaf_off(static_cast<size_t>(nl));
    for (int pl = 0; pl < nl; ++pl) leaf_off[size_t(pl)] = e->leaves[p.prog.leaf_src[pl]].offset;
    bool empty_red = total == 0;  // every output is the identity / initial value
    for (int k = 0; k < nd && !empty_red; ++k) {
        if (in[size_t(k)] == 1) continue;
        RDim d{};
        d.n = in[size_t(k)];
        d.red = out_dim_of[size_t(k)] < 0;
        int od = r->keep_dims ? k : out_dim_of[size_t(k)];
        d.out = d.red ? 0 : (out->shape[od] == 1 ? 0 : out->strides[od]);
        for (int pl = 0; pl < nl; ++pl) d.leaf[pl] = p.prog.leaf_gather[pl] ? 0 : ls[size_t(p.prog.leaf_src[pl])][size_t(k)];
        dims.push_back(d);
    }
    if (empty_red) {  // keep the kept dims for output addressing, reduce nothing
        for (int k = 0; k < nd; ++k) {
            if (out_dim_of[size_t(k)] < 0 || in[size_t(k)] == 1) continue;
            RDim d{};
            d.n = in[size_t(k)];
            d.red = false;
            int od = r->keep_dims ? k : out_dim_of[size_t(k)];
            d.out = out->strides[od];
            dims.push_back(d);
        }
    }
    // dominant leaf: the one reading the most bytes
    int dom = -1;
    double best = -1;
    for (int pl = 0; pl < nl; ++pl) {
        double fp = double(dtype_size(e->leaves[p.prog.leaf_src[pl]].dtype));
        for (auto& d : dims)
            if (d.leaf[pl] != 0) fp *= double(d.n);
        if (fp > best) { best = fp; dom = pl; }
    }
    // iterate dims with a negative dominant stride backwards
    if (dom >= 0) {
        for (auto& d : dims) {
            if (d.leaf[dom] < 0) {
                for (int pl = 0; pl < nl; ++pl) {
                    leaf_off[size_t(pl)] += (d.n - 1) * d.leaf[pl];
                    d.leaf[pl] = -d.leaf[pl];
                }
                if (!d.red) {
                    out_off += (d.n - 1) * d.out;
                    d.out = -d.out;
                }
            }
        }
    }
    dims = merge_dims(dims, nl);
    // fast dim: smallest non-zero dominant stride (last dim when none)
    int fast = dims.empty() ? -1 : int(dims.size()) - 1;
    if (dom >= 0) {
        int64_t bs = std::numeric_limits<int64_t>::max();
        for (int k = int(dims.size()) - 1; k >= 0; --k) {
            int64_t s = dims[size_t(k)].leaf[dom];
            if (s != 0 && s < bs) { bs = s; fast = k; }
        }
    }
    std::vector<RDim> kept, red;
    for (int k = 0; k < int(dims.size()); ++k) (dims[size_t(k)].red ? red : kept).push_back(dims[size_t(k)]);
    auto by_stride = [&](const RDim& a, const RDim& b) {
        int64_t sa = dom >= 0 ? a.leaf[dom] : 0, sb = dom >= 0 ? b.leaf[dom] : 0;
        return sa > sb;
    };
    // an empty reduction reads nothing: rows layout with zero chunks per output
    bool fast_red = empty_red || fast < 0 || dims[size_t(fast)].red || kept.empty();
    std::vector<RDim> order;
    RDim unit{};
    unit.n = 1;
    unit.red = true;
    if (fast_red) {
        p.layout = RL_ROWS;
        std::stable_sort(red.begin(), red.end(), by_stride);
        red = merge_dims(red, nl);
        if (red.empty() || empty_red) {  // nothing (or no element) to reduce: one chunk of length 0/1
            unit.n = empty_red ? 0 : 1;
            red.assign(1, unit);
        }
        kept = merge_dims(kept, nl);
        order = kept;
        order.insert(order.end(), red.begin(), red.end());
        p.n_kept = int(kept.size());
        p.n_red = int(red.size());
    } else {
        p.layout = RL_COLS;
        RDim f = dims[size_t(fast)];
        std::vector<RDim> ko;
        for (int k = 0; k < int(dims.size()); ++k)
            if (!dims[size_t(k)].red && k != fast) ko.push_back(dims[size_t(k)]);
        ko = merge_dims(ko, nl);
        std::stable_sort(red.begin(), red.end(), by_stride);
        red = merge_dims(red, nl);
        if (red.empty()) red.assign(1, unit);
        order = ko;
        order.insert(order.end(), red.begin(), red.end());
        order.push_back(f);
        p.n_kept = int(ko.size());
        p.n_red = int(red.size());
    }
    p.rank = int(order.size());
    if (p.rank > kMaxRank) raise(XG_UNSUPPORTED, "reduction rank too large");
    int64_t max_off = 0;
    for (int k = 0; k < p.rank; ++k) {
        const RDim& d = order[size_t(k)];
        if (d.n >= (int64_t(1) << 31)) raise(XG_UNSUPPORTED, "dimension too large for 32-bit indexing");
        p.shape[k] = d.n;
        p.div[k] = make_fastdiv(uint32_t(std::max<int64_t>(d.n, 1)));
        p.out_stride[k] = d.red ? 0 : d.out;
        max_off += (std::max<int64_t>(d.n, 1) - 1) * std::abs(d.out);
    }
    p.n_out = uint32_t(n_out);
    p.lf = uint32_t(order.back().n);
    p.nchunks_f = uint32_t((p.lf + uint32_t(V) - 1) / uint32_t(V));
    p.div_nchunks_f = make_fastdiv(std::max<uint32_t>(p.nchunks_f, 1));
    int osz = dtype_size(out->dtype);
    p.out_ptr = static_cast<char*>(out->data) + out_off * osz;
    for (int pl = 0; pl < nl; ++pl) {
        LeafView& lv = p.leaves[pl];
        const xg_tensor& tl = e->leaves[p.prog.leaf_src[pl]];
        if (p.prog.leaf_gather[pl]) {
            set_gather_leaf(tl, &lv);
            continue;
        }
        int esz = dtype_size(tl.dtype);
        lv.ptr = static_cast<const char*>(tl.data) + leaf_off[size_t(pl)] * esz;
        lv.dtype = tl.dtype;
        int64_t lmax = 0;
        bool al = reinterpret_cast<uintptr_t>(lv.ptr) % uintptr_t(esz * V) == 0;
        for (int k = 0; k < p.rank; ++k) {
            lv.stride[k] = order[size_t(k)].leaf[pl];
            lmax += (std::max<int64_t>(order[size_t(k)].n, 1) - 1) * std::abs(lv.stride[k]);
            if (k + 1 < p.rank && lv.stride[k] % V != 0) al = false;
        }
        max_off = std::max(max_off, lmax);
        int64_t is = lv.stride[p.rank - 1];
        lv.mode = is == 0 ? LM_BCAST : (is == 1 && al) ? LM_CONTIG : LM_STRIDED;
        if (lv.mode != LM_BCAST && streams(lv.stride, p.shape, p.rank, esz)) lv.mode |= LM_NT;
    }
    p.index64 = max_off >= (int64_t(1) << 31) ? 1 : 0;

    // work decomposition
    if (best < 0) best = 0;
    if (p.layout == RL_ROWS) {
        uint64_t outer = 1;
        for (int k = p.n_kept; k < p.rank - 1; ++k) outer *= uint64_t(p.shape[k]);
        uint64_t chunks = outer * p.nchunks_f;
        if (chunks >= (uint64_t(1) << 32)) raise(XG_UNSUPPORTED, "reduction too long for 32-bit indexing");
        p.red_count = uint32_t(chunks);
        uint32_t g = pow2_ceil((chunks + 3) / 4);
        p.group = int(std::min<uint32_t>(std::max<uint32_t>(g, 1), 256));
        uint64_t want = (uint64_t(cus) * 1024 + uint64_t(n_out) * p.group - 1) / (uint64_t(n_out) * p.group);
        uint64_t maxs = std::max<uint64_t>(1, chunks / (uint64_t(p.group) * 2));
        p.splits = int(std::max<uint64_t>(1, std::min<uint64_t>(std::min(want, maxs), 65535)));
        p.per_split = uint32_t((chunks + p.splits - 1) / p.splits);
    } else {
        uint64_t rows = 1;
        for (int k = p.n_kept; k < p.rank - 1; ++k) rows *= uint64_t(p.shape[k]);
        if (rows >= (uint64_t(1) << 32)) raise(XG_UNSUPPORTED, "reduction too long for 32-bit indexing");
        p.red_count = uint32_t(rows);
        p.group = 256;
        p.col_tiles = (p.nchunks_f + 63) / 64;
        p.div_col_tiles = make_fastdiv(std::max<uint32_t>(p.col_tiles, 1));
        uint64_t n_ko = uint64_t(n_out) / std::max<uint32_t>(p.lf, 1);
        uint64_t items = std::max<uint64_t>(n_ko * p.col_tiles, 1);
        uint64_t want = (uint64_t(cus) * 4 + items - 1) / items;
        uint64_t maxs = std::max<uint64_t>(1, rows / 32);
        // keep the partial-result traffic to a few percent of the input
        uint64_t traffic = std::max<uint64_t>(1, uint64_t(best) / (uint64_t(n_out) * 16 * 32));
        p.splits = int(std::max<uint64_t>(1, std::min<uint64_t>({want, maxs, traffic, 65535})));
        p.per_split = uint32_t((rows + p.splits - 1) / p.splits);
    }
    while (p.splits > 1 && uint64_t(n_out) * uint64_t(p.splits) >= (uint64_t(1) << 31)) p.splits /= 2;
    if (p.layout == RL_ROWS) p.per_split = uint32_t((uint64_t(p.red_count) + p.splits - 1) / p.splits);
    else p.per_split = uint32_t((uint64_t(p.red_count) + p.splits - 1) / p.splits);
    p.div_splits = make_fastdiv(uint32_t(p.splits));

    // compact rank <= 2 form and the work split of the compiled fast kernels
    const bool rows_ok = p.layout == RL_ROWS && p.rank <= 2 && p.n_red == 1;
    const bool cols_ok = p.layout == RL_COLS && p.rank == 2 && p.n_kept == 0;
    if (!p.index64 && (rows_ok || cols_ok) && !empty_red) {
        p.compact = 1;
        for (int pl = 0; pl < nl; ++pl)
            for (int k = 0; k < p.rank; ++k) p.cst[pl][k] = p.prog.leaf_gather[pl] ? 0 : int32_t(p.leaves[pl].stride[k]);
        // enough loads in flight to cover HBM latency
        const uint64_t target_waves = uint64_t(cus) * uint64_t(g_reduce_waves_per_cu);
        constexpr uint64_t kU = 4;                          // chunks (rows) in flight per lane
        if (p.layout == RL_ROWS) {
            const uint64_t chunks = p.nchunks_f;
            uint64_t g = pow2_ceil(uint32_t(std::max<uint64_t>(1, (chunks + 2 * kU - 1) / (2 * kU))));
            if (uint64_t(n_out) < target_waves) g = 64;  // few outputs: whole waves, then splits
            p.f_group = int(std::min<uint64_t>(std::max<uint64_t>(g, 1), 64));
            const uint64_t waves = (uint64_t(n_out) * p.f_group + 63) / 64;
            uint64_t s = 1;
            if (waves < target_waves) {
                s = (target_waves + waves - 1) / waves;
                s = std::min<uint64_t>(s, std::max<uint64_t>(1, chunks / (uint64_t(p.f_group) * kU * 2)));
            }
            p.f_splits = int(std::min<uint64_t>(s, 65535));
            p.f_per_split = uint32_t((chunks + p.f_splits - 1) / p.f_splits);
        } else {
            const uint64_t rows = p.red_count;
            // tile width in lanes: 128 when that leaves at most an eighth of
            // the lanes without a column, else 64 (on a [8192, 4096] column
            // sum, float32 / bf16: 64 lanes 4.23 / 2.45 TB/s, 128 lanes
            // 4.19 / 2.55, 256 lanes 3.90 / 2.03; MI355X)
            const uint64_t nc = p.nchunks_f;
            const uint64_t tw = (nc >= 128 && ((nc + 127) / 128 * 128 - nc) * 8 <= nc) ? 128 : 64;
            const uint64_t tiles = (nc + tw - 1) / tw;
            const uint64_t row_groups = 512 / tw;
            // blocks of 512 threads: one per CU at the default 16 waves per CU
            // (4.2 TB/s on a [8192, 4096] float32 column sum, 3.7 with two per CU)
            const uint64_t target_blocks = std::max<uint64_t>(1, target_waves / 16 * uint64_t(g_cols_blocks_per_cu));
            uint64_t s = (target_blocks + tiles - 1) / tiles;
            s = std::min<uint64_t>(s, std::max<uint64_t>(1, rows / (row_groups * kU * 2)));
            // partial results stay within about 6% of the input bytes
            s = std::min<uint64_t>(s, std::max<uint64_t>(1, uint64_t(best) * uint64_t(g_cols_partial_mult) /
                                                                         (uint64_t(n_out) * 16 * 16)));
            p.f_group = int(tw);
            p.f_splits = int(std::min<uint64_t>(s, 65535));
            p.f_per_split = uint32_t((rows + p.f_splits - 1) / p.f_splits);
        }
        while (p.f_splits > 1 && uint64_t(n_out) * uint64_t(p.f_splits) >= (uint64_t(1) << 31)) p.f_splits /= 2;
        if (p.layout == RL_ROWS) p.f_per_split = uint32_t((uint64_t(p.nchunks_f) + p.f_splits - 1) / p.f_splits);
        else p.f_per_split = uint32_t((uint64_t(p.red_count) + p.f_splits - 1) / p.f_splits);
    }
    return p;
}

size_t reduce_workspace_bytes(const RedPlan& p) {
    const int s = std::max(p.splits, p.compact ? p.f_splits : 1);
    return s > 1 ? size_t(s) * size_t(p.n_out) * 16 : 0;
}

}  // namespace xg
