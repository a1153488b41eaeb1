// planner.cpp — see planner.hpp.
//
// Semantics restated from xtensor 0.27.1:
//  * value type of a node = decltype(functor(args...)), core/xfunction.hpp:133-142;
//    the functors are C++ operators (core/xoperation.hpp:86-109), std:: math
//    (core/xmath.hpp:152-217), maximum/minimum/clamp via xtl::select
//    (core/xmath.hpp:573-617) and where (core/xoperation.hpp:775-784).  The
//    types below are computed by the C++ compiler on the same expressions.
//  * broadcast shape = right-aligned fold of the leaf shapes in tree order,
//    core/xstrides.hpp:737-780 through core/xfunction.hpp:537-544, 709-727.
//  * broadcast strides: size-1 and missing leading dims read with stride 0,
//    core/xstrides.hpp:503-530 and xstepper offsets, core/xiterator.hpp:475-587.
#include "planner.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <numeric>
#include <type_traits>

namespace xg {

int g_nt_loads = 1;   // non-temporal loads of streaming leaves
int g_nt_stores = 1;  // non-temporal stores of large outputs
int g_fill_granules = 1;  // granule-fill reads for rows that share 64-byte granules with unwritten memory


[[noreturn]] void raise(int code, const std::string& msg) { throw Error{code, msg}; }

int64_t numel(const Shape& s) {
    int64_t n = 1;
    for (auto x : s) n *= x;
    return n;
}

std::string shape_str(const Shape& s) {
    std::string r = "(";
    for (size_t i = 0; i < s.size(); ++i) r += (i ? ", " : "") + std::to_string(s[i]);
    return r + ")";
}

// ------------------------------------------------------------------ types
namespace {

template <class T> struct tg { using type = T; };

template <class T> constexpr int dt_of() {
    if constexpr (std::is_same_v<T, bool>) return XG_BOOL;
    else if constexpr (std::is_same_v<T, int8_t>) return XG_INT8;
    else if constexpr (std::is_same_v<T, uint8_t>) return XG_UINT8;
    else if constexpr (std::is_same_v<T, int16_t>) return XG_INT16;
    else if constexpr (std::is_same_v<T, uint16_t>) return XG_UINT16;
    else if constexpr (std::is_same_v<T, int32_t>) return XG_INT32;
    else if constexpr (std::is_same_v<T, uint32_t>) return XG_UINT32;
    else if constexpr (std::is_same_v<T, int64_t> || std::is_same_v<T, long long>) return XG_INT64;
    else if constexpr (std::is_same_v<T, uint64_t> || std::is_same_v<T, unsigned long long>) return XG_UINT64;
    else if constexpr (std::is_same_v<T, float>) return XG_FLOAT32;
    else if constexpr (std::is_same_v<T, double>) return XG_FLOAT64;
    else return -1;
}

template <class F> int visit_dt(int dt, F&& f) {
    switch (dt) {
    case XG_BOOL: return f(tg<bool>{});
    case XG_INT8: return f(tg<int8_t>{});
    case XG_UINT8: return f(tg<uint8_t>{});
    case XG_INT16: return f(tg<int16_t>{});
    case XG_UINT16: return f(tg<uint16_t>{});
    case XG_INT32: return f(tg<int32_t>{});
    case XG_UINT32: return f(tg<uint32_t>{});
    case XG_INT64: return f(tg<int64_t>{});
    case XG_UINT64: return f(tg<uint64_t>{});
    case XG_FLOAT32: return f(tg<float>{});
    case XG_FLOAT64: return f(tg<double>{});
    default: raise(XG_INVALID_ARGUMENT, "bad dtype " + std::to_string(dt));
    }
}

#define XG_UN(NAME, EXPR) \
    struct NAME { template <class A> static auto f(A a) -> decltype(EXPR); };
#define XG_BIN(NAME, EXPR) \
    struct NAME { template <class A, class B> static auto f(A a, B b) -> decltype(EXPR); };

XG_UN(t_negate, -a)
XG_UN(t_identity, +a)
XG_UN(t_lnot, !a)
XG_UN(t_bnot, ~a)
XG_BIN(t_arith, a + b)   // usual arithmetic conversions (+ - * / and compares)
XG_BIN(t_mod, a % b)     // integral only
XG_BIN(t_shl, a << b)
XG_BIN(t_land, a && b)
XG_BIN(t_common, (std::common_type_t<A, B>)a)  // xtl::select result
XG_BIN(t_pow, std::pow(a, b))
// math::abs is the identity on unsigned integer types (core/xmath.hpp:59-63, 283-287)
template <class A> auto xabs(A a) {
    if constexpr (std::is_unsigned_v<A> && !std::is_same_v<A, bool>) return a;
    else return std::abs(a);
}
XG_UN(t_abs, xabs(a))
XG_UN(t_fabs, std::fabs(a))
XG_UN(t_exp, std::exp(a))   // every unary std:: math function has this signature set

template <class F, class... A> constexpr bool callable = requires(A... a) { F::f(a...); };

template <class F> int un_type(int a) {
    return visit_dt(a, [](auto ta) {
        using A = typename decltype(ta)::type;
        if constexpr (callable<F, A>) return dt_of<decltype(F::f(A()))>();
        else return -1;
    });
}

template <class F> int bin_type(int a, int b) {
    return visit_dt(a, [&](auto ta) {
        using A = typename decltype(ta)::type;
        return visit_dt(b, [](auto tb) {
            using B = typename decltype(tb)::type;
            if constexpr (callable<F, A, B>) return dt_of<decltype(F::f(A(), B()))>();
            else return -1;
        });
    });
}

bool is_int_dt(int dt) { return dt != XG_FLOAT32 && dt != XG_FLOAT64 && dt != XG_BFLOAT16; }

}  // namespace

int result_dtype(int op, int nargs, const int* t, int index) {
    auto need = [&](int n) {
        if (nargs != n) raise(XG_INVALID_ARGUMENT, "op " + std::to_string(op) + " takes " + std::to_string(n) + " arguments");
    };
    int r = -1;
    switch (op) {
    case XG_OP_NEGATE: need(1); r = un_type<t_negate>(t[0]); break;
    case XG_OP_IDENTITY: need(1); r = un_type<t_identity>(t[0]); break;
    case XG_OP_LOGICAL_NOT: need(1); r = XG_BOOL; break;
    case XG_OP_BITWISE_NOT: need(1); r = un_type<t_bnot>(t[0]); break;
    case XG_OP_CAST:
        need(1);
        if (index < 0 || index >= XG_NUM_DTYPES || index == XG_BFLOAT16) raise(XG_INVALID_ARGUMENT, "bad cast dtype");
        r = index;
        break;
    case XG_OP_PLUS: case XG_OP_MINUS: case XG_OP_MULTIPLIES: case XG_OP_DIVIDES:
        need(2); r = bin_type<t_arith>(t[0], t[1]); break;
    case XG_OP_MODULUS: case XG_OP_BITWISE_AND: case XG_OP_BITWISE_OR: case XG_OP_BITWISE_XOR:
        need(2); r = bin_type<t_mod>(t[0], t[1]); break;
    case XG_OP_LEFT_SHIFT: case XG_OP_RIGHT_SHIFT: need(2); r = bin_type<t_shl>(t[0], t[1]); break;
    case XG_OP_LOGICAL_AND: case XG_OP_LOGICAL_OR: need(2); r = XG_BOOL; break;
    case XG_OP_LESS: case XG_OP_LESS_EQUAL: case XG_OP_GREATER: case XG_OP_GREATER_EQUAL: case XG_OP_EQUAL_TO:
    case XG_OP_NOT_EQUAL_TO:
        need(2);
        r = bin_type<t_arith>(t[0], t[1]) < 0 ? -1 : XG_BOOL;
        break;
    case XG_OP_MAXIMUM: case XG_OP_MINIMUM: need(2); r = bin_type<t_common>(t[0], t[1]); break;
    case XG_OP_POW: case XG_OP_FMOD: case XG_OP_ATAN2: case XG_OP_HYPOT:
        need(2); r = bin_type<t_pow>(t[0], t[1]); break;
    case XG_OP_WHERE: need(3); r = bin_type<t_common>(t[1], t[2]); break;
    case XG_OP_CLIP: {
        need(3);
        // select(lo < hi, select(v < lo, lo, select(hi < v, hi, v)), hi)
        int inner = bin_type<t_common>(t[2], t[0]);
        int mid = bin_type<t_common>(t[1], inner);
        r = bin_type<t_common>(mid, t[2]);
        break;
    }
    case XG_OP_ABS: need(1); r = un_type<t_abs>(t[0]); break;
    case XG_OP_FABS: need(1); r = un_type<t_fabs>(t[0]); break;
    case XG_OP_EXP: case XG_OP_EXP2: case XG_OP_EXPM1: case XG_OP_LOG: case XG_OP_LOG2: case XG_OP_LOG10:
    case XG_OP_LOG1P: case XG_OP_SQRT: case XG_OP_CBRT: case XG_OP_SIN: case XG_OP_COS: case XG_OP_TAN:
    case XG_OP_TANH: case XG_OP_SINH: case XG_OP_COSH: case XG_OP_ASIN: case XG_OP_ACOS: case XG_OP_ATAN:
    case XG_OP_ERF: case XG_OP_FLOOR: case XG_OP_CEIL: case XG_OP_TRUNC: case XG_OP_ROUND:
        need(1); r = un_type<t_exp>(t[0]); break;
    default: raise(XG_INVALID_ARGUMENT, "unknown op " + std::to_string(op));
    }
    if (r < 0) raise(XG_TYPE_ERROR, "operator " + std::to_string(op) + " is not defined for the operand types");
    return r;
}

// ------------------------------------------------------------------ analyse
namespace {
constexpr int64_t UNSET = std::numeric_limits<int64_t>::max();

void broadcast_into(const Shape& in, Shape& out) {
    if (out.size() < in.size()) raise(XG_BROADCAST_ERROR, "broadcast: input rank exceeds output rank");
    size_t oi = out.size(), ii = in.size();
    for (; ii != 0; --ii, --oi) {
        int64_t& o = out[oi - 1];
        int64_t i = in[ii - 1];
        if (o == UNSET || o == 1) o = i;
        else if (i == 1) {}
        else if (i != o) raise(XG_BROADCAST_ERROR, "Incompatible dimension of arrays: " + shape_str(in) + " vs " + shape_str(out));
    }
}

void dfs_leaves(const xg_expr* e, int node, std::vector<int>& order) {
    const xg_node& n = e->nodes[node];
    if (n.op == XG_OP_LEAF) { order.push_back(n.index); return; }
    for (int k = 0; k < n.nargs; ++k) dfs_leaves(e, n.args[k], order);
}
}  // namespace

bool broadcastable(const Shape& src, const Shape& dst) {
    if (dst.size() < src.size()) return false;
    for (size_t k = 0; k < src.size(); ++k) {
        int64_t s = src[src.size() - 1 - k], d = dst[dst.size() - 1 - k];
        if (!(s == d || s == 1)) return false;
    }
    return true;
}

ExprInfo analyse(const xg_expr* e) {
    if (!e || !e->nodes || e->n_nodes <= 0 || e->n_nodes > XG_MAX_NODES || e->n_leaves < 0 ||
        e->n_leaves > XG_MAX_LEAVES || e->n_consts < 0 || e->n_consts > XG_MAX_CONSTS)
        raise(XG_INVALID_ARGUMENT, "malformed expression header");
    if ((e->n_leaves && !e->leaves) || (e->n_consts && !e->consts)) raise(XG_INVALID_ARGUMENT, "missing leaves/consts");
    ExprInfo info;
    info.node_dt.resize(size_t(e->n_nodes));
    std::vector<int> rank(size_t(e->n_nodes), 0);
    for (int l = 0; l < e->n_leaves; ++l) {
        const xg_tensor& t = e->leaves[l];
        if (t.ndim < 0 || t.ndim > XG_MAX_RANK) raise(XG_INVALID_ARGUMENT, "leaf rank out of range");
        if (t.dtype < 0 || t.dtype >= XG_NUM_DTYPES) raise(XG_INVALID_ARGUMENT, "leaf dtype out of range");
        for (int k = 0; k < t.ndim; ++k)
            if (t.shape[k] < 0) raise(XG_INVALID_ARGUMENT, "negative leaf extent");
    }
    for (int c = 0; c < e->n_consts; ++c)
        if (e->consts[c].dtype < 0 || e->consts[c].dtype >= XG_NUM_DTYPES || e->consts[c].dtype == XG_BFLOAT16)
            raise(XG_INVALID_ARGUMENT, "constant dtype out of range");
    for (int i = 0; i < e->n_nodes; ++i) {
        const xg_node& n = e->nodes[i];
        if (n.nargs < 0 || n.nargs > 3) raise(XG_INVALID_ARGUMENT, "bad argument count");
        for (int k = 0; k < n.nargs; ++k)
            if (n.args[k] < 0 || n.args[k] >= i) raise(XG_INVALID_ARGUMENT, "node arguments must precede the node");
        if (n.op == XG_OP_LEAF) {
            if (n.nargs != 0 || n.index < 0 || n.index >= e->n_leaves) raise(XG_INVALID_ARGUMENT, "bad leaf node");
            info.node_dt[size_t(i)] = value_dtype(e->leaves[n.index].dtype);
            rank[size_t(i)] = e->leaves[n.index].ndim;
        } else if (n.op == XG_OP_GATHER) {
            if (n.nargs != 1 || n.index < 0 || n.index >= e->n_leaves) raise(XG_INVALID_ARGUMENT, "bad gather node");
            const int odt = info.node_dt[size_t(n.args[0])];
            if (!is_int_dt(odt) || odt == XG_BOOL) raise(XG_INVALID_ARGUMENT, "gather offsets must be integers");
            info.node_dt[size_t(i)] = value_dtype(e->leaves[n.index].dtype);
            rank[size_t(i)] = rank[size_t(n.args[0])];
        } else if (n.op == XG_OP_CONST) {
            if (n.nargs != 0 || n.index < 0 || n.index >= e->n_consts) raise(XG_INVALID_ARGUMENT, "bad constant node");
            info.node_dt[size_t(i)] = e->consts[n.index].dtype;
            rank[size_t(i)] = 0;
        } else {
            int at[3] = {0, 0, 0};
            int r = 0;
            for (int k = 0; k < n.nargs; ++k) {
                at[k] = info.node_dt[size_t(n.args[k])];
                r = std::max(r, rank[size_t(n.args[k])]);
            }
            info.node_dt[size_t(i)] = result_dtype(n.op, n.nargs, at, n.index);
            rank[size_t(i)] = r;
        }
    }
    int root = e->n_nodes - 1;
    info.dtype = info.node_dt[size_t(root)];
    info.shape.assign(size_t(rank[size_t(root)]), UNSET);
    std::vector<int> order;
    dfs_leaves(e, root, order);
    for (int l : order) {
        const xg_tensor& t = e->leaves[l];
        broadcast_into(Shape(t.shape, t.shape + t.ndim), info.shape);
    }
    return info;
}

std::vector<std::vector<int64_t>> leaf_strides(const xg_expr* e, const Shape& iter) {
    std::vector<std::vector<int64_t>> st(size_t(e->n_leaves), std::vector<int64_t>(iter.size(), 0));
    for (int l = 0; l < e->n_leaves; ++l) {
        const xg_tensor& t = e->leaves[l];
        int off = int(iter.size()) - t.ndim;
        for (int k = 0; k < t.ndim; ++k) {
            if (off + k < 0) continue;
            st[size_t(l)][size_t(off + k)] = (t.shape[k] == 1) ? 0 : t.strides[k];
        }
    }
    return st;
}

// ------------------------------------------------------------------ lowering
namespace {

struct Lowerer {
    const xg_expr* e;
    const ExprInfo& info;
    struct Op {
        MicroOp m;
        int dst, a, b, c;  // SSA ids (-1 none)
    };
    std::vector<Op> ops;
    std::vector<int> ssa_dt;
    std::vector<int> node_ssa;
    std::vector<int> leaf_ssa;
    std::vector<int> const_ssa;
    std::vector<int> leaf_order;   // expression leaf number of program leaf k
    std::vector<uint8_t> leaf_is_gather;
    std::vector<int> gather_slot;  // program leaf of an expression leaf read by MO_GATHER
    std::vector<int> const_order;

    Lowerer(const xg_expr* ex, const ExprInfo& in) : e(ex), info(in) {
        node_ssa.assign(size_t(e->n_nodes), -1);
        leaf_ssa.assign(size_t(e->n_leaves), -1);
        gather_slot.assign(size_t(e->n_leaves), -1);
        const_ssa.assign(size_t(e->n_consts), -1);
    }

    int emit(uint8_t op, int t, int rdt, int a = -1, int b = -1, int c = -1, int aux = 0) {
        Op o{};
        o.m.op = op;
        o.m.t = uint8_t(t);
        o.m.aux = uint16_t(aux);
        o.dst = int(ssa_dt.size());
        o.a = a;
        o.b = b;
        o.c = c;
        ssa_dt.push_back(rdt);
        ops.push_back(o);
        return o.dst;
    }

    // static_cast<to>(value of type from)
    int cvt(int id, int to) {
        int from = ssa_dt[size_t(id)];
        if (from == to) return id;
        return emit(MO_CVT, from, to, id, -1, -1, to);
    }

    int value(int node) {
        int& s = node_ssa[size_t(node)];
        if (s >= 0) return s;
        const xg_node& n = e->nodes[node];
        int rdt = info.node_dt[size_t(node)];
        if (n.op == XG_OP_LEAF) {
            int& ls = leaf_ssa[size_t(n.index)];
            if (ls < 0) {
                if (leaf_order.size() >= size_t(kMaxLeaves)) raise(XG_UNSUPPORTED, "too many leaves");
                ls = emit(MO_LOAD, e->leaves[n.index].dtype, rdt, -1, -1, -1, int(leaf_order.size()));
                leaf_order.push_back(n.index);
                leaf_is_gather.push_back(0);
            }
            return s = ls;
        }
        if (n.op == XG_OP_GATHER) {
            const int off = cvt(value(n.args[0]), XG_INT64);
            int& slot = gather_slot[size_t(n.index)];
            if (slot < 0) {
                if (leaf_order.size() >= size_t(kMaxLeaves)) raise(XG_UNSUPPORTED, "too many leaves");
                slot = int(leaf_order.size());
                leaf_order.push_back(n.index);
                leaf_is_gather.push_back(1);
            }
            return s = emit(MO_GATHER, e->leaves[n.index].dtype, rdt, off, -1, -1, slot);
        }
        if (n.op == XG_OP_CONST) {
            int& cs = const_ssa[size_t(n.index)];
            if (cs < 0) {
                cs = emit(MO_CONST, rdt, rdt, -1, -1, -1, int(const_order.size()));
                const_order.push_back(n.index);
            }
            return s = cs;
        }
        int a = n.nargs > 0 ? value(n.args[0]) : -1;
        int b = n.nargs > 1 ? value(n.args[1]) : -1;
        int c = n.nargs > 2 ? value(n.args[2]) : -1;
        int ta = a >= 0 ? ssa_dt[size_t(a)] : 0;
        int tb = b >= 0 ? ssa_dt[size_t(b)] : 0;
        auto arith_t = [&](int x, int y) {
            int at[2] = {x, y};
            return result_dtype(XG_OP_PLUS, 2, at, 0);
        };
        auto common_t = [&](int x, int y) {
            int at[2] = {x, y};
            return result_dtype(XG_OP_MAXIMUM, 2, at, 0);
        };
        auto select = [&](int cond, int x, int y) {  // xtl::select
            int r = common_t(ssa_dt[size_t(x)], ssa_dt[size_t(y)]);
            int xx = cvt(x, r), yy = cvt(y, r);
            return emit(MO_SEL, r, r, cond, xx, yy);
        };
        auto compare = [&](uint8_t mo, int x, int y) {
            int t = arith_t(ssa_dt[size_t(x)], ssa_dt[size_t(y)]);
            return emit(mo, t, XG_BOOL, cvt(x, t), cvt(y, t));
        };
        auto to_bool = [&](int x) { return cvt(x, XG_BOOL); };
        switch (n.op) {
        case XG_OP_IDENTITY: return s = cvt(a, rdt);
        case XG_OP_CAST: return s = cvt(a, rdt);
        case XG_OP_NEGATE: return s = emit(MO_NEG, rdt, rdt, cvt(a, rdt));
        case XG_OP_BITWISE_NOT: return s = emit(MO_BNOT, rdt, rdt, cvt(a, rdt));
        case XG_OP_LOGICAL_NOT: return s = emit(MO_LNOT, XG_BOOL, XG_BOOL, to_bool(a));
        case XG_OP_PLUS: case XG_OP_MINUS: case XG_OP_MULTIPLIES: case XG_OP_DIVIDES: case XG_OP_MODULUS:
        case XG_OP_BITWISE_AND: case XG_OP_BITWISE_OR: case XG_OP_BITWISE_XOR: {
            static const uint8_t map[] = {MO_ADD, MO_SUB, MO_MUL, MO_DIV, MO_MOD, 0, 0, MO_BAND, MO_BOR, MO_BXOR};
            uint8_t mo = map[n.op - XG_OP_PLUS];
            return s = emit(mo, rdt, rdt, cvt(a, rdt), cvt(b, rdt));
        }
        case XG_OP_LEFT_SHIFT: case XG_OP_RIGHT_SHIFT: {
            int ti[1] = {tb};
            int rt = result_dtype(XG_OP_IDENTITY, 1, ti, 0);  // promoted right operand
            return s = emit(n.op == XG_OP_LEFT_SHIFT ? MO_SHL : MO_SHR, rdt, rdt, cvt(a, rdt), cvt(b, rt));
        }
        case XG_OP_LOGICAL_AND: return s = emit(MO_LAND, XG_BOOL, XG_BOOL, to_bool(a), to_bool(b));
        case XG_OP_LOGICAL_OR: return s = emit(MO_LOR, XG_BOOL, XG_BOOL, to_bool(a), to_bool(b));
        case XG_OP_LESS: return s = compare(MO_LT, a, b);
        case XG_OP_LESS_EQUAL: return s = compare(MO_LE, a, b);
        case XG_OP_GREATER: return s = compare(MO_GT, a, b);
        case XG_OP_GREATER_EQUAL: return s = compare(MO_GE, a, b);
        case XG_OP_EQUAL_TO: return s = compare(MO_EQ, a, b);
        case XG_OP_NOT_EQUAL_TO: return s = compare(MO_NE, a, b);
        case XG_OP_MAXIMUM: return s = select(compare(MO_GT, a, b), a, b);
        case XG_OP_MINIMUM: return s = select(compare(MO_LT, a, b), a, b);
        case XG_OP_WHERE: return s = select(to_bool(a), b, c);
        case XG_OP_CLIP: {
            // clamp_fun: select(lo < hi, select(v < lo, lo, select(hi < v, hi, v)), hi)
            int v = a, lo = b, hi = c;
            int inner = select(compare(MO_LT, hi, v), hi, v);
            int mid = select(compare(MO_LT, v, lo), lo, inner);
            return s = select(compare(MO_LT, lo, hi), mid, hi);
        }
        case XG_OP_POW: case XG_OP_FMOD: case XG_OP_ATAN2: case XG_OP_HYPOT: {
            uint8_t mo = n.op == XG_OP_POW ? MO_POW : n.op == XG_OP_FMOD ? MO_FMOD : n.op == XG_OP_ATAN2 ? MO_ATAN2 : MO_HYPOT;
            return s = emit(mo, rdt, rdt, cvt(a, rdt), cvt(b, rdt));
        }
        case XG_OP_ABS:
            if (is_unsigned_int(rdt) && rdt != XG_BOOL) return s = cvt(a, rdt);  // identity
            return s = emit(MO_ABS, rdt, rdt, cvt(a, rdt));
        case XG_OP_FABS: return s = emit(MO_ABS, rdt, rdt, cvt(a, rdt));
        default: break;
        }
        static const struct { int op; uint8_t mo; } math[] = {
            {XG_OP_EXP, MO_EXP}, {XG_OP_EXP2, MO_EXP2}, {XG_OP_EXPM1, MO_EXPM1}, {XG_OP_LOG, MO_LOG},
            {XG_OP_LOG2, MO_LOG2}, {XG_OP_LOG10, MO_LOG10}, {XG_OP_LOG1P, MO_LOG1P}, {XG_OP_SQRT, MO_SQRT},
            {XG_OP_CBRT, MO_CBRT}, {XG_OP_SIN, MO_SIN}, {XG_OP_COS, MO_COS}, {XG_OP_TAN, MO_TAN},
            {XG_OP_TANH, MO_TANH}, {XG_OP_SINH, MO_SINH}, {XG_OP_COSH, MO_COSH}, {XG_OP_ASIN, MO_ASIN},
            {XG_OP_ACOS, MO_ACOS}, {XG_OP_ATAN, MO_ATAN}, {XG_OP_ERF, MO_ERF}, {XG_OP_FLOOR, MO_FLOOR},
            {XG_OP_CEIL, MO_CEIL}, {XG_OP_TRUNC, MO_TRUNC}, {XG_OP_ROUND, MO_ROUND}};
        for (auto& m : math)
            if (m.op == n.op) return s = emit(m.mo, rdt, rdt, cvt(a, rdt));
        (void)ta;
        raise(XG_UNSUPPORTED, "operator " + std::to_string(n.op) + " has no device lowering");
    }
};

}  // namespace

Program lower(const xg_expr* e, const ExprInfo& info, int final_dt, int map) {
    Lowerer L(e, info);
    int root = L.value(e->n_nodes - 1);
    int rdt = info.dtype;
    // per-element terms of the norms, reducers/xnorm.hpp:225-290
    if (map == MAP_SQUARE) {  // norm_sq(t) = t * t
        int at[2] = {rdt, rdt};
        int t = result_dtype(XG_OP_MULTIPLIES, 2, at, 0);
        int x = L.cvt(root, t);
        root = L.emit(MO_MUL, t, t, x, x);
        rdt = t;
    } else if (map == MAP_ABS) {  // std::abs(t) for signed types, t for unsigned ones
        if (!is_unsigned_int(rdt)) {
            int at[1] = {rdt};
            int t = result_dtype(XG_OP_ABS, 1, at, 0);
            root = L.emit(MO_ABS, t, t, L.cvt(root, t));
            rdt = t;
        }
    } else if (map == MAP_NONZERO) {  // t != 0
        root = L.cvt(root, XG_BOOL);
        rdt = XG_BOOL;
    }
    if (final_dt >= 0 && final_dt != XG_BFLOAT16) root = L.cvt(root, final_dt), rdt = final_dt;
    if (final_dt == XG_BFLOAT16) root = L.cvt(root, XG_FLOAT32), rdt = XG_FLOAT32;
    if (int(L.ops.size()) > kMaxOps) raise(XG_UNSUPPORTED, "expression too large for one kernel");
    // liveness: last op index that reads each SSA value
    size_t nssa = L.ssa_dt.size();
    std::vector<int> last(nssa, -1);
    for (int i = 0; i < int(L.ops.size()); ++i) {
        for (int v : {L.ops[size_t(i)].a, L.ops[size_t(i)].b, L.ops[size_t(i)].c})
            if (v >= 0) last[size_t(v)] = i;
    }
    last[size_t(root)] = int(L.ops.size());  // result stays live to the end
    // linear-scan register assignment; a destination may reuse an operand's
    // register freed by the same op (the kernels read operands before writing)
    std::vector<int> reg(nssa, -1);
    std::vector<int> free_regs;
    int nregs = 0;
    Program p{};
    for (int i = 0; i < int(L.ops.size()); ++i) {
        auto& o = L.ops[size_t(i)];
        for (int v : {o.a, o.b, o.c}) {
            if (v >= 0 && last[size_t(v)] == i && reg[size_t(v)] >= 0) {
                bool dup = std::find(free_regs.begin(), free_regs.end(), reg[size_t(v)]) != free_regs.end();
                if (!dup) free_regs.push_back(reg[size_t(v)]);
            }
        }
        int r;
        if (!free_regs.empty()) {
            std::sort(free_regs.begin(), free_regs.end(), std::greater<int>());
            r = free_regs.back();
            free_regs.pop_back();
        } else {
            r = nregs++;
        }
        if (r >= kMaxRegs) raise(XG_UNSUPPORTED, "expression needs more than " + std::to_string(kMaxRegs) + " registers");
        reg[size_t(o.dst)] = r;
        o.m.dst = uint8_t(r);
        o.m.a = uint8_t(o.a >= 0 ? reg[size_t(o.a)] : 0);
        o.m.b = uint8_t(o.b >= 0 ? reg[size_t(o.b)] : 0);
        o.m.c = uint8_t(o.c >= 0 ? reg[size_t(o.c)] : 0);
        // a value never read again frees its register right away
        if (last[size_t(o.dst)] < 0) free_regs.push_back(r);
        p.ops[i] = o.m;
    }
    p.nops = int(L.ops.size());
    p.nregs = std::max(nregs, 1);
    p.nleaves = int(L.leaf_order.size());
    p.nconsts = int(L.const_order.size());
    for (int k = 0; k < p.nleaves; ++k) {
        p.leaf_src[k] = int8_t(L.leaf_order[size_t(k)]);
        p.leaf_gather[k] = L.leaf_is_gather[size_t(k)];
    }
    for (int k = 0; k < p.nconsts; ++k) p.const_src[k] = int8_t(L.const_order[size_t(k)]);
    p.result_reg = reg[size_t(root)];
    p.result_dtype = rdt;
    // constants in register representation of their own type
    for (int c = 0; c < p.nconsts; ++c) {
        const xg_scalar& s = e->consts[p.const_src[c]];
        uint64_t bits = 0;
        switch (s.dtype) {
        case XG_FLOAT32: { float f = float(s.v.f); uint32_t u; std::memcpy(&u, &f, 4); bits = u; break; }
        case XG_FLOAT64: std::memcpy(&bits, &s.v.f, 8); break;
        case XG_BOOL: bits = s.v.i != 0; break;
        case XG_INT8: bits = uint32_t(int32_t(int8_t(s.v.i))); break;
        case XG_UINT8: bits = uint8_t(s.v.u); break;
        case XG_INT16: bits = uint32_t(int32_t(int16_t(s.v.i))); break;
        case XG_UINT16: bits = uint16_t(s.v.u); break;
        case XG_INT32: bits = uint32_t(int32_t(s.v.i)); break;
        case XG_UINT32: bits = uint32_t(s.v.u); break;
        case XG_INT64: bits = uint64_t(s.v.i); break;
        case XG_UINT64: bits = s.v.u; break;
        default: raise(XG_INVALID_ARGUMENT, "bad constant dtype");
        }
        p.consts[c] = bits;
    }
    return p;
}

int choose_vector_width(const xg_expr* e, const xg_tensor* out) {
    // each lane moves 16 bytes of the widest operand per access, so a wave's
    // access is 1 KiB of consecutive memory
    int w = out ? dtype_size(out->dtype) : 1;
    for (int l = 0; e && l < e->n_leaves; ++l) w = std::max(w, dtype_size(e->leaves[l].dtype));
    return std::max(2, std::min(16, 16 / w));
}

// ------------------------------------------------------------------ assign plan
namespace {
struct Dim {
    int64_t n;
    int64_t out;
    int64_t leaf[kMaxLeaves];
};
}  // namespace

EwPlan plan_assign(const xg_expr* e, const xg_tensor* out, int V) {
    if (!out || out->ndim < 0 || out->ndim > XG_MAX_RANK || out->dtype < 0 || out->dtype >= XG_NUM_DTYPES)
        raise(XG_INVALID_ARGUMENT, "malformed output tensor");
    ExprInfo info = analyse(e);
    Shape os(out->shape, out->shape + out->ndim);
    for (auto x : os)
        if (x < 0) raise(XG_INVALID_ARGUMENT, "negative output extent");
    if (!broadcastable(info.shape, os))
        raise(XG_BROADCAST_ERROR, "Incompatible dimension of arrays: cannot broadcast " + shape_str(info.shape) +
                                      " into " + shape_str(os));
    EwPlan p{};
    p.prog = lower(e, info, out->dtype);
    p.out_dtype = out->dtype;
    auto ls = leaf_strides(e, os);
    {  // leaves read only through MO_GATHER are not iterated: no strides
        std::vector<bool> iterated(size_t(e->n_leaves), false);
        for (int pl = 0; pl < p.prog.nleaves; ++pl)
            if (!p.prog.leaf_gather[pl]) iterated[size_t(p.prog.leaf_src[pl])] = true;
        for (int l = 0; l < e->n_leaves; ++l)
            if (!iterated[size_t(l)]) std::fill(ls[size_t(l)].begin(), ls[size_t(l)].end(), 0);
    }
    // collect dims, dropping size-1 ones
    std::vector<Dim> dims;
    int64_t total = numel(os);
    if (total == 0) { p.units = 0; return p; }
    int64_t out_off = out->offset;
    std::vector<int64_t> leaf_off(size_t(e->n_leaves));
    for (int l = 0; l < e->n_leaves; ++l) leaf_off[size_t(l)] = e->leaves[l].offset;
    for (int k = 0; k < int(os.size()); ++k) {
        if (os[size_t(k)] == 1) continue;
        Dim d{};
        d.n = os[size_t(k)];
        d.out = out->strides[k];
        for (int l = 0; l < e->n_leaves; ++l) d.leaf[l] = ls[size_t(l)][size_t(k)];
        if (d.out == 0) raise(XG_INVALID_ARGUMENT, "output view overlaps itself (zero stride)");
        if (d.out < 0) {  // iterate the dim backwards: same element pairs, positive output stride
            out_off += (d.n - 1) * d.out;
            d.out = -d.out;
            for (int l = 0; l < e->n_leaves; ++l) {
                leaf_off[size_t(l)] += (d.n - 1) * d.leaf[l];
                d.leaf[l] = -d.leaf[l];
            }
        }
        dims.push_back(d);
    }
    // outermost = largest output stride, so the output's contiguous dim is innermost
    std::stable_sort(dims.begin(), dims.end(), [](const Dim& a, const Dim& b) { return a.out > b.out; });
    // merge neighbours that are contiguous in every operand
    std::vector<Dim> merged;
    for (auto& d : dims) {
        if (!merged.empty()) {
            Dim& o = merged.back();
            bool ok = o.out == d.out * d.n;
            for (int l = 0; l < e->n_leaves && ok; ++l) ok = o.leaf[l] == d.leaf[l] * d.n;
            if (ok) {
                o.n *= d.n;
                o.out = d.out;
                for (int l = 0; l < e->n_leaves; ++l) o.leaf[l] = d.leaf[l];
                continue;
            }
        }
        merged.push_back(d);
    }
    if (merged.empty()) {  // a single element
        Dim d{};
        d.n = 1;
        d.out = 1;
        merged.push_back(d);
    }
    int rank = int(merged.size());
    if (rank > kMaxRank) raise(XG_UNSUPPORTED, "rank too large");
    p.rank = rank;
    int64_t max_off = 0;
    for (int k = 0; k < rank; ++k) {
        p.shape[k] = merged[size_t(k)].n;
        p.out_stride[k] = merged[size_t(k)].out;
        max_off += (merged[size_t(k)].n - 1) * std::abs(merged[size_t(k)].out);
    }
    int64_t inner = merged.back().n;
    int64_t rows = total / inner;
    if (inner >= (int64_t(1) << 31) || rows >= (int64_t(1) << 31))
        raise(XG_UNSUPPORTED, "iteration space too large for 32-bit unit indexing");
    p.inner = uint32_t(inner);
    p.nchunks = uint32_t((inner + V - 1) / V);
    p.div_chunks = make_fastdiv(p.nchunks);
    p.units = uint64_t(rows) * p.nchunks;
    if (p.units >= (uint64_t(1) << 32)) raise(XG_UNSUPPORTED, "iteration space too large for 32-bit unit indexing");
    for (int k = 0; k < rank; ++k) p.div[k] = make_fastdiv(uint32_t(std::min<int64_t>(merged[size_t(k)].n, 0xffffffff)));
    int osz = dtype_size(out->dtype);
    p.out_ptr = static_cast<char*>(out->data) + out_off * osz;
    auto aligned = [&](const char* ptr, int esz, int64_t inner_stride, const int64_t* strides) {
        if (inner_stride != 1) return false;
        if (reinterpret_cast<uintptr_t>(ptr) % uintptr_t(esz * V) != 0) return false;
        for (int k = 0; k + 1 < rank; ++k)
            if (strides[k] % V != 0) return false;
        return true;
    };
    p.out_mode = aligned(p.out_ptr, osz, p.out_stride[rank - 1], p.out_stride) ? LM_CONTIG : LM_STRIDED;
    if (g_nt_stores && int64_t(total) * osz >= kNtMinBytes) p.out_mode |= LM_NT;
    p.fill_zero = 0;
    if (g_fill_granules && lm_base(p.out_mode) == LM_CONTIG && rank >= 2 &&
        int64_t(total) * osz >= kNtMinBytes && p.out_stride[rank - 2] != inner) {
        // rows apart: a row whose bytes do not start and end on 64-byte
        // boundaries shares its edge granules with memory this launch leaves alone
        const bool start_ok = reinterpret_cast<uintptr_t>(p.out_ptr) % 64 == 0 && (p.out_stride[rank - 2] * osz) % 64 == 0;
        bool strides_ok = true;
        for (int k = 0; k + 2 < rank; ++k) strides_ok = strides_ok && (p.out_stride[k] * osz) % 64 == 0;
        if (!(start_ok && strides_ok) || (inner * osz) % 64 != 0) p.out_mode |= LM_FILL;
    }
    for (int pl = 0; pl < p.prog.nleaves; ++pl) {
        int l = p.prog.leaf_src[pl];
        LeafView& lv = p.leaves[pl];
        const xg_tensor& t = e->leaves[l];
        if (p.prog.leaf_gather[pl]) {  // read at computed offsets: the tensor's whole extent
            set_gather_leaf(t, &lv);
            continue;
        }
        int esz = dtype_size(t.dtype);
        lv.ptr = static_cast<const char*>(t.data) + leaf_off[size_t(l)] * esz;
        lv.dtype = t.dtype;
        int64_t lmax = 0;
        for (int k = 0; k < rank; ++k) {
            lv.stride[k] = merged[size_t(k)].leaf[l];
            lmax += (merged[size_t(k)].n - 1) * std::abs(lv.stride[k]);
        }
        max_off = std::max(max_off, lmax);
        int64_t is = lv.stride[rank - 1];
        lv.mode = is == 0 ? LM_BCAST : aligned(lv.ptr, esz, is, lv.stride) ? LM_CONTIG : LM_STRIDED;
        if (lv.mode != LM_BCAST && streams(lv.stride, p.shape, rank, esz)) lv.mode |= LM_NT;
    }
    p.index64 = max_off >= (int64_t(1) << 31) ? 1 : 0;
    p.compact = (!p.index64 && rank <= 4) ? 1 : 0;
    if (p.compact) {
        for (int k = 0; k < rank; ++k) {
            p.shape32[k] = uint32_t(p.shape[k]);
            p.cst[kMaxLeaves][k] = int32_t(p.out_stride[k]);
            for (int pl = 0; pl < p.prog.nleaves; ++pl)
                p.cst[pl][k] = p.prog.leaf_gather[pl] ? 0 : int32_t(p.leaves[pl].stride[k]);
        }
    }
    return p;
}

void set_nontemporal(int loads, int stores) {
    g_nt_loads = loads;
    g_nt_stores = stores;
}

void set_fill_granules(int on) { g_fill_granules = on; }

void set_gather_leaf(const xg_tensor& t, LeafView* lv) {
    int64_t lo = 0, hi = 0;
    bool empty = false;
    for (int k = 0; k < t.ndim; ++k) {
        if (t.shape[k] == 0) empty = true;
        const int64_t ext = (std::max<int64_t>(t.shape[k], 1) - 1) * t.strides[k];
        (ext < 0 ? lo : hi) += ext;
    }
    for (int k = 0; k < kMaxRank; ++k) lv->stride[k] = 0;
    lv->ptr = static_cast<const char*>(t.data) + t.offset * dtype_size(t.dtype);
    lv->dtype = t.dtype;
    lv->stride[0] = empty ? 1 : lo;  // an empty tensor has no valid offset
    lv->stride[1] = empty ? 0 : hi;
    lv->mode = LM_GATHER;
}

bool streams(const int64_t* stride, const int64_t* shape, int rank, int esz) {
    if (!g_nt_loads) return false;
    int64_t n = 1;
    for (int k = 0; k < rank; ++k) {
        if (shape[k] > 1 && stride[k] == 0) return false;
        n *= std::max<int64_t>(shape[k], 0);
    }
    return n * esz >= kNtMinBytes;
}

SpecKey spec_key(const EwPlan& p) {
    SpecKey k{};
    k.nops = p.prog.nops;
    for (int i = 0; i < p.prog.nops; ++i) k.ops[i] = p.prog.ops[i];
    k.nleaves = p.prog.nleaves;
    for (int l = 0; l < p.prog.nleaves; ++l) {
        k.leaf_dtype[l] = uint8_t(p.leaves[l].dtype);
        k.leaf_mode[l] = uint8_t(lm_base(p.leaves[l].mode));
    }
    k.out_dtype = uint8_t(p.out_dtype);
    k.out_mode = uint8_t(lm_base(p.out_mode));
    k.result_reg = p.prog.result_reg;
    return k;
}

bool spec_equal(const SpecKey& a, const SpecKey& b) {
    if (a.nops != b.nops || a.nleaves != b.nleaves || a.out_dtype != b.out_dtype || a.out_mode != b.out_mode ||
        a.result_reg != b.result_reg)
        return false;
    if (std::memcmp(a.ops, b.ops, sizeof(MicroOp) * size_t(a.nops)) != 0) return false;
    for (int l = 0; l < a.nleaves; ++l)
        if (a.leaf_dtype[l] != b.leaf_dtype[l] || a.leaf_mode[l] != b.leaf_mode[l]) return false;
    return true;
}

}  // namespace xg

namespace xg {
int mapped_dtype(int map, int dt) {
    if (map == MAP_SQUARE) {
        int at[2] = {dt, dt};
        return result_dtype(XG_OP_MULTIPLIES, 2, at, 0);
    }
    if (map == MAP_ABS) {
        if (is_unsigned_int(dt)) return dt;
        int at[1] = {dt};
        return result_dtype(XG_OP_ABS, 1, at, 0);
    }
    if (map == MAP_NONZERO) return XG_BOOL;
    return dt;
}
}  // namespace xg
