// xdevice.hpp — C++ front end of the MI355X evaluator: xtensor's expression
// vocabulary over device arrays, flattened into the C-ABI descriptor of
// include/xg.h.
//
// Mirrors the pieces of xtensor a drop-in needs (the expression tag,
// a device container, lazy functions and scalars, strided views,
// noalias assignment, reducers, cumsum):
//   xt::noalias(c) = a * b + xt::exp(-a * a);     // core/xnoalias.hpp:160-165
//   auto s = xt::eval(xt::sum(c, {1}));           // core/xmath.hpp:1803
//   xt::noalias(xt::view(y, xt::all(), xt::range(0, 56, 2))) = f(x);   // views/xview.hpp
//   auto cs = xt::cumsum(c, 1);                   // core/xmath.hpp:2247-2266
// Every evaluation is one call of xg_assign / xg_reduce / xg_scan on the
// front end's stream; element access from the host goes through to_host(),
// which synchronises first, so host reads see the same values as xtensor's
// synchronous evaluation would produce.
//
// Scalars keep their C++ type, as xt::xscalar does (containers/xscalar.hpp):
// `1e-5` is a double and `3` an int, so promotion follows C++ exactly.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <initializer_list>
#include <memory>
#include <stdexcept>
#include <string>
#include <tuple>
#include <type_traits>
#include <utility>
#include <vector>

#include "xg.h"

namespace xtensor_amd {

// ---------------------------------------------------------------- errors
// xt::broadcast_error (utils/xexception.hpp:63)
class broadcast_error : public std::runtime_error {
  public:
    using std::runtime_error::runtime_error;
};

inline void check(xg_status st) {
    if (st == XG_OK) return;
    const std::string m = xg_last_error_message();
    switch (st) {
        case XG_BROADCAST_ERROR: throw broadcast_error(m);
        case XG_BAD_AXES: throw std::runtime_error(m);
        case XG_TYPE_ERROR: throw std::invalid_argument(m);
        case XG_OUT_OF_MEMORY: throw std::bad_alloc();
        default: throw std::runtime_error(m);
    }
}

inline void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

// stream every evaluation is enqueued on (nullptr: the null stream)
inline xg_stream_t& current_stream() {
    static thread_local xg_stream_t s = nullptr;
    return s;
}

// ---------------------------------------------------------------- dtypes
template <class T> struct dtype_of;
template <> struct dtype_of<bool> { static constexpr int value = XG_BOOL; };
template <> struct dtype_of<int8_t> { static constexpr int value = XG_INT8; };
template <> struct dtype_of<uint8_t> { static constexpr int value = XG_UINT8; };
template <> struct dtype_of<int16_t> { static constexpr int value = XG_INT16; };
template <> struct dtype_of<uint16_t> { static constexpr int value = XG_UINT16; };
template <> struct dtype_of<int32_t> { static constexpr int value = XG_INT32; };
template <> struct dtype_of<uint32_t> { static constexpr int value = XG_UINT32; };
template <> struct dtype_of<int64_t> { static constexpr int value = XG_INT64; };
template <> struct dtype_of<uint64_t> { static constexpr int value = XG_UINT64; };
template <> struct dtype_of<float> { static constexpr int value = XG_FLOAT32; };
template <> struct dtype_of<double> { static constexpr int value = XG_FLOAT64; };
template <> struct dtype_of<long long> : dtype_of<int64_t> {};
template <> struct dtype_of<unsigned long long> : dtype_of<uint64_t> {};

using shape_type = std::vector<int64_t>;

// ---------------------------------------------------------------- descriptor
// The flattened tree: nodes in post-order, leaves and constants numbered in
// order of appearance.
class descriptor {
  public:
    int leaf(const xg_tensor& t) {
        return node(XG_OP_LEAF, {}, leaf_slot(t));
    }
    // a leaf that no node reads directly (the source of a gather)
    int leaf_slot(const xg_tensor& t) {
        leaves_.push_back(t);
        return int(leaves_.size()) - 1;
    }
    template <class T> int constant(T v) {
        xg_scalar s{};
        s.dtype = dtype_of<T>::value;
        if constexpr (std::is_floating_point_v<T>) s.v.f = double(v);
        else if constexpr (std::is_signed_v<T>) s.v.i = int64_t(v);
        else s.v.u = uint64_t(v);
        consts_.push_back(s);
        return node(XG_OP_CONST, {}, int(consts_.size()) - 1);
    }
    int node(int op, std::initializer_list<int> args, int index = 0) {
        xg_node n{};
        n.op = op;
        n.nargs = int(args.size());
        int k = 0;
        for (int a : args) n.args[k++] = a;
        n.index = index;
        nodes_.push_back(n);
        return int(nodes_.size()) - 1;
    }
    xg_expr expr() const {
        xg_expr e{};
        e.n_nodes = int(nodes_.size());
        e.n_leaves = int(leaves_.size());
        e.n_consts = int(consts_.size());
        e.nodes = nodes_.data();
        e.leaves = leaves_.data();
        e.consts = consts_.data();
        return e;
    }

  private:
    std::vector<xg_node> nodes_;
    std::vector<xg_tensor> leaves_;
    std::vector<xg_scalar> consts_;
};

// ---------------------------------------------------------------- expressions
// the expression tag of device expressions (core/xexpression.hpp:303-393)
struct xdevice_expression_tag {};

template <class D> struct xdev_expression {
    using expression_tag = xdevice_expression_tag;
    const D& derived_cast() const { return static_cast<const D&>(*this); }
};

template <class E> using is_dev_expr = std::is_base_of<xdev_expression<E>, E>;

// a scalar leaf: its C++ type drives promotion (containers/xscalar.hpp)
template <class T> struct xdev_scalar : xdev_expression<xdev_scalar<T>> {
    T value;
    explicit xdev_scalar(T v) : value(v) {}
    int flatten(descriptor& d) const { return d.constant(value); }
};

// a lazy n-ary function node (core/xfunction.hpp): opcode + arguments
template <int OP, class... E> struct xdev_function : xdev_expression<xdev_function<OP, E...>> {
    std::tuple<E...> args;
    explicit xdev_function(E... a) : args(std::move(a)...) {}
    int flatten(descriptor& d) const {
        return std::apply([&](const E&... a) { return d.node(OP, {a.flatten(d)...}); }, args);
    }
};

template <class T> class xdevice_array;

// scalars become xdev_scalar, expressions stay themselves
template <class X> auto as_expr(const X& x) {
    if constexpr (std::is_arithmetic_v<X>) return xdev_scalar<X>(x);
    else return x;
}

// ---------------------------------------------------------------- container
// Device storage (hipMalloc), shape and strides in elements, an offset; views
// share the storage.  Element access from the host is only through
// to_host() / from_host().
template <class T> class xdevice_array : public xdev_expression<xdevice_array<T>> {
  public:
    using value_type = T;

    xdevice_array() = default;
    explicit xdevice_array(shape_type shape) : shape_(std::move(shape)) {
        strides_ = row_major(shape_);
        const size_t n = size_t(size());
        void* p = nullptr;
        if (n) hip_check(hipMalloc(&p, n * sizeof(T)), "hipMalloc");
        mem_ = std::shared_ptr<void>(p, [](void* q) {
            if (q) (void)hipFree(q);
        });
        data_ = static_cast<T*>(p);
    }
    // wrap memory the caller owns (xt::adapt, containers/xadapt.hpp)
    static xdevice_array adapt(T* data, shape_type shape, shape_type strides = {}, int64_t offset = 0) {
        xdevice_array a;
        a.data_ = data;
        a.shape_ = std::move(shape);
        a.strides_ = strides.empty() ? row_major(a.shape_) : std::move(strides);
        a.offset_ = offset;
        return a;
    }
    static xdevice_array from_host(const std::vector<T>& h, shape_type shape) {
        xdevice_array a(std::move(shape));
        if (int64_t(h.size()) != a.size()) throw std::invalid_argument("host data size mismatch");
        if (!h.empty()) hip_check(hipMemcpy(a.data_, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice), "hipMemcpy");
        return a;
    }
    // values in row-major order (synchronises the current stream first)
    std::vector<T> to_host() const {
        hip_check(hipStreamSynchronize(reinterpret_cast<hipStream_t>(current_stream())), "hipStreamSynchronize");
        std::vector<T> out(static_cast<size_t>(size()));
        if (out.empty()) return out;
        int64_t lo = offset_, hi = offset_;
        for (size_t k = 0; k < shape_.size(); ++k) {
            const int64_t ext = (shape_[k] - 1) * strides_[k];
            (ext < 0 ? lo : hi) += ext;
        }
        std::vector<T> span(size_t(hi - lo + 1));
        hip_check(hipMemcpy(span.data(), data_ + lo, span.size() * sizeof(T), hipMemcpyDeviceToHost), "hipMemcpy");
        std::vector<int64_t> idx(shape_.size(), 0);
        for (size_t i = 0; i < out.size(); ++i) {
            int64_t off = offset_ - lo;
            for (size_t k = 0; k < shape_.size(); ++k) off += idx[k] * strides_[k];
            out[i] = span[size_t(off)];
            for (int k = int(shape_.size()) - 1; k >= 0; --k) {
                if (++idx[size_t(k)] < shape_[size_t(k)]) break;
                idx[size_t(k)] = 0;
            }
        }
        return out;
    }

    const shape_type& shape() const { return shape_; }
    const shape_type& strides() const { return strides_; }
    int64_t offset() const { return offset_; }
    T* data() const { return data_; }
    int64_t size() const {
        int64_t n = 1;
        for (int64_t s : shape_) n *= s;
        return n;
    }
    size_t dimension() const { return shape_.size(); }

    xg_tensor tensor() const {
        if (shape_.size() > XG_MAX_RANK) throw std::invalid_argument("rank above 8");
        xg_tensor t{};
        t.data = data_;
        t.dtype = dtype_of<T>::value;
        t.ndim = int32_t(shape_.size());
        t.offset = offset_;
        for (size_t k = 0; k < shape_.size(); ++k) {
            t.shape[k] = shape_[k];
            t.strides[k] = strides_[k];
        }
        return t;
    }
    int flatten(descriptor& d) const { return d.leaf(tensor()); }

    // a view sharing this storage
    xdevice_array reshaped_view(shape_type shape, shape_type strides, int64_t offset) const {
        xdevice_array v = *this;
        v.shape_ = std::move(shape);
        v.strides_ = std::move(strides);
        v.offset_ = offset;
        return v;
    }

    static shape_type row_major(const shape_type& s) {
        shape_type st(s.size(), 1);
        for (int k = int(s.size()) - 2; k >= 0; --k) st[size_t(k)] = st[size_t(k) + 1] * s[size_t(k) + 1];
        return st;
    }

  private:
    std::shared_ptr<void> mem_;
    T* data_ = nullptr;
    int64_t offset_ = 0;
    shape_type shape_;
    shape_type strides_;
};

// ---------------------------------------------------------------- operators
#define XG_DEV_BINARY(NAME, OP)                                                                        \
    template <class A, class B,                                                                      \
              std::enable_if_t<is_dev_expr<A>::value || is_dev_expr<B>::value, int> = 0>             \
    auto NAME(const A& a, const B& b) {                                                              \
        auto x = as_expr(a);                                                                         \
        auto y = as_expr(b);                                                                         \
        return xdev_function<OP, decltype(x), decltype(y)>(x, y);                                    \
    }
XG_DEV_BINARY(operator+, XG_OP_PLUS)
XG_DEV_BINARY(operator-, XG_OP_MINUS)
XG_DEV_BINARY(operator*, XG_OP_MULTIPLIES)
XG_DEV_BINARY(operator/, XG_OP_DIVIDES)
XG_DEV_BINARY(operator%, XG_OP_MODULUS)
XG_DEV_BINARY(operator&, XG_OP_BITWISE_AND)
XG_DEV_BINARY(operator|, XG_OP_BITWISE_OR)
XG_DEV_BINARY(operator^, XG_OP_BITWISE_XOR)
XG_DEV_BINARY(operator<<, XG_OP_LEFT_SHIFT)
XG_DEV_BINARY(operator>>, XG_OP_RIGHT_SHIFT)
XG_DEV_BINARY(operator&&, XG_OP_LOGICAL_AND)
XG_DEV_BINARY(operator||, XG_OP_LOGICAL_OR)
XG_DEV_BINARY(operator<, XG_OP_LESS)
XG_DEV_BINARY(operator<=, XG_OP_LESS_EQUAL)
XG_DEV_BINARY(operator>, XG_OP_GREATER)
XG_DEV_BINARY(operator>=, XG_OP_GREATER_EQUAL)
XG_DEV_BINARY(equal, XG_OP_EQUAL_TO)
XG_DEV_BINARY(not_equal, XG_OP_NOT_EQUAL_TO)
XG_DEV_BINARY(maximum, XG_OP_MAXIMUM)
XG_DEV_BINARY(minimum, XG_OP_MINIMUM)
XG_DEV_BINARY(pow, XG_OP_POW)
XG_DEV_BINARY(fmod, XG_OP_FMOD)
XG_DEV_BINARY(atan2, XG_OP_ATAN2)
XG_DEV_BINARY(hypot, XG_OP_HYPOT)
#undef XG_DEV_BINARY

#define XG_DEV_UNARY(NAME, OP)                                                                         \
    template <class A, std::enable_if_t<is_dev_expr<A>::value, int> = 0> auto NAME(const A& a) {     \
        return xdev_function<OP, A>(a);                                                              \
    }
XG_DEV_UNARY(operator-, XG_OP_NEGATE)
XG_DEV_UNARY(operator+, XG_OP_IDENTITY)
XG_DEV_UNARY(operator!, XG_OP_LOGICAL_NOT)
XG_DEV_UNARY(operator~, XG_OP_BITWISE_NOT)
XG_DEV_UNARY(abs, XG_OP_ABS)
XG_DEV_UNARY(fabs, XG_OP_FABS)
XG_DEV_UNARY(exp, XG_OP_EXP)
XG_DEV_UNARY(exp2, XG_OP_EXP2)
XG_DEV_UNARY(expm1, XG_OP_EXPM1)
XG_DEV_UNARY(log, XG_OP_LOG)
XG_DEV_UNARY(log2, XG_OP_LOG2)
XG_DEV_UNARY(log10, XG_OP_LOG10)
XG_DEV_UNARY(log1p, XG_OP_LOG1P)
XG_DEV_UNARY(sqrt, XG_OP_SQRT)
XG_DEV_UNARY(cbrt, XG_OP_CBRT)
XG_DEV_UNARY(sin, XG_OP_SIN)
XG_DEV_UNARY(cos, XG_OP_COS)
XG_DEV_UNARY(tan, XG_OP_TAN)
XG_DEV_UNARY(tanh, XG_OP_TANH)
XG_DEV_UNARY(sinh, XG_OP_SINH)
XG_DEV_UNARY(cosh, XG_OP_COSH)
XG_DEV_UNARY(asin, XG_OP_ASIN)
XG_DEV_UNARY(acos, XG_OP_ACOS)
XG_DEV_UNARY(atan, XG_OP_ATAN)
XG_DEV_UNARY(erf, XG_OP_ERF)
XG_DEV_UNARY(floor, XG_OP_FLOOR)
XG_DEV_UNARY(ceil, XG_OP_CEIL)
XG_DEV_UNARY(trunc, XG_OP_TRUNC)
XG_DEV_UNARY(round, XG_OP_ROUND)
#undef XG_DEV_UNARY

// xt::where (core/xoperation.hpp:775-784) and xt::clip (core/xmath.hpp:814-823)
template <class C, class A, class B> auto where(const C& c, const A& a, const B& b) {
    auto x = as_expr(c);
    auto y = as_expr(a);
    auto z = as_expr(b);
    return xdev_function<XG_OP_WHERE, decltype(x), decltype(y), decltype(z)>(x, y, z);
}
template <class E, class L, class H> auto clip(const E& e, const L& lo, const H& hi) {
    auto x = as_expr(e);
    auto y = as_expr(lo);
    auto z = as_expr(hi);
    return xdev_function<XG_OP_CLIP, decltype(x), decltype(y), decltype(z)>(x, y, z);
}

// xt::cast<T> (core/xoperation.hpp:145-165): the target dtype rides in the node index
template <class T, class E> struct xdev_cast : xdev_expression<xdev_cast<T, E>> {
    E arg;
    explicit xdev_cast(E a) : arg(std::move(a)) {}
    int flatten(descriptor& d) const {
        const int a = arg.flatten(d);
        return d.node(XG_OP_CAST, {a}, dtype_of<T>::value);
    }
};
template <class T, class E, std::enable_if_t<is_dev_expr<E>::value, int> = 0> auto cast(const E& e) {
    return xdev_cast<T, E>(e);
}

// ---------------------------------------------------------------- views
// xt::range / xt::all / xt::newaxis (views/xslice.hpp), xt::view (views/xview.hpp)
struct xrange {
    int64_t start, stop, step;
};
struct xall_tag {};
struct xnewaxis_tag {};
inline xrange range(int64_t start, int64_t stop, int64_t step = 1) { return {start, stop, step}; }
inline xall_tag all() { return {}; }
inline xnewaxis_tag newaxis() { return {}; }

namespace detail {
template <class T> void apply_slices(const xdevice_array<T>&, size_t, shape_type&, shape_type&, int64_t&) {}
template <class T, class S, class... R>
void apply_slices(const xdevice_array<T>& a, size_t k, shape_type& sh, shape_type& st, int64_t& off, const S& s,
                  const R&... rest) {
    if constexpr (std::is_same_v<S, xnewaxis_tag>) {
        sh.push_back(1);
        st.push_back(0);
        apply_slices(a, k, sh, st, off, rest...);
    } else {
        if (k >= a.dimension()) throw std::out_of_range("too many slices");
        const int64_t n = a.shape()[k], stride = a.strides()[k];
        if constexpr (std::is_same_v<S, xall_tag>) {
            sh.push_back(n);
            st.push_back(stride);
        } else if constexpr (std::is_same_v<S, xrange>) {
            if (s.step == 0) throw std::invalid_argument("range step 0");
            auto clampi = [&](int64_t v, int64_t lo, int64_t hi) { return v < lo ? lo : (v > hi ? hi : v); };
            int64_t b = s.start < 0 ? s.start + n : s.start;
            int64_t e = s.stop < 0 ? s.stop + n : s.stop;
            int64_t cnt;
            if (s.step > 0) {
                b = clampi(b, 0, n);
                e = clampi(e, 0, n);
                cnt = e > b ? (e - b + s.step - 1) / s.step : 0;
            } else {
                b = clampi(b, -1, n - 1);
                e = clampi(e, -1, n - 1);
                cnt = b > e ? (b - e - s.step - 1) / (-s.step) : 0;
            }
            off += b * stride;
            sh.push_back(cnt);
            st.push_back(stride * s.step);
        } else {  // an integer index drops the dimension
            const int64_t i = int64_t(s) < 0 ? int64_t(s) + n : int64_t(s);
            if (i < 0 || i >= n) throw std::out_of_range("view index out of range");
            off += i * stride;
        }
        apply_slices(a, k + 1, sh, st, off, rest...);
    }
}
}  // namespace detail

template <class T, class... S> xdevice_array<T> view(const xdevice_array<T>& a, const S&... slices) {
    shape_type sh, st;
    int64_t off = a.offset();
    detail::apply_slices(a, 0, sh, st, off, slices...);
    size_t consumed = 0;
    ((consumed += std::is_same_v<S, xnewaxis_tag> ? 0 : 1), ...);
    for (size_t k = consumed; k < a.dimension(); ++k) {  // missing trailing slices are all()
        sh.push_back(a.shape()[k]);
        st.push_back(a.strides()[k]);
    }
    return a.reshaped_view(sh, st, off);
}

// xt::transpose (misc/xmanipulation.hpp:238-259): reversed shape and strides, no copy
template <class T> xdevice_array<T> transpose(const xdevice_array<T>& a) {
    shape_type sh(a.shape().rbegin(), a.shape().rend()), st(a.strides().rbegin(), a.strides().rend());
    return a.reshaped_view(sh, st, a.offset());
}

// ---------------------------------------------------------------- reducers
// A lazy reducer (reducers/xreducer.hpp): evaluated when assigned or eval'd.
template <class E> struct xdev_reducer {
    E expr;
    xg_reduce_desc desc{};
};

namespace detail {
template <class E>
xdev_reducer<E> make_reducer(int op, const E& e, std::initializer_list<int> axes, bool keep_dims, int acc_dtype,
                             bool all_axes) {
    xdev_reducer<E> r{e, {}};
    r.desc.op = op;
    r.desc.acc_dtype = acc_dtype;
    r.desc.n_axes = all_axes ? -1 : int32_t(axes.size());
    int k = 0;
    for (int a : axes) r.desc.axes[k++] = a;
    r.desc.keep_dims = keep_dims ? 1 : 0;
    return r;
}
}  // namespace detail

// keep_dims mirrors xt::keep_dims (reducers/xreducer.hpp:35-161)
struct keep_dims_t {};
inline constexpr keep_dims_t keep_dims{};

#define XG_DEV_REDUCER(NAME, OP)                                                                       \
    template <class T = void, class E, std::enable_if_t<is_dev_expr<E>::value, int> = 0>             \
    auto NAME(const E& e, std::initializer_list<int> axes) {                                         \
        return detail::make_reducer(OP, e, axes, false, acc_code<T>(), false);                       \
    }                                                                                                \
    template <class T = void, class E, std::enable_if_t<is_dev_expr<E>::value, int> = 0>             \
    auto NAME(const E& e, std::initializer_list<int> axes, keep_dims_t) {                            \
        return detail::make_reducer(OP, e, axes, true, acc_code<T>(), false);                        \
    }                                                                                                \
    template <class T = void, class E, std::enable_if_t<is_dev_expr<E>::value, int> = 0>             \
    auto NAME(const E& e) {                                                                          \
        return detail::make_reducer(OP, e, {}, false, acc_code<T>(), true);                          \
    }

// the T of sum<T>(e, axes): the accumulator / result type (core/xmath.hpp:367-423)
template <class T> constexpr int acc_code() {
    if constexpr (std::is_void_v<T>) return -1;
    else return dtype_of<T>::value;
}

XG_DEV_REDUCER(sum, XG_RED_SUM)
XG_DEV_REDUCER(prod, XG_RED_PROD)
XG_DEV_REDUCER(amax, XG_RED_AMAX)
XG_DEV_REDUCER(amin, XG_RED_AMIN)
XG_DEV_REDUCER(mean, XG_RED_MEAN)
XG_DEV_REDUCER(norm_sq, XG_RED_NORM_SQ)
XG_DEV_REDUCER(norm_l2, XG_RED_NORM_L2)
XG_DEV_REDUCER(norm_l1, XG_RED_NORM_L1)
XG_DEV_REDUCER(norm_linf, XG_RED_NORM_LINF)
XG_DEV_REDUCER(variance, XG_RED_VARIANCE)
XG_DEV_REDUCER(stddev, XG_RED_STDDEV)
#undef XG_DEV_REDUCER

// xt::variance<T>(e, axes, ddof) (core/xmath.hpp:2082-2103): the divisor is
// (size - ddof) / outputs
template <class T = void, class E, std::enable_if_t<is_dev_expr<E>::value, int> = 0>
auto variance(const E& e, std::initializer_list<int> axes, int ddof) {
    auto r = detail::make_reducer(XG_RED_VARIANCE, e, axes, false, acc_code<T>(), false);
    r.desc.ddof = ddof;
    return r;
}

// xt::minmax(e) (core/xmath.hpp:2195-2230): {min, max} of every element as a
// [2] result of the value type
template <class E, std::enable_if_t<is_dev_expr<E>::value, int> = 0> auto minmax(const E& e) {
    return detail::make_reducer(XG_RED_MINMAX, e, {}, false, -1, true);
}

// ---------------------------------------------------------------- evaluation
template <class E> xg_expr flatten_expr(const E& e, descriptor& d) {
    as_expr(e).flatten(d);
    return d.expr();
}

// dtype code and broadcast shape of an expression (xfunction::shape / value_type)
template <class E> std::pair<int, shape_type> expr_info(const E& e) {
    descriptor d;
    xg_expr x = flatten_expr(e, d);
    int32_t dt = 0, nd = 0;
    int64_t shp[XG_MAX_RANK];
    check(xg_expr_info(&x, &dt, &nd, shp));
    return {dt, shape_type(shp, shp + nd)};
}

template <class E> std::pair<int, shape_type> reduce_info(const xdev_reducer<E>& r) {
    descriptor d;
    xg_expr x = flatten_expr(r.expr, d);
    int32_t dt = 0, nd = 0;
    int64_t shp[XG_MAX_RANK];
    check(xg_reduce_info(&x, &r.desc, &dt, &nd, shp));
    return {dt, shape_type(shp, shp + nd)};
}

// xt::noalias (core/xnoalias.hpp:18-227): assignment without a temporary
template <class T> class noalias_proxy {
  public:
    explicit noalias_proxy(const xdevice_array<T>& out) : out_(out) {}

    template <class E, std::enable_if_t<is_dev_expr<E>::value, int> = 0> noalias_proxy& operator=(const E& e) {
        descriptor d;
        xg_expr x = flatten_expr(e, d);
        xg_tensor t = out_.tensor();
        check(xg_assign(&x, &t, current_stream()));
        return *this;
    }
    template <class E> noalias_proxy& operator=(const xdev_reducer<E>& r) {
        descriptor d;
        xg_expr x = flatten_expr(r.expr, d);
        xg_tensor t = out_.tensor();
        check(xg_reduce(&x, &r.desc, &t, current_stream()));
        return *this;
    }
    template <class E> noalias_proxy& operator+=(const E& e) { return *this = out_ + e; }
    template <class E> noalias_proxy& operator-=(const E& e) { return *this = out_ - e; }
    template <class E> noalias_proxy& operator*=(const E& e) { return *this = out_ * e; }
    template <class E> noalias_proxy& operator/=(const E& e) { return *this = out_ / e; }

  private:
    xdevice_array<T> out_;
};

template <class T> noalias_proxy<T> noalias(const xdevice_array<T>& out) { return noalias_proxy<T>(out); }

// xt::eval (core/xeval.hpp:45-57) of a reducer into a new array of its
// result type; T must be that type
template <class T, class E> xdevice_array<T> eval(const xdev_reducer<E>& r) {
    auto [dt, shape] = reduce_info(r);
    if (dt != dtype_of<T>::value) throw std::invalid_argument("eval<T>: T differs from the reducer's value type");
    xdevice_array<T> out(shape);
    noalias(out) = r;
    return out;
}

template <class T, class E, std::enable_if_t<is_dev_expr<E>::value, int> = 0> xdevice_array<T> eval(const E& e) {
    auto [dt, shape] = expr_info(e);
    (void)dt;
    xdevice_array<T> out(shape);
    noalias(out) = e;
    return out;
}

// xt::cumsum / xt::cumprod (reducers/xaccumulator.hpp:343-367) with an axis;
// T is the result type (the expression's value type unless given)
template <class T, class E> xdevice_array<T> scan(int op, const E& e, int axis) {
    descriptor d;
    xg_expr x = flatten_expr(e, d);
    xg_scan_desc s{};
    s.op = op;
    s.acc_dtype = -1;
    s.axis = axis;
    int32_t dt = 0, nd = 0;
    int64_t shp[XG_MAX_RANK];
    check(xg_scan_info(&x, &s, &dt, &nd, shp));
    if (dt != dtype_of<T>::value) s.acc_dtype = dtype_of<T>::value;
    xdevice_array<T> out(shape_type(shp, shp + nd));
    xg_tensor t = out.tensor();
    check(xg_scan(&x, &s, &t, current_stream()));
    return out;
}
template <class T, class E> xdevice_array<T> cumsum(const E& e, int axis) { return scan<T>(XG_SCAN_CUMSUM, e, axis); }
template <class T, class E> xdevice_array<T> cumprod(const E& e, int axis) { return scan<T>(XG_SCAN_CUMPROD, e, axis); }

// ---------------------------------------------------------------- index views
// xt::index_view(e, indices) (views/xindex_view.hpp:786-797): the 1-d
// sequence of e's elements at the multi-indices `indices` ([n, e.dimension()],
// int64, on the device).  As an operand it is a gather fused into the
// expression; assigning to it scatters (xg_assign_indexed).  Indices outside
// e read as 0 and are skipped on assignment.
template <class T> struct xdev_index_view : xdev_expression<xdev_index_view<T>> {
    using value_type = T;
    xdevice_array<T> source;
    xdevice_array<int64_t> index;

    xdev_index_view(xdevice_array<T> s, xdevice_array<int64_t> i) : source(std::move(s)), index(std::move(i)) {
        if (index.dimension() != 2 || size_t(index.shape()[1]) != source.dimension() || source.dimension() == 0)
            throw std::invalid_argument("index_view: indices must be [n, e.dimension()]");
    }
    int64_t size() const { return index.shape()[0]; }
    // element offsets: sum over k of indices[:, k] * stride_k
    int offsets(descriptor& d) const {
        int off = -1;
        for (size_t k = 0; k < source.dimension(); ++k) {
            auto col = index.reshaped_view({index.shape()[0]}, {index.strides()[0]},
                                           index.offset() + int64_t(k) * index.strides()[1]);
            const int t = d.node(XG_OP_MULTIPLIES, {col.flatten(d), d.constant(int64_t(source.strides()[k]))});
            off = off < 0 ? t : d.node(XG_OP_PLUS, {off, t});
        }
        return off;
    }
    int flatten(descriptor& d) const {
        const int off = offsets(d);
        return d.node(XG_OP_GATHER, {off}, d.leaf_slot(source.tensor()));
    }
    template <class E> xdev_index_view& operator=(const E& e) {
        descriptor dv, doff;
        xg_expr xv = flatten_expr(as_expr(e), dv);
        offsets(doff);
        xg_expr xo = doff.expr();
        xg_tensor t = source.tensor();
        check(xg_assign_indexed(&xv, &xo, &t, current_stream()));
        return *this;
    }
    xdev_index_view& operator=(const xdev_index_view& o) { return this->operator=<xdev_index_view>(o); }
    template <class E> xdev_index_view& operator+=(const E& e) { return *this = *this + e; }
    template <class E> xdev_index_view& operator-=(const E& e) { return *this = *this - e; }
    template <class E> xdev_index_view& operator*=(const E& e) { return *this = *this * e; }
    template <class E> xdev_index_view& operator/=(const E& e) { return *this = *this / e; }
    void fill(T v) { *this = v; }
};

template <class T> xdev_index_view<T> index_view(const xdevice_array<T>& e, const xdevice_array<int64_t>& indices) {
    return xdev_index_view<T>(e, indices);
}
template <class T>
xdev_index_view<T> index_view(const xdevice_array<T>& e, const std::vector<std::vector<int64_t>>& indices) {
    std::vector<int64_t> flat;
    for (const auto& ix : indices) {
        if (ix.size() != e.dimension()) throw std::invalid_argument("index_view: index of the wrong rank");
        flat.insert(flat.end(), ix.begin(), ix.end());
    }
    return xdev_index_view<T>(e, xdevice_array<int64_t>::from_host(flat, {int64_t(indices.size()), int64_t(e.dimension())}));
}

// xt::argwhere (row-major order) of a device condition: positions listed on the host
template <class C> std::vector<std::vector<int64_t>> argwhere(const C& cond) {
    auto c = eval<uint8_t>(cond);
    const std::vector<uint8_t> h = c.to_host();
    std::vector<std::vector<int64_t>> out;
    std::vector<int64_t> idx(c.dimension(), 0);
    for (size_t i = 0; i < h.size(); ++i) {
        if (h[i]) out.push_back(idx);
        for (int k = int(idx.size()) - 1; k >= 0; --k) {
            if (++idx[size_t(k)] < c.shape()[size_t(k)]) break;
            idx[size_t(k)] = 0;
        }
    }
    return out;
}

// xt::filter(e, cond) = index_view(e, argwhere(cond)) (views/xindex_view.hpp:820-826)
template <class T, class C> xdev_index_view<T> filter(const xdevice_array<T>& e, const C& cond) {
    auto pos = argwhere(cond);
    if (pos.empty()) return xdev_index_view<T>(e, xdevice_array<int64_t>({0, int64_t(e.dimension())}));
    return index_view(e, pos);
}

// xt::filtration(e, cond) (views/xindex_view.hpp:845-850) and masked_view
// assignment (views/xmasked_view.hpp): e = where(cond, value, e), one launch
template <class T, class C> struct xdev_filtration {
    xdevice_array<T> e;
    C cond;
    template <class V> xdev_filtration& operator=(const V& v) {
        noalias(e) = where(cond, as_expr(v), e);
        return *this;
    }
    template <class V> xdev_filtration& operator+=(const V& v) { noalias(e) = where(cond, e + v, e); return *this; }
    template <class V> xdev_filtration& operator-=(const V& v) { noalias(e) = where(cond, e - v, e); return *this; }
    template <class V> xdev_filtration& operator*=(const V& v) { noalias(e) = where(cond, e * v, e); return *this; }
    template <class V> xdev_filtration& operator/=(const V& v) { noalias(e) = where(cond, e / v, e); return *this; }
};
template <class T, class C> xdev_filtration<T, C> filtration(const xdevice_array<T>& e, const C& cond) {
    return xdev_filtration<T, C>{e, cond};
}
template <class T, class C> xdev_filtration<T, C> masked_view(const xdevice_array<T>& e, const C& mask) {
    return xdev_filtration<T, C>{e, mask};
}

// ---------------------------------------------------------------- .npy files
// xt::load_npy<T>(path) (io/xnpy.hpp:775-800): the file's data in a new device
// array; a Fortran-order file gives column-major strides, as the reference
// adapts its buffer.  The file's type string must be T's ("Cast error").
template <class T> xdevice_array<T> load_npy(const std::string& path) {
    int32_t dt = 0, nd = 0, fortran = 0;
    int64_t shp[XG_MAX_RANK];
    check(xg_npy_info(path.c_str(), &dt, &nd, shp, &fortran));
    if (dt != dtype_of<T>::value) throw std::runtime_error("Cast error: formats not matching");
    shape_type sh(shp, shp + nd);
    xdevice_array<T> a(sh);
    if (fortran) {
        shape_type st(sh.size(), 1);
        for (size_t k = 1; k < sh.size(); ++k) st[k] = st[k - 1] * std::max<int64_t>(sh[k - 1], 1);
        a = a.reshaped_view(sh, st, 0);
    }
    xg_tensor t = a.tensor();
    check(xg_npy_load(path.c_str(), &t, current_stream()));
    return a;
}

// xt::dump_npy(path, e) (io/xnpy.hpp:709-755): writes a C-order file with
// the reference's header layout; an expression is evaluated first (as T).
template <class T> void dump_npy(const std::string& path, const xdevice_array<T>& a) {
    xg_tensor t = a.tensor();
    check(xg_npy_dump(path.c_str(), &t, current_stream()));
}
template <class T, class E, std::enable_if_t<is_dev_expr<E>::value, int> = 0>
void dump_npy(const std::string& path, const E& e) {
    dump_npy(path, eval<T>(e));
}

}  // namespace xtensor_amd
