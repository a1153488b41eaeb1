// acc.hpp — accumulators of the reduction and scan trees on gfx950.
//
// AK_F32C / AK_F64C: running sum s plus compensation c; every addition is
// an error-free TwoSum (Knuth), the rounding error going into c
// (Neumaier / Kahan-Babuska summation), and two partial sums combine the
// same way, so the pairwise tree keeps the compensation.  The other kinds
// are plain float products / maxima / minima and exact two's-complement
// integer arithmetic.
#pragma once

#include "interp.hpp"
#include "reduce.hpp"

namespace xg {

// ---------------------------------------------------------------- accumulators
template <int AK, int OP> struct Acc;

template <int OP> struct Acc<AK_F32C, OP> {
    float s, c;
    __device__ void init(uint64_t id) { s = as_f32(id); c = 0.f; }
    __device__ void add(Reg x) {
        float v = as_f32(x);
        float t = s + v;
        float bp = t - s;
        c += (s - (t - bp)) + (v - bp);
        s = t;
    }
    __device__ void merge(const Acc& o) {
        float t = s + o.s;
        float bp = t - s;
        float err = (s - (t - bp)) + (o.s - bp);
        s = t;
        c = c + o.c + err;
    }
    __device__ Acc shfl_xor(int m) const { Acc r; r.s = __shfl_xor(s, m); r.c = __shfl_xor(c, m); return r; }
    __device__ Acc shfl_up(int d) const { Acc r; r.s = __shfl_up(s, d); r.c = __shfl_up(c, d); return r; }
    __device__ Acc shfl_idx(int src) const { Acc r; r.s = __shfl(s, src); r.c = __shfl(c, src); return r; }
    __device__ Store store() const { return Store{__float_as_uint(s), __float_as_uint(c)}; }
    __device__ void load(const Store& st) { s = __uint_as_float(uint32_t(st.a)); c = __uint_as_float(uint32_t(st.b)); }
    // the accumulator as it sits in LDS (8 bytes, not a 16-byte Store)
    using Lds = float2;
    __device__ Lds lds() const { return float2{s, c}; }
    __device__ void from_lds(const Lds& v) { s = v.x; c = v.y; }
    __device__ Reg value(int vdt) const {
        if (vdt == XG_FLOAT32) return from_f32(s + c);
        return from_f64(double(s) + double(c));
    }
};

template <int OP> struct Acc<AK_F64C, OP> {
    double s, c;
    __device__ void init(uint64_t id) { s = as_f64(id); c = 0.0; }
    __device__ void add(Reg x) {
        double v = as_f64(x);
        double t = s + v;
        double bp = t - s;
        c += (s - (t - bp)) + (v - bp);
        s = t;
    }
    __device__ void merge(const Acc& o) {
        double t = s + o.s;
        double bp = t - s;
        double err = (s - (t - bp)) + (o.s - bp);
        s = t;
        c = c + o.c + err;
    }
    __device__ Acc shfl_xor(int m) const { Acc r; r.s = __shfl_xor(s, m); r.c = __shfl_xor(c, m); return r; }
    __device__ Acc shfl_up(int d) const { Acc r; r.s = __shfl_up(s, d); r.c = __shfl_up(c, d); return r; }
    __device__ Acc shfl_idx(int src) const { Acc r; r.s = __shfl(s, src); r.c = __shfl(c, src); return r; }
    __device__ Store store() const { return Store{uint64_t(__double_as_longlong(s)), uint64_t(__double_as_longlong(c))}; }
    __device__ void load(const Store& st) { s = __longlong_as_double(int64_t(st.a)); c = __longlong_as_double(int64_t(st.b)); }
    using Lds = double2;
    __device__ Lds lds() const { return double2{s, c}; }
    __device__ void from_lds(const Lds& v) { s = v.x; c = v.y; }
    __device__ Reg value(int) const { return from_f64(s + c); }
};

template <int AK> struct AccType;
template <> struct AccType<AK_F32> { using T = float; static constexpr int dt = XG_FLOAT32; };
template <> struct AccType<AK_F64> { using T = double; static constexpr int dt = XG_FLOAT64; };
template <> struct AccType<AK_I32> { using T = int32_t; static constexpr int dt = XG_INT32; };
template <> struct AccType<AK_U32> { using T = uint32_t; static constexpr int dt = XG_UINT32; };
template <> struct AccType<AK_I64> { using T = int64_t; static constexpr int dt = XG_INT64; };
template <> struct AccType<AK_U64> { using T = uint64_t; static constexpr int dt = XG_UINT64; };

template <class T> __device__ __forceinline__ T reg_get(Reg r) {
    if constexpr (std::is_same_v<T, float>) return as_f32(r);
    else if constexpr (std::is_same_v<T, double>) return as_f64(r);
    else if constexpr (sizeof(T) == 4) return T(uint32_t(r));
    else return T(r);
}
template <class T> __device__ __forceinline__ Reg reg_put(T x) {
    if constexpr (std::is_same_v<T, float>) return from_f32(x);
    else if constexpr (std::is_same_v<T, double>) return from_f64(x);
    else if constexpr (sizeof(T) == 4) return Reg(uint32_t(x));
    else return Reg(uint64_t(x));
}

// plain accumulators; integer + and * wrap (computed unsigned)
template <int AK, int OP> struct Acc {
    using T = typename AccType<AK>::T;
    T v;
    __device__ static T combine(T a, T b) {
        if constexpr (OP == RO_ADD) {
            if constexpr (std::is_integral_v<T>) return T(std::make_unsigned_t<T>(a) + std::make_unsigned_t<T>(b));
            else return a + b;
        } else if constexpr (OP == RO_MUL) {
            if constexpr (std::is_integral_v<T>) return T(std::make_unsigned_t<T>(a) * std::make_unsigned_t<T>(b));
            else return a * b;
        } else if constexpr (OP == RO_MAX) {
            return a > b ? a : b;  // math::maximum: select(a > b, a, b)
        } else if constexpr (OP == RO_MIN) {
            return a < b ? a : b;  // math::minimum: select(a < b, a, b)
        } else if constexpr (OP == RO_MAXNUM) {
            return a < b ? b : a;  // std::max(a, b)
        } else {
            return b < a ? b : a;  // std::min(a, b)
        }
    }
    __device__ void init(uint64_t id) { v = reg_get<T>(id); }
    __device__ void add(Reg x) { v = combine(v, reg_get<T>(x)); }
    __device__ void merge(const Acc& o) { v = combine(v, o.v); }
    __device__ Acc shfl_xor(int m) const { Acc r; r.v = __shfl_xor(v, m); return r; }
    __device__ Acc shfl_up(int d) const { Acc r; r.v = __shfl_up(v, d); return r; }
    __device__ Acc shfl_idx(int src) const { Acc r; r.v = __shfl(v, src); return r; }
    __device__ Store store() const { return Store{uint64_t(reg_put<T>(v)), 0}; }
    __device__ void load(const Store& st) { v = reg_get<T>(st.a); }
    using Lds = T;
    __device__ Lds lds() const { return v; }
    __device__ void from_lds(const Lds& x) { v = x; }
    __device__ Reg value(int vdt) const { return cvt(reg_put<T>(v), AccType<AK>::dt, vdt); }
};

}  // namespace xg
