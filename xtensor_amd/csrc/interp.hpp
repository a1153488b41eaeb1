// interp.hpp — device side of the micro-op program: value conversions,
// operator semantics, leaf loads and result stores on gfx950.
//
// One source serves two kinds of kernels.  A dynamic kernel reads each
// MicroOp from its kernel arguments (wave-uniform, scalar registers) and
// switches on it.  A specialised kernel gets the same MicroOps as constexpr
// values; `exec_op` is force-inlined, so the compiler folds every switch and
// emits straight-line code for that program.
//
// Semantics follow C++ on the value types the planner assigned (xtensor's
// functors are C++ operators, core/xoperation.hpp:86-109): integer + - * and
// shifts wrap in two's complement, float ops round once each (the library
// is built with -ffp-contract=off), float / and sqrt are IEEE correctly
// rounded (HIP's default), transcendental functions are the ROCm device
// library's (<= 2 ulp).
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>

#include "xg_internal.hpp"

namespace xg {

using Reg = uint64_t;

__device__ __forceinline__ float as_f32(Reg r) { return __uint_as_float(uint32_t(r)); }
__device__ __forceinline__ double as_f64(Reg r) { return __longlong_as_double(int64_t(r)); }
__device__ __forceinline__ Reg from_f32(float f) { return Reg(__float_as_uint(f)); }
__device__ __forceinline__ Reg from_f64(double d) { return Reg(__double_as_longlong(d)); }

// static_cast<to>(value of type from)
__device__ __forceinline__ Reg cvt(Reg r, int from, int to) {
    if (from == XG_FLOAT32 || from == XG_FLOAT64) {
        if (from == XG_FLOAT32) {
            float f = as_f32(r);
            switch (to) {
            case XG_FLOAT32: return r;
            case XG_FLOAT64: return from_f64(double(f));
            case XG_BOOL: return f != 0.0f;
            case XG_UINT64: return Reg(uint64_t(f));
            case XG_INT64: return Reg(int64_t(f));
            case XG_UINT32: return Reg(uint32_t(f));
            case XG_INT32: return Reg(uint32_t(int32_t(f)));
            case XG_INT8: return Reg(uint32_t(int32_t(int8_t(int32_t(f)))));
            case XG_UINT8: return Reg(uint8_t(int32_t(f)));
            case XG_INT16: return Reg(uint32_t(int32_t(int16_t(int32_t(f)))));
            case XG_UINT16: return Reg(uint16_t(int32_t(f)));
            default: return 0;
            }
        } else {
            double d = as_f64(r);
            switch (to) {
            case XG_FLOAT32: return from_f32(float(d));
            case XG_FLOAT64: return r;
            case XG_BOOL: return d != 0.0;
            case XG_UINT64: return Reg(uint64_t(d));
            case XG_INT64: return Reg(int64_t(d));
            case XG_UINT32: return Reg(uint32_t(d));
            case XG_INT32: return Reg(uint32_t(int32_t(d)));
            case XG_INT8: return Reg(uint32_t(int32_t(int8_t(int32_t(d)))));
            case XG_UINT8: return Reg(uint8_t(int32_t(d)));
            case XG_INT16: return Reg(uint32_t(int32_t(int16_t(int32_t(d)))));
            case XG_UINT16: return Reg(uint16_t(int32_t(d)));
            default: return 0;
            }
        }
    }
    // integer (or bool) source: its exact value as a 64-bit integer
    int64_t v;
    if (from == XG_INT64 || from == XG_UINT64) v = int64_t(r);
    else if (from == XG_UINT32) v = int64_t(uint32_t(r));
    else v = int64_t(int32_t(uint32_t(r)));  // i32 and the sign/zero-extended small types
    switch (to) {
    case XG_BOOL: return v != 0;
    case XG_INT8: return Reg(uint32_t(int32_t(int8_t(v))));
    case XG_UINT8: return Reg(uint8_t(v));
    case XG_INT16: return Reg(uint32_t(int32_t(int16_t(v))));
    case XG_UINT16: return Reg(uint16_t(v));
    case XG_INT32: case XG_UINT32: return Reg(uint32_t(v));
    case XG_INT64: case XG_UINT64: return Reg(v);
    case XG_FLOAT32:
        if (from == XG_UINT64) return from_f32(float(uint64_t(v)));
        if (from == XG_INT64) return from_f32(float(v));
        if (from == XG_UINT32) return from_f32(float(uint32_t(v)));
        return from_f32(float(int32_t(v)));
    case XG_FLOAT64:
        if (from == XG_UINT64) return from_f64(double(uint64_t(v)));
        if (from == XG_INT64) return from_f64(double(v));
        if (from == XG_UINT32) return from_f64(double(uint32_t(v)));
        return from_f64(double(int32_t(v)));
    default: return 0;
    }
}

template <class T> __device__ __forceinline__ T iget(Reg r) {
    if constexpr (sizeof(T) == 4) return T(uint32_t(r));
    else return T(r);
}

template <class T> __device__ __forceinline__ Reg iput(T x) {
    if constexpr (sizeof(T) == 4) return Reg(uint32_t(x));
    else return Reg(uint64_t(x));
}

// integer binary op on type T (S = same-width signed, U = unsigned)
template <class T> __device__ __forceinline__ Reg int_binop(int op, Reg x, Reg y) {
    using U = std::conditional_t<sizeof(T) == 4, uint32_t, uint64_t>;
    T a = iget<T>(x), b = iget<T>(y);
    U ua = U(a), ub = U(b);
    constexpr int W = sizeof(T) * 8;
    switch (op) {
    case MO_ADD: return iput<U>(ua + ub);
    case MO_SUB: return iput<U>(ua - ub);
    case MO_MUL: return iput<U>(ua * ub);
    case MO_DIV: return b == 0 ? 0 : iput<T>(a / b);   // C++: undefined; the kernel stays defined
    case MO_MOD: return b == 0 ? 0 : iput<T>(a % b);
    case MO_BAND: return iput<U>(ua & ub);
    case MO_BOR: return iput<U>(ua | ub);
    case MO_BXOR: return iput<U>(ua ^ ub);
    case MO_SHL: return iput<U>(ua << (uint32_t(y) & (W - 1)));
    case MO_SHR: return iput<T>(a >> (uint32_t(y) & (W - 1)));
    case MO_LT: return a < b;
    case MO_LE: return a <= b;
    case MO_GT: return a > b;
    case MO_GE: return a >= b;
    case MO_EQ: return a == b;
    case MO_NE: return a != b;
    default: return 0;
    }
}

template <class T> __device__ __forceinline__ T fget(Reg r) {
    if constexpr (sizeof(T) == 4) return as_f32(r);
    else return as_f64(r);
}
template <class T> __device__ __forceinline__ Reg fput(T x) {
    if constexpr (sizeof(T) == 4) return from_f32(x);
    else return from_f64(x);
}

template <class T> __device__ __forceinline__ Reg flt_binop(int op, Reg x, Reg y) {
    T a = fget<T>(x), b = fget<T>(y);
    switch (op) {
    case MO_ADD: return fput<T>(a + b);
    case MO_SUB: return fput<T>(a - b);
    case MO_MUL: return fput<T>(a * b);
    case MO_DIV: return fput<T>(a / b);
    case MO_LT: return a < b;
    case MO_LE: return a <= b;
    case MO_GT: return a > b;
    case MO_GE: return a >= b;
    case MO_EQ: return a == b;
    case MO_NE: return a != b;
    case MO_POW: return fput<T>(pow(a, b));
    case MO_FMOD: return fput<T>(fmod(a, b));
    case MO_ATAN2: return fput<T>(atan2(a, b));
    case MO_HYPOT: return fput<T>(hypot(a, b));
    default: return 0;
    }
}

// cube root in double: the device library's value refined by one Newton step
// whose residual y^3 - x is evaluated with an exact product (y*y split by
// FMA), which brings it within 1 ulp of the exact cube root
__device__ __forceinline__ double cbrt_refined(double x) {
    double y = cbrt(x);
    if (!(y != 0.0) || !isfinite(y)) return y;
    double y2 = y * y;
    double y2e = fma(y, y, -y2);          // y*y = y2 + y2e exactly
    double r = fma(y2, y, -x) + y2e * y;  // y^3 - x
    return y - r / (3.0 * y2);
}

template <class T> __device__ __forceinline__ Reg flt_unop(int op, Reg x) {
    T a = fget<T>(x);
    switch (op) {
    case MO_NEG: return fput<T>(-a);
    case MO_ABS: return fput<T>(fabs(a));
    case MO_EXP: return fput<T>(exp(a));
    case MO_EXP2: return fput<T>(exp2(a));
    case MO_EXPM1: return fput<T>(expm1(a));
    case MO_LOG: return fput<T>(log(a));
    case MO_LOG2: return fput<T>(log2(a));
    case MO_LOG10: return fput<T>(log10(a));
    case MO_LOG1P: return fput<T>(log1p(a));
    case MO_SQRT: return fput<T>(sqrt(a));
    case MO_CBRT:
        if constexpr (sizeof(T) == 8) return fput<T>(cbrt_refined(a));
        else return fput<T>(cbrt(a));
    case MO_SIN: return fput<T>(sin(a));
    case MO_COS: return fput<T>(cos(a));
    case MO_TAN: return fput<T>(tan(a));
    case MO_TANH: return fput<T>(tanh(a));
    case MO_SINH: return fput<T>(sinh(a));
    case MO_COSH: return fput<T>(cosh(a));
    case MO_ASIN: return fput<T>(asin(a));
    case MO_ACOS: return fput<T>(acos(a));
    case MO_ATAN: return fput<T>(atan(a));
    case MO_ERF: return fput<T>(erf(a));
    case MO_FLOOR: return fput<T>(floor(a));
    case MO_CEIL: return fput<T>(ceil(a));
    case MO_TRUNC: return fput<T>(trunc(a));
    case MO_ROUND: return fput<T>(round(a));
    default: return 0;
    }
}

template <class T> __device__ __forceinline__ Reg int_unop(int op, Reg x) {
    using U = std::conditional_t<sizeof(T) == 4, uint32_t, uint64_t>;
    T a = iget<T>(x);
    switch (op) {
    case MO_NEG: return iput<U>(U(0) - U(a));
    case MO_BNOT: return iput<U>(~U(a));
    case MO_ABS: return iput<U>(a < 0 ? U(0) - U(a) : U(a));
    default: return 0;
    }
}

// one scalar micro-op (everything except MO_LOAD)
__device__ __forceinline__ Reg exec_scalar(const MicroOp& m, Reg a, Reg b, Reg c, const uint64_t* kconst) {
    switch (m.op) {
    case MO_CONST: return kconst[m.aux];
    case MO_CVT: return cvt(a, m.t, m.aux);
    case MO_LNOT: return uint32_t(a) == 0;
    case MO_LAND: return (uint32_t(a) != 0) && (uint32_t(b) != 0);
    case MO_LOR: return (uint32_t(a) != 0) || (uint32_t(b) != 0);
    case MO_SEL: return uint32_t(a) != 0 ? b : c;
    default: break;
    }
    bool unary = m.op == MO_NEG || m.op == MO_BNOT || m.op >= MO_ABS;
    switch (m.t) {
    case XG_INT32: return unary ? int_unop<int32_t>(m.op, a) : int_binop<int32_t>(m.op, a, b);
    case XG_UINT32: return unary ? int_unop<uint32_t>(m.op, a) : int_binop<uint32_t>(m.op, a, b);
    case XG_INT64: return unary ? int_unop<int64_t>(m.op, a) : int_binop<int64_t>(m.op, a, b);
    case XG_UINT64: return unary ? int_unop<uint64_t>(m.op, a) : int_binop<uint64_t>(m.op, a, b);
    case XG_FLOAT32: return unary ? flt_unop<float>(m.op, a) : flt_binop<float>(m.op, a, b);
    case XG_FLOAT64: return unary ? flt_unop<double>(m.op, a) : flt_binop<double>(m.op, a, b);
    default: return 0;
    }
}

// ---------------------------------------------------------------- loads
// register representation of one stored element (raw little-endian bits)
__device__ __forceinline__ Reg from_storage(uint64_t raw, int dt) {
    switch (dt) {
    case XG_BOOL: return Reg(uint8_t(raw) != 0);
    case XG_INT8: return Reg(uint32_t(int32_t(int8_t(uint8_t(raw)))));
    case XG_UINT8: return Reg(uint8_t(raw));
    case XG_INT16: return Reg(uint32_t(int32_t(int16_t(uint16_t(raw)))));
    case XG_UINT16: return Reg(uint16_t(raw));
    case XG_BFLOAT16: return Reg(uint32_t(uint16_t(raw)) << 16);
    case XG_INT32: case XG_UINT32: case XG_FLOAT32: return Reg(uint32_t(raw));
    default: return Reg(raw);
    }
}

typedef uint32_t u32x2_t __attribute__((ext_vector_type(2)));
typedef uint32_t u32x4_t __attribute__((ext_vector_type(4)));

// one load of type T; NT = non-temporal (streaming) cache policy
template <bool NT, class T> __device__ __forceinline__ T ld(const T* p) {
    if constexpr (NT) return __builtin_nontemporal_load(p);
    else return *p;
}

template <bool NT = false>
__device__ __forceinline__ uint64_t load_raw(const char* p, int esz) {
    switch (esz) {
    case 1: return ld<NT>(reinterpret_cast<const uint8_t*>(p));
    case 2: return ld<NT>(reinterpret_cast<const uint16_t*>(p));
    case 4: return ld<NT>(reinterpret_cast<const uint32_t*>(p));
    default: return ld<NT>(reinterpret_cast<const uint64_t*>(p));
    }
}

// V consecutive, naturally aligned elements of size esz in as few wide loads as possible
template <int V, bool NT = false>
__device__ __forceinline__ void load_vec(const char* p, int esz, int dt, Reg (&out)[V]) {
    const int bytes = V * esz;
    if constexpr (V == 1) {
        out[0] = from_storage(load_raw<NT>(p, esz), dt);
    } else if (bytes >= 16) {
        constexpr int MAXW = V * 2;  // up to 8-byte elements
        uint32_t w[MAXW];
#pragma unroll
        for (int q = 0; q < MAXW / 4; ++q) {
            if (q * 16 < bytes) {
                u32x4_t x = ld<NT>(reinterpret_cast<const u32x4_t*>(p) + q);
                w[4 * q] = x.x; w[4 * q + 1] = x.y; w[4 * q + 2] = x.z; w[4 * q + 3] = x.w;
            }
        }
#pragma unroll
        for (int v = 0; v < V; ++v) {
            uint64_t raw;
            if (esz == 8) raw = uint64_t(w[2 * v]) | (uint64_t(w[2 * v + 1]) << 32);
            else if (esz == 4) raw = w[v];
            else if (esz == 2) raw = (w[v >> 1] >> (16 * (v & 1))) & 0xffff;
            else raw = (w[v >> 2] >> (8 * (v & 3))) & 0xff;
            out[v] = from_storage(raw, dt);
        }
    } else if (bytes == 8) {
        u32x2_t x = ld<NT>(reinterpret_cast<const u32x2_t*>(p));
        uint64_t all = uint64_t(x.x) | (uint64_t(x.y) << 32);
#pragma unroll
        for (int v = 0; v < V; ++v) out[v] = from_storage(all >> (v * esz * 8), dt);
    } else if (bytes == 4) {
        uint32_t all = ld<NT>(reinterpret_cast<const uint32_t*>(p));
#pragma unroll
        for (int v = 0; v < V; ++v) out[v] = from_storage(all >> (v * esz * 8), dt);
    } else {
#pragma unroll
        for (int v = 0; v < V; ++v) out[v] = from_storage(load_raw<NT>(p + v * esz, esz), dt);
    }
}

// MO_GATHER: element offsets `off` (int64) into a tensor at `base`; offsets
// outside [lo, hi] (and lanes past nvalid) read as 0, so a bad index never
// touches memory outside the tensor
template <int V>
__device__ __forceinline__ void gather_vals(const char* base, int dt, int64_t lo, int64_t hi, int nvalid,
                                            const Reg (&off)[V], Reg (&out)[V]) {
    const int esz = dtype_size(dt);
    Reg tmp[V];
#pragma unroll
    for (int v = 0; v < V; ++v) {
        const int64_t o = int64_t(off[v]);
        tmp[v] = (v < nvalid && o >= lo && o <= hi) ? from_storage(load_raw(base + o * esz, esz), dt) : Reg(0);
    }
#pragma unroll
    for (int v = 0; v < V; ++v) out[v] = tmp[v];
}

// ---------------------------------------------------------------- stores
__device__ __forceinline__ uint64_t to_storage(Reg r, int dt) {
    if (dt == XG_BFLOAT16) {  // float -> bf16, round to nearest even, NaN kept quiet
        uint32_t u = uint32_t(r);
        if ((u & 0x7fffffffu) > 0x7f800000u) return (u >> 16) | 0x40;
        return (u + 0x7fffu + ((u >> 16) & 1u)) >> 16;
    }
    return r;
}

// one store of type T; NT = non-temporal (streaming) cache policy
template <bool NT, class T> __device__ __forceinline__ void st(T* p, T v) {
    if constexpr (NT) __builtin_nontemporal_store(v, p);
    else *p = v;
}

// Placed around the non-temporal store of a branch whose other side stores
// the same value plainly: an empty asm statement the optimiser cannot move
// or merge, so it does not hoist or sink the two stores into one and drop
// the non-temporal hint.  No instruction is emitted.
__device__ __forceinline__ void nt_store_fence() { __asm__ volatile(""); }

template <bool NT = false>
__device__ __forceinline__ void store_raw(char* p, int esz, uint64_t raw) {
    switch (esz) {
    case 1: st<NT>(reinterpret_cast<uint8_t*>(p), uint8_t(raw)); break;
    case 2: st<NT>(reinterpret_cast<uint16_t*>(p), uint16_t(raw)); break;
    case 4: st<NT>(reinterpret_cast<uint32_t*>(p), uint32_t(raw)); break;
    default: st<NT>(reinterpret_cast<uint64_t*>(p), raw); break;
    }
}

template <int V, bool NT = false>
__device__ __forceinline__ void store_vec(char* p, int esz, int dt, const Reg (&in)[V]) {
    const int bytes = V * esz;
    // one element per lane: a plain store (the packing below masks each
    // element to esz bytes, which needs esz < bytes)
    if constexpr (V == 1) {
        store_raw<NT>(p, esz, to_storage(in[0], dt));
    } else if (bytes >= 16) {
        constexpr int MAXW = V * 2;
        uint32_t w[MAXW];
#pragma unroll
        for (int q = 0; q < MAXW; ++q) w[q] = 0;
#pragma unroll
        for (int v = 0; v < V; ++v) {
            uint64_t raw = to_storage(in[v], dt);
            if (esz == 8) { w[2 * v] = uint32_t(raw); w[2 * v + 1] = uint32_t(raw >> 32); }
            else if (esz == 4) w[v] = uint32_t(raw);
            else if (esz == 2) w[v >> 1] |= (uint32_t(raw) & 0xffff) << (16 * (v & 1));
            else w[v >> 2] |= (uint32_t(raw) & 0xff) << (8 * (v & 3));
        }
#pragma unroll
        for (int q = 0; q < MAXW / 4; ++q)
            if (q * 16 < bytes) {
                u32x4_t x = {w[4 * q], w[4 * q + 1], w[4 * q + 2], w[4 * q + 3]};
                st<NT>(reinterpret_cast<u32x4_t*>(p) + q, x);
            }
    } else if (bytes == 8) {
        uint64_t all = 0;
#pragma unroll
        for (int v = 0; v < V; ++v) all |= (to_storage(in[v], dt) & ((1ull << (esz * 8)) - 1)) << (v * esz * 8);
        u32x2_t x = {uint32_t(all), uint32_t(all >> 32)};
        st<NT>(reinterpret_cast<u32x2_t*>(p), x);
    } else if (bytes == 4) {
        uint32_t all = 0;
#pragma unroll
        for (int v = 0; v < V; ++v) all |= (uint32_t(to_storage(in[v], dt)) & ((1u << (esz * 8)) - 1)) << (v * esz * 8);
        st<NT>(reinterpret_cast<uint32_t*>(p), all);
    } else {
#pragma unroll
        for (int v = 0; v < V; ++v) store_raw<NT>(p + v * esz, esz, to_storage(in[v], dt));
    }
}

// store_vec with the non-temporal hint chosen at run time; the two forms are
// kept apart by nt_store_fence, else the optimiser merges them into one plain
// store
template <int V>
__device__ __forceinline__ void store_vec_rt(bool nt, char* p, int esz, int dt, const Reg (&in)[V]) {
    if (nt) {
        nt_store_fence();
        store_vec<V, true>(p, esz, dt, in);
        nt_store_fence();
    } else {
        store_vec<V>(p, esz, dt, in);
    }
}

}  // namespace xg
