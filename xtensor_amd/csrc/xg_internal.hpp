// xg_internal.hpp — definitions shared by the host planner and the gfx950
// kernels: dtype helpers, the typed micro-op program the planner lowers an
// expression tree into, and the launch plans handed to the kernels by value
// (kernel arguments live in SGPR-loadable kernarg memory, so every field here
// is wave-uniform on the device).
#pragma once

#include <stdint.h>

#include "../../include/xg.h"

#if defined(__HIPCC__)
#define XG_HD __host__ __device__
#else
#define XG_HD
#endif

namespace xg {

// ---------------------------------------------------------------- dtypes
XG_HD constexpr int dtype_size(int dt) {
    return (dt == XG_BOOL || dt == XG_INT8 || dt == XG_UINT8) ? 1
         : (dt == XG_INT16 || dt == XG_UINT16 || dt == XG_BFLOAT16) ? 2
         : (dt == XG_INT32 || dt == XG_UINT32 || dt == XG_FLOAT32) ? 4
         : 8;
}
XG_HD constexpr bool is_float(int dt) { return dt == XG_FLOAT32 || dt == XG_FLOAT64 || dt == XG_BFLOAT16; }
XG_HD constexpr bool is_signed_int(int dt) {
    return dt == XG_INT8 || dt == XG_INT16 || dt == XG_INT32 || dt == XG_INT64;
}
XG_HD constexpr bool is_unsigned_int(int dt) {
    return dt == XG_BOOL || dt == XG_UINT8 || dt == XG_UINT16 || dt == XG_UINT32 || dt == XG_UINT64;
}
XG_HD constexpr bool is_64bit(int dt) { return dt == XG_INT64 || dt == XG_UINT64 || dt == XG_FLOAT64; }
// type a leaf of storage type `dt` has as a value (bf16 is storage only)
XG_HD constexpr int value_dtype(int dt) { return dt == XG_BFLOAT16 ? XG_FLOAT32 : dt; }

// A partial result of a reduction or scan tree as it sits in memory: a
// compensated pair (sum, compensation), or a plain value in `a`.
struct Store {
    uint64_t a, b;
};

// ---------------------------------------------------------------- micro-ops
// Registers are 64-bit.  A value of a type of at most 32 bits lives in the
// low 32 bits (small integer types sign- or zero-extended to 32 bits as a
// C++ promotion would), a 64-bit value uses all of them.
enum MicroOpCode : uint8_t {
    MO_LOAD = 0,    // dst <- leaf[aux] converted from its storage type (t)
    MO_CONST,       // dst <- constant[aux]
    MO_CVT,         // dst <- static_cast<aux>(a of type t)
    MO_NEG,         // arithmetic, t in {I32,U32,I64,U64,F32,F64}
    MO_BNOT,        // ~a, integer t
    MO_LNOT,        // !a, a is bool
    MO_ADD, MO_SUB, MO_MUL, MO_DIV, MO_MOD,
    MO_BAND, MO_BOR, MO_BXOR, MO_SHL, MO_SHR,  // shift: t = left type, b = count
    MO_LT, MO_LE, MO_GT, MO_GE, MO_EQ, MO_NE,  // t = operand type, result bool
    MO_LAND, MO_LOR,                           // bool x bool -> bool
    MO_SEL,         // dst <- a ? b : c  (a bool; b, c same type)
    MO_POW, MO_FMOD, MO_ATAN2, MO_HYPOT,       // t in {F32, F64}
    MO_ABS,         // t in {I32, I64, F32, F64}
    MO_EXP, MO_EXP2, MO_EXPM1, MO_LOG, MO_LOG2, MO_LOG10, MO_LOG1P, MO_SQRT, MO_CBRT,
    MO_SIN, MO_COS, MO_TAN, MO_TANH, MO_SINH, MO_COSH, MO_ASIN, MO_ACOS, MO_ATAN, MO_ERF,
    MO_FLOOR, MO_CEIL, MO_TRUNC, MO_ROUND,
    MO_GATHER,      // dst <- leaf[aux] at element offset a (int64), storage type t; 0 outside its extent
    MO_NUM_CODES
};

struct MicroOp {
    uint8_t op;
    uint8_t t;    // operand dtype
    uint8_t dst;
    uint8_t a;
    uint8_t b;
    uint8_t c;
    uint16_t aux; // leaf / constant number, or target dtype of MO_CVT
};

constexpr int kMaxOps = 64;
constexpr int kMaxRegs = 12;
constexpr int kMaxLeaves = XG_MAX_LEAVES;
constexpr int kMaxConsts = XG_MAX_CONSTS;
constexpr int kMaxRank = XG_MAX_RANK;

// How a leaf is read along the innermost iteration dimension.
enum LeafMode : uint8_t {
    LM_CONTIG = 0,  // inner stride 1 and vector-aligned: wide loads
    LM_BCAST = 1,   // inner stride 0: one load per chunk
    LM_STRIDED = 2, // anything else: one load per element
    LM_GATHER = 3,  // read at computed offsets (MO_GATHER); stride[0..1] = valid offset range
    // flag on CONTIG / STRIDED: a large leaf whose every element is read by
    // exactly one work item (no broadcast reuse) is loaded non-temporally,
    // so the stream does not evict cached lines (dirty ones included) that
    // it will never hit
    LM_NT = 8,
    // flag on a CONTIG output whose rows end (or start) inside a 64-byte
    // memory granule that the launch writes only in part (a strided view of
    // rows apart): before storing a row's last (first) unit, the lane reads
    // one word of the rest of that granule (see fill_granules)
    LM_FILL = 16
};
XG_HD constexpr int lm_base(int mode) { return mode & 7; }
XG_HD constexpr bool lm_nt(int mode) { return (mode & LM_NT) != 0; }
XG_HD constexpr bool lm_fill(int mode) { return (mode & LM_FILL) != 0; }
// leaves spanning at least this many bytes stream non-temporally
constexpr int64_t kNtMinBytes = int64_t(64) << 20;

// Division by a run-time constant with a precomputed multiplier
// (Granlund & Montgomery, "Division by invariant integers using
// multiplication", PLDI 1994, round-up variant for 32-bit unsigned n < 2^31).
struct FastDiv {
    uint32_t d;
    uint32_t mul;
    uint32_t shift;
    uint32_t pad_;
};

XG_HD inline FastDiv make_fastdiv(uint32_t d) {
    FastDiv f{d, 0, 0, 0};
    if (d <= 1) { f.mul = 0; f.shift = 0; return f; }
    uint32_t l = 0;
    while ((1ull << l) < d) ++l;  // l = ceil(log2 d)
    uint64_t m = ((1ull << 32) * ((1ull << l) - d)) / d + 1;
    f.mul = uint32_t(m);
    f.shift = l;
    return f;
}

#if defined(__HIPCC__)
__device__ __forceinline__ uint32_t fastdiv(uint32_t n, const FastDiv& f) {
    if (f.d == 1) return n;
    uint32_t t = __umulhi(n, f.mul);
    return (t + ((n - t) >> 1)) >> (f.shift - 1);
}
#endif

XG_HD inline uint32_t fastdiv_host(uint32_t n, const FastDiv& f) {
    if (f.d == 1) return n;
    uint32_t t = uint32_t((uint64_t(n) * f.mul) >> 32);
    return (t + ((n - t) >> 1)) >> (f.shift - 1);
}

// ---------------------------------------------------------------- programs
struct Program {
    int32_t nops;
    int32_t nregs;
    int32_t nleaves;
    int32_t nconsts;
    int32_t result_reg;
    int32_t result_dtype;   // value type of the result register
    // the program numbers leaves and constants in order of first use; these
    // map that numbering back to the expression's leaf / constant numbers
    int8_t leaf_src[kMaxLeaves];
    uint8_t leaf_gather[kMaxLeaves];  // 1: the leaf is read by MO_GATHER only
    int8_t const_src[kMaxConsts];
    MicroOp ops[kMaxOps];
    uint64_t consts[kMaxConsts];
};

struct LeafView {
    const char* ptr;         // base pointer + offset applied
    int64_t stride[kMaxRank];// element strides against the iteration dims
    int32_t dtype;           // storage dtype
    int32_t mode;            // LeafMode along the innermost dim
};

// ---------------------------------------------------------------- elementwise
// Iteration space: rank dims (outer..inner) after dropping size-1 dims and
// merging dims contiguous in every operand.  Work unit = V consecutive
// elements of one inner row.
struct EwPlan {
    int32_t rank;
    int32_t out_dtype;
    int32_t out_mode;        // LM_CONTIG or LM_STRIDED, | LM_NT, | LM_FILL
    int32_t index64;         // 1 if any offset needs 64-bit arithmetic
    uint32_t fill_zero;      // always 0: ties a granule-fill read into the stored value
    int64_t shape[kMaxRank];
    FastDiv div[kMaxRank];   // div[k] divides by shape[k] (outer dims)
    uint32_t inner;          // shape[rank-1]
    uint32_t nchunks;        // ceil(inner / V)
    FastDiv div_chunks;
    uint64_t units;          // rows * nchunks
    char* out_ptr;
    int64_t out_stride[kMaxRank];
    LeafView leaves[kMaxLeaves];
    Program prog;
    // compact form for compiled kernels (rank <= 4, 32-bit offsets):
    // cst[l] = strides of leaf l, cst[kMaxLeaves] = output strides
    int32_t compact;
    uint32_t shape32[4];
    int32_t cst[kMaxLeaves + 1][4];
};

}  // namespace xg
