// batch.hip — one persistent launch over many uneven instances.
//
// The plain path evaluates a row softmax of one instance as three launches
// (amax with keep_dims, the fused exp-sum reduction, the normalising
// assignment; the pattern of test/test_xmath.cpp:985-1027).  Here a single
// persistent kernel handles a whole batch: every wave works through a list
// of tasks (a few rows of one instance) that the host balanced across the
// grid, then takes left-over tasks from a shared work queue (an atomic
// counter); it loads each row once into registers and writes the result
// once.  A row of `cols` elements is owned by `lanes_per_row(cols)` lanes;
// its maximum is exact, its sum is a compensated (TwoSum) sum folded in a
// fixed butterfly order, so the value of every row depends only on that
// row — never on the batch around it or on the schedule.
//
// Host side: task split and schedule (longest task first onto the least
// loaded wave), a two-node HIP graph (queue reset, kernel) per batch, and an
// autotuner over task size, launch shape and queue share whose choice is
// cached per shape class.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <climits>
#include <cmath>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <queue>
#include <tuple>
#include <vector>

#include "acc.hpp"
#include "batch.hpp"
#include "xg_api_internal.hpp"

namespace xg {

template <class T> struct BTraits;
template <> struct BTraits<float> {
    static constexpr int AK = AK_F32C;
    static constexpr int DT = XG_FLOAT32;
    __device__ static float ex(float x) { return expf(x); }
    __device__ static Reg reg(float x) { return from_f32(x); }
    __device__ static float val(Reg r) { return as_f32(r); }
    __device__ static float lowest() { return -FLT_MAX; }
};
template <> struct BTraits<double> {
    static constexpr int AK = AK_F64C;
    static constexpr int DT = XG_FLOAT64;
    __device__ static double ex(double x) { return exp(x); }
    __device__ static Reg reg(double x) { return from_f64(x); }
    __device__ static double val(Reg r) { return as_f64(r); }
    __device__ static double lowest() { return -DBL_MAX; }
};
template <> struct BTraits<int32_t> {
    static constexpr int AK = AK_I32;
    static constexpr int DT = XG_INT32;
    __device__ static int32_t ex(int32_t x) { return x; }
    __device__ static Reg reg(int32_t x) { return Reg(uint32_t(x)); }
    __device__ static int32_t val(Reg r) { return int32_t(uint32_t(r)); }
    __device__ static int32_t lowest() { return INT32_MIN; }
};

// running maximum as xtensor's `maximum` folds it: acc = acc > x ? acc : x,
// starting from numeric_limits<T>::lowest() (core/xmath.hpp:573-602, 1823)
template <class T> __device__ __forceinline__ T fmax_x(T acc, T x) { return acc > x ? acc : x; }

// One-instruction maximum (IEEE maxNum: a NaN operand yields the other).
// Used for the softmax shift only: a NaN anywhere in a row makes that row's
// sum, and so every output of the row, NaN either way.
template <class T> __device__ __forceinline__ T max_num(T a, T b) {
    if constexpr (std::is_same_v<T, float>) return __builtin_fmaxf(a, b);
    else if constexpr (std::is_same_v<T, double>) return __builtin_fmax(a, b);
    else return fmax_x(a, b);
}

// butterfly combine inside aligned groups of `lpr` lanes, lower lane first
template <class A> __device__ __forceinline__ A sub_combine(A a, const int lpr) {
    const int lane = threadIdx.x & 63;
    for (int m = 1; m < lpr; m <<= 1) {
        A o = a.shfl_xor(m);
        if (lane & m) {
            o.merge(a);
            a = o;
        } else {
            a.merge(o);
        }
    }
    return a;
}

template <class T> __device__ __forceinline__ T sub_max(T m, int lpr) {
    const int lane = threadIdx.x & 63;
    for (int k = 1; k < lpr; k <<= 1) {
        T o = __shfl_xor(m, k);
        m = (lane & k) ? fmax_x(o, m) : fmax_x(m, o);
    }
    return m;
}

// exp(x - m) for x <= m (m is the row maximum), without the rounding error
// of forming x - m in the input precision: d + e == x - m exactly (TwoSum).
//   float:  2^p with p = (d + e) * log2(e) carried as ph + pl (Dekker
//           product of d and log2(e) split in two floats, plus the small
//           terms), p = n + f with n = rint(ph), 2^f by the hardware exp2 on
//           |f| <= 0.5 + tiny, scaled by 2^n.  Float arithmetic only.
//   double: exp(d) * (1 + e).
// x - m is in [-inf, 0] or NaN; -inf (and anything below the float range)
// gives 0, NaN stays NaN.
template <class T> __device__ __forceinline__ T exp_shifted(T x, T m) {
    const T nm = -m;
    const T d = x + nm;
    const T bp = d - x;
    const T e = (x - (d - bp)) + (nm - bp);
    if constexpr (std::is_same_v<T, float>) {
        constexpr float L = 1.44269502f;         // log2(e) rounded to float
        constexpr float Llo = 1.92596303e-08f;    // log2(e) - L
        const float ph = d * L;
        const float pl = __builtin_fmaf(d, L, -ph) + __builtin_fmaf(d, Llo, e * L);
        const float n = __builtin_rintf(ph);
        const float f = (ph - n) + pl;
        // the clamp keeps the exponent conversion defined for d = -inf / NaN
        const float r = __builtin_ldexpf(__builtin_amdgcn_exp2f(f), int(__builtin_fmaxf(n, -200.f)));
        return d < -160.f ? 0.f : r;
    } else {
        const double v = exp(d);
        return (d >= -DBL_MAX) ? __builtin_fma(v, e, v) : v;
    }
}

// Divide by a row's sum: q = RN(v * r) with r = 1 / sum, then one
// residual correction step (Markstein), which gives the correctly rounded
// quotient except in rare double-rounding cases (then 1 ulp off).
template <class T> __device__ __forceinline__ T div_by(T v, T sum, T r) {
    if constexpr (std::is_same_v<T, float>) {
        const float q = v * r;
        const float res = __builtin_fmaf(-q, sum, v);
        return __builtin_fmaf(res, r, q);
    } else {
        const double q = v * r;
        const double res = __builtin_fma(-q, sum, v);
        return __builtin_fma(res, r, q);
    }
}

// Buffer resources (gfx9 raw buffers, stride 0): a load whose byte offset
// lies at or past `bytes` returns 0 and a store there is dropped, so row
// tails and rows past the end of a task need no branches.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t rsrc(const void* p, uint32_t bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), short(0), int(bytes), 0x00020000);
}
template <class T> __device__ __forceinline__ T bload(__amdgpu_buffer_rsrc_t r, uint32_t off) {
    if constexpr (sizeof(T) == 4) return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b32(r, int(off), 0, 0));
    else return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b64(r, int(off), 0, 0));
}
template <class T> __device__ __forceinline__ void bstore(__amdgpu_buffer_rsrc_t r, uint32_t off, T v) {
    if constexpr (sizeof(T) == 4) __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, v), r, int(off), 0, 0);
    else __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(decltype(__builtin_amdgcn_raw_buffer_load_b64(r, 0, 0, 0)), v), r, int(off), 0, 0);
}
constexpr uint32_t kOffNone = 0x80000000u;  // an offset past every resource

// RB row groups at once, NE elements per lane per row (cols <= NE * LPR,
// cols >= 1).  A row group is PER = 64 / LPR consecutive rows, row `sub`
// of it in lanes [sub * LPR, sub * LPR + LPR); its elements are
// x[sl + k * LPR], k < NE.  One buffer resource spans a row group (its
// first row to the end of its last row inside the task), so all RB * NE
// loads go out back to back, unmasked.
template <class T, int OP, int LPR, int NE, int RB>
__device__ __forceinline__ void rows_regs(const BItem& it, int nrows, int row0, int rb, int sub, int sl) {
    using Tr = BTraits<T>;
    constexpr int PER = 64 / LPR;
    constexpr uint32_t E = sizeof(T);
    const int cols = it.cols;
    T v[RB][NE];
    __amdgpu_buffer_rsrc_t rin[RB];
#pragma unroll
    for (int g = 0; g < RB; ++g) {
        const int first = rb + g * PER;                 // wave-uniform
        const int nvalid = min(PER, nrows - first);     // rows of the group inside the task
        const int64_t row = int64_t(row0) + (first < nrows ? first : 0);
        const uint32_t bytes = nvalid > 0 ? uint32_t((int64_t(nvalid - 1) * it.in_rs + cols) * E) : 0u;
        rin[g] = rsrc(it.in + row * it.in_rs * E, bytes);
        const uint32_t loff = (uint32_t(sub) * uint32_t(it.in_rs) + uint32_t(sl)) * E;
#pragma unroll
        for (int k = 0; k < NE; ++k) {
            // element sl + k * LPR.  For NE >= 2 the class guarantees
            // cols > (NE / 2) * LPR, so the first half is always inside the
            // row.  (Skipping chunks wholly past the row end with a uniform
            // branch was measured slower: it splits the batch of loads.)
            // The lane-dependent part stays on the left of the comparison so
            // that no per-k lane value is kept live.
            const bool in = (NE >= 2 && k < NE / 2) || sl < cols - k * LPR;
            const T x = bload<T>(rin[g], loff + uint32_t(k * LPR) * E);
            if constexpr (OP == BO_SUM) v[g][k] = (PER == 1 || in) ? x : T(0);
            else v[g][k] = in ? x : Tr::lowest();
        }
    }
    T m[RB];
    if constexpr (OP != BO_SUM) {
#pragma unroll
        for (int g = 0; g < RB; ++g) {
            m[g] = v[g][0];
#pragma unroll
            for (int k = 1; k < NE; ++k) m[g] = OP == BO_SOFTMAX ? max_num(m[g], v[g][k]) : fmax_x(m[g], v[g][k]);
        }
#pragma unroll
        for (int g = 0; g < RB; ++g) m[g] = sub_max(m[g], LPR);
    }
    if constexpr (OP == BO_MAX) {
#pragma unroll
        for (int g = 0; g < RB; ++g) {
            const int r = rb + g * PER + sub;
            if (sl == 0 && r < nrows) reinterpret_cast<T*>(it.out)[int64_t(row0 + r) * it.out_rs] = m[g];
        }
    } else {
        Acc<Tr::AK, RO_ADD> s[RB];
#pragma unroll
        for (int g = 0; g < RB; ++g) {
            s[g].init(0);
#pragma unroll
            for (int k = 0; k < NE; ++k) {
                if constexpr (OP == BO_SOFTMAX) {
                    const bool in = (NE >= 2 && k < NE / 2) || sl < cols - k * LPR;
                    v[g][k] = in ? exp_shifted(v[g][k], m[g]) : T(0);
                }
                s[g].add(Tr::reg(v[g][k]));
            }
        }
#pragma unroll
        for (int g = 0; g < RB; ++g) s[g] = sub_combine(s[g], LPR);
#pragma unroll
        for (int g = 0; g < RB; ++g) {
            const int r = rb + g * PER + sub;
            const T sum = Tr::val(s[g].value(Tr::DT));
            if constexpr (OP == BO_SUM) {
                if (sl == 0 && r < nrows) reinterpret_cast<T*>(it.out)[int64_t(row0 + r) * it.out_rs] = sum;
            } else {
                const T rcp = T(1) / sum;
                const int first = rb + g * PER;
                const int nvalid = min(PER, nrows - first);
                const int64_t row = int64_t(row0) + (first < nrows ? first : 0);
                const uint32_t bytes = nvalid > 0 ? uint32_t((int64_t(nvalid - 1) * it.out_rs + cols) * E) : 0u;
                const __amdgpu_buffer_rsrc_t ro = rsrc(it.out + row * it.out_rs * E, bytes);
                const uint32_t loff = (uint32_t(sub) * uint32_t(it.out_rs) + uint32_t(sl)) * E;
#pragma unroll
                for (int k = 0; k < NE; ++k) {
                    const uint32_t off = (sub < nvalid && sl < cols - k * LPR) ? loff + uint32_t(k * LPR) * E : kOffNone;
                    bstore<T>(ro, off, div_by(v[g][k], sum, rcp));
                }
            }
        }
    }
}

// rows too long for registers: streaming passes over memory
template <class T, int OP>
__device__ __forceinline__ void row_stream(const BItem& it, int row, bool valid, int sl) {
    using Tr = BTraits<T>;
    const T* p = reinterpret_cast<const T*>(it.in) + int64_t(row) * it.in_rs;
    const int cols = valid ? it.cols : 0;
    T m = Tr::lowest();
    if (OP != BO_SUM) {
        for (int j = sl; j < cols; j += 64) m = fmax_x(m, p[j]);
        m = sub_max(m, 64);
    }
    if (OP == BO_MAX) {
        if (valid && sl == 0) reinterpret_cast<T*>(it.out)[int64_t(row) * it.out_rs] = m;
        return;
    }
    Acc<Tr::AK, RO_ADD> s;
    s.init(0);
    for (int j = sl; j < cols; j += 64) s.add(Tr::reg(OP == BO_SUM ? p[j] : exp_shifted(p[j], m)));
    s = sub_combine(s, 64);
    const T sum = Tr::val(s.value(Tr::DT));
    if (OP == BO_SUM) {
        if (valid && sl == 0) reinterpret_cast<T*>(it.out)[int64_t(row) * it.out_rs] = sum;
        return;
    }
    const T rcp = T(1) / sum;
    T* q = reinterpret_cast<T*>(it.out) + int64_t(row) * it.out_rs;
    for (int j = sl; j < cols; j += 64) q[j] = div_by(exp_shifted(p[j], m), sum, rcp);
}

template <class T, int OP, int LPR, int NE>
__device__ __forceinline__ void rows_loop(const BItem& it, int nrows, int row0, int lane) {
    constexpr int PER = 64 / LPR;
    constexpr int RB = batch_rows_at_once(NE, int(sizeof(T)));
    for (int rb = 0; rb < nrows; rb += PER * RB) {
        // The lane number is made opaque once per iteration: otherwise the
        // compiler hoists every per-element lane offset of every variant to
        // the kernel entry, where they stay live (and spill) for its whole run.
        int l = lane;
        asm volatile("" : "+v"(l));
        rows_regs<T, OP, LPR, NE, RB>(it, nrows, row0, rb, l / LPR, l % LPR);
    }
}

// rows of an empty axis: sum -> 0, amax -> lowest() (the reducers' initial values)
template <class T, int OP>
__device__ __forceinline__ void rows_empty(const BItem& it, int nrows, int row0, int lane) {
    if constexpr (OP != BO_SOFTMAX) {
        const T id = OP == BO_SUM ? T(0) : BTraits<T>::lowest();
        for (int r = lane; r < nrows; r += 64) reinterpret_cast<T*>(it.out)[int64_t(row0 + r) * it.out_rs] = id;
    }
}

template <class T, int OP>
__global__ __launch_bounds__(256, 5) void batch_rows_kernel(const BatchArgs a) {
    const int lane = threadIdx.x & 63;
    const uint32_t w = __builtin_amdgcn_readfirstlane(blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6));
    uint32_t t = a.wbeg[w];
    uint32_t tend = a.wbeg[w + 1];
    for (;; ++t) {
        if (t >= tend) {  // own list done: take from the shared queue
            uint32_t q = 0;
            if (lane == 0) q = atomicAdd(a.counter, 1u);
            t = a.dyn_begin + __builtin_amdgcn_readlane(q, 0);  // wave-uniform from here on
            if (t >= a.ntasks) break;  // every wave leaves once the queue is drained
            tend = t + 1;
        }
        const BTask task = a.tasks[t];
        const BItem it = a.items[task.item];
        const int nrows = task.nrows;
        const int row0 = task.row0;
        const int ne = batch_ne_class(it.cols, it.lpr, int(sizeof(T)));  // the same for the whole task
        if (it.cols == 0) rows_empty<T, OP>(it, nrows, row0, lane);
        else if (it.lpr == 4) rows_loop<T, OP, 4, 1>(it, nrows, row0, lane);
        else if (it.lpr == 8) rows_loop<T, OP, 8, 1>(it, nrows, row0, lane);
        else if (it.lpr == 16) rows_loop<T, OP, 16, 1>(it, nrows, row0, lane);
        else if (it.lpr == 32) rows_loop<T, OP, 32, 1>(it, nrows, row0, lane);
        else if (ne == 1) rows_loop<T, OP, 64, 1>(it, nrows, row0, lane);
        else if (ne == 2) rows_loop<T, OP, 64, 2>(it, nrows, row0, lane);
        else if (ne == 4) rows_loop<T, OP, 64, 4>(it, nrows, row0, lane);
        else if (ne == 8) rows_loop<T, OP, 64, 8>(it, nrows, row0, lane);
        else if (ne == 16) rows_loop<T, OP, 64, 16>(it, nrows, row0, lane);
        else if (sizeof(T) == 4 && ne == 32) rows_loop<T, OP, 64, sizeof(T) == 4 ? 32 : 16>(it, nrows, row0, lane);
        else {
            for (int r = 0; r < nrows; ++r) row_stream<T, OP>(it, row0 + r, true, lane);
        }
    }
}

// queue reset as a kernel (not a memset) so that graph replays stay exact
__global__ void batch_reset_kernel(uint32_t* c) {
    if (threadIdx.x == 0) *c = 0u;
}

using BatchKernel = void (*)(const BatchArgs);

static BatchKernel batch_kernel_for(int op, int dtype) {
    if (dtype == XG_FLOAT32) {
        return op == BO_SOFTMAX ? batch_rows_kernel<float, BO_SOFTMAX>
             : op == BO_SUM     ? batch_rows_kernel<float, BO_SUM>
                                : batch_rows_kernel<float, BO_MAX>;
    }
    if (dtype == XG_FLOAT64) {
        return op == BO_SOFTMAX ? batch_rows_kernel<double, BO_SOFTMAX>
             : op == BO_SUM     ? batch_rows_kernel<double, BO_SUM>
                                : batch_rows_kernel<double, BO_MAX>;
    }
    if (dtype == XG_INT32) {
        return op == BO_SUM ? batch_rows_kernel<int32_t, BO_SUM>
             : op == BO_MAX ? batch_rows_kernel<int32_t, BO_MAX>
                            : nullptr;
    }
    return nullptr;
}

// ---------------------------------------------------------------- host side
int64_t batch_task_cost(const BItem& it, const BTask& t, int dtype) {
    // register slots the task occupies (padded rows x padded row length)
    // plus a fixed cost per row-group iteration
    const int esz = dtype_size(dtype);
    const int ne = batch_ne_class(it.cols, it.lpr, esz);
    const int64_t per = 64 / it.lpr;
    const int64_t step = per * batch_rows_at_once(ne, esz);
    const int64_t groups = (t.nrows + per - 1) / per;
    const int64_t slots = ne ? int64_t(ne) * it.lpr : int64_t(it.cols) * 3;  // streamed rows: 3 passes
    return groups * std::max<int64_t>(slots, 64) + 512 * ((t.nrows + step - 1) / step);
}

void batch_split_tasks(const BItem* items, size_t n, int op, int dtype, int task_elems, std::vector<BTask>& tasks) {
    tasks.clear();
    const int esz = dtype_size(dtype);
    for (size_t i = 0; i < n; ++i) {
        const BItem& it = items[i];
        if (it.rows == 0) continue;
        if (it.cols == 0 && op == BO_SOFTMAX) continue;
        const int64_t step = int64_t(64 / it.lpr) * batch_rows_at_once(batch_ne_class(it.cols, it.lpr, esz), esz);
        const int64_t want = std::max<int64_t>(1, task_elems / std::max<int64_t>(1, it.cols));
        const int64_t rpt = ((want + step - 1) / step) * step;
        for (int64_t r0 = 0; r0 < it.rows; r0 += rpt) {
            const int64_t k = std::min<int64_t>(rpt, it.rows - r0);
            tasks.push_back(BTask{int32_t(i), int32_t(r0), int32_t(k), 0});
        }
    }
    if (tasks.size() >= (size_t(1) << 31)) raise(XG_UNSUPPORTED, "batch has too many tasks");
}

BItem batch_item(const void* in, void* out, int64_t rows, int64_t cols, int64_t in_rs, int64_t out_rs, int op,
                 int dtype) {
    BItem d{};
    d.in = static_cast<const char*>(in);
    d.out = static_cast<char*>(out);
    // a single row's strides are never used to step; normalise them
    d.in_rs = rows > 1 ? in_rs : cols;
    d.out_rs = rows > 1 ? out_rs : (op == BO_SOFTMAX ? cols : 1);
    d.rows = int32_t(rows);
    d.cols = int32_t(cols);
    d.lpr = lanes_per_row(cols);
    // A wave addresses the rows of a group through one buffer resource with
    // 32-bit offsets; rows too far apart fall back to one row per wave, which
    // leaves every result unchanged (the extra lanes only fold in exact
    // zeros / lowest()).
    const int64_t esz = dtype_size(dtype);
    const int64_t rs = std::max<int64_t>(d.in_rs, op == BO_SOFTMAX ? d.out_rs : 0);
    if (d.lpr < 64 && (int64_t(64 / d.lpr - 1) * rs + cols) * esz >= (int64_t(1) << 31)) d.lpr = 64;
    return d;
}

hipError_t launch_batch_rows(int op, int dtype, const BatchArgs& a, int blocks, int block, hipStream_t stream) {
    BatchKernel k = batch_kernel_for(op, dtype);
    if (!k) return hipErrorInvalidValue;
    hipLaunchKernelGGL(batch_reset_kernel, dim3(1), dim3(64), 0, stream, a.counter);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess || a.ntasks == 0) return e;
    hipLaunchKernelGGL(k, dim3(unsigned(blocks)), dim3(unsigned(block)), 0, stream, a);
    return hipGetLastError();
}

namespace {

// shape class: op, dtype, log2 of the element count, log2 of the mean row length
using ClassKey = std::tuple<int, int, int, int>;
std::mutex g_tune_mu;
std::map<ClassKey, xg_batch_tuning> g_tune_cache;

int ilog2(double x) { return x < 1.0 ? 0 : int(std::floor(std::log2(x))); }

xg_batch_tuning default_tuning() { return xg_batch_tuning{256, 4, 2048, 10}; }

void check_tuning(const xg_batch_tuning& t) {
    if (t.block < 64 || t.block > 256 || t.block % 64 != 0)
        raise(XG_INVALID_ARGUMENT, "batch tuning: block must be a multiple of 64 in [64, 256] (the kernel is built for at most 256 threads)");
    if (t.blocks_per_cu < 1 || t.blocks_per_cu > 64)
        raise(XG_INVALID_ARGUMENT, "batch tuning: blocks_per_cu must be in [1, 64]");
    if (t.task_elems < 64 || t.task_elems > (1 << 24))
        raise(XG_INVALID_ARGUMENT, "batch tuning: task_elems must be in [64, 2^24]");
    if (t.dyn_percent < 0 || t.dyn_percent > 100)
        raise(XG_INVALID_ARGUMENT, "batch tuning: dyn_percent must be in [0, 100]");
}

}  // namespace

}  // namespace xg

using namespace xg;

struct xg_batch_s {
    int op = 0;
    int dtype = 0;
    int device = 0;
    int cus = 0;
    std::vector<BItem> items;
    std::vector<BTask> tasks;
    int64_t elements = 0;
    ClassKey cls{};
    xg_batch_tuning tune{};
    BItem* d_items = nullptr;
    BTask* d_tasks = nullptr;
    size_t d_tasks_cap = 0;
    std::vector<uint32_t> wbeg;
    uint32_t* d_wbeg = nullptr;
    size_t d_wbeg_cap = 0;
    uint32_t dyn_begin = 0;
    uint32_t* d_counter = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;

    ~xg_batch_s() { release(); }

    void release() {
        drop_graph();
        if (d_items) (void)hipFree(d_items);
        if (d_tasks) (void)hipFree(d_tasks);
        if (d_wbeg) (void)hipFree(d_wbeg);
        if (d_counter) (void)hipFree(d_counter);
        d_items = nullptr;
        d_tasks = nullptr;
        d_wbeg = nullptr;
        d_counter = nullptr;
    }

    void drop_graph() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
    }

    BatchArgs args() const {
        return BatchArgs{d_items, d_tasks, d_wbeg, dyn_begin, uint32_t(tasks.size()), d_counter};
    }
    int grid_blocks() const { return std::max(1, cus * tune.blocks_per_cu); }
    int grid_waves() const { return grid_blocks() * (tune.block / 64); }

    int64_t task_cost(const BTask& t) const { return batch_task_cost(items[size_t(t.item)], t, dtype); }

    // Split every instance into tasks of about tune.task_elems elements
    // (whole row-group iterations of one wave), then schedule them: all but
    // the smallest tune.dyn_percent of the work is assigned longest first to
    // the least loaded wave of the grid (no atomics, balanced to within one
    // task); the rest, largest first, goes to the shared work queue.
    void build_schedule() {
        batch_split_tasks(items.data(), items.size(), op, dtype, tune.task_elems, tasks);
        std::vector<int64_t> cost(tasks.size());
        std::vector<size_t> order(tasks.size());
        int64_t total = 0;
        for (size_t k = 0; k < tasks.size(); ++k) {
            cost[k] = task_cost(tasks[k]);
            order[k] = k;
            total += cost[k];
        }
        std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) { return cost[a] > cost[b]; });
        // the dynamic share: a suffix of the sorted list
        size_t nstatic = order.size();
        {
            const int64_t dyn_target = total * int64_t(tune.dyn_percent) / 100;
            int64_t acc = 0;
            while (nstatic > 0 && acc + cost[order[nstatic - 1]] <= dyn_target) acc += cost[order[--nstatic]];
        }
        // longest processing time first onto the least loaded wave
        const int W = grid_waves();
        std::vector<std::vector<uint32_t>> lists(static_cast<size_t>(W));
        {
            using Load = std::pair<int64_t, int>;
            std::priority_queue<Load, std::vector<Load>, std::greater<Load>> heap;
            for (int w = 0; w < W; ++w) heap.push({0, w});
            for (size_t k = 0; k < nstatic; ++k) {
                Load l = heap.top();
                heap.pop();
                lists[size_t(l.second)].push_back(uint32_t(order[k]));
                l.first += cost[order[k]];
                heap.push(l);
            }
        }
        std::vector<BTask> sched;
        sched.reserve(tasks.size());
        wbeg.assign(size_t(W) + 1, 0);
        for (int w = 0; w < W; ++w) {
            wbeg[size_t(w)] = uint32_t(sched.size());
            for (uint32_t k : lists[size_t(w)]) sched.push_back(tasks[k]);
        }
        wbeg[size_t(W)] = uint32_t(sched.size());
        dyn_begin = uint32_t(sched.size());
        for (size_t k = nstatic; k < order.size(); ++k) sched.push_back(tasks[order[k]]);
        tasks.swap(sched);
        upload(d_tasks, d_tasks_cap, tasks);
        upload(d_wbeg, d_wbeg_cap, wbeg);
    }

    template <class V> void upload(V*& dptr, size_t& cap, const std::vector<V>& h) {
        if (h.size() > cap) {
            drop_graph();
            if (dptr) check_hip(hipFree(dptr), "hipFree");
            dptr = nullptr;
            cap = 0;
            if (hipMalloc(&dptr, std::max<size_t>(1, h.size()) * sizeof(V)) != hipSuccess)
                raise(XG_OUT_OF_MEMORY, "batch schedule allocation failed");
            cap = std::max<size_t>(1, h.size());
        }
        if (!h.empty()) check_hip(hipMemcpy(dptr, h.data(), h.size() * sizeof(V), hipMemcpyHostToDevice), "hipMemcpy");
    }

    void enqueue(hipStream_t stream) const {
        BatchKernel k = batch_kernel_for(op, dtype);
        hipLaunchKernelGGL(batch_reset_kernel, dim3(1), dim3(64), 0, stream, d_counter);
        check_hip(hipGetLastError(), "batch reset launch");
        if (tasks.empty()) return;
        hipLaunchKernelGGL(k, dim3(grid_blocks()), dim3(tune.block), 0, stream, args());
        check_hip(hipGetLastError(), "batch kernel launch");
    }

    // two kernel nodes: queue reset -> persistent kernel
    void build_graph() {
        drop_graph();
        check_hip(hipGraphCreate(&graph, 0), "hipGraphCreate");
        uint32_t* cptr = d_counter;
        void* reset_args[] = {&cptr};
        hipKernelNodeParams rp{};
        rp.func = reinterpret_cast<void*>(batch_reset_kernel);
        rp.gridDim = dim3(1);
        rp.blockDim = dim3(64);
        rp.sharedMemBytes = 0;
        rp.kernelParams = reset_args;
        rp.extra = nullptr;
        hipGraphNode_t n_reset = nullptr;
        check_hip(hipGraphAddKernelNode(&n_reset, graph, nullptr, 0, &rp), "hipGraphAddKernelNode");
        if (!tasks.empty()) {
            BatchArgs a = args();
            void* kargs[] = {&a};
            hipKernelNodeParams kp{};
            kp.func = reinterpret_cast<void*>(batch_kernel_for(op, dtype));
            kp.gridDim = dim3(unsigned(grid_blocks()));
            kp.blockDim = dim3(unsigned(tune.block));
            kp.sharedMemBytes = 0;
            kp.kernelParams = kargs;
            kp.extra = nullptr;
            hipGraphNode_t n_k = nullptr;
            check_hip(hipGraphAddKernelNode(&n_k, graph, &n_reset, 1, &kp), "hipGraphAddKernelNode");
        }
        check_hip(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0), "hipGraphInstantiate");
    }

    void apply(const xg_batch_tuning& t) {
        check_tuning(t);
        tune = t;
        drop_graph();
        build_schedule();
    }
};

namespace {

void check_item(const xg_batch_item& it, int op, int64_t i) {
    const std::string at = "batch item " + std::to_string(i) + ": ";
    if (it.rows < 0 || it.cols < 0) raise(XG_INVALID_ARGUMENT, at + "negative extent");
    if (it.rows >= (int64_t(1) << 31) || it.cols >= (int64_t(1) << 31))
        raise(XG_UNSUPPORTED, at + "rows and cols must be below 2^31");
    if (it.rows == 0) return;
    if (!it.out) raise(XG_INVALID_ARGUMENT, at + "null output");
    if (it.cols > 0 && !it.in) raise(XG_INVALID_ARGUMENT, at + "null input");
    if (it.rows > 1 && it.in_row_stride < it.cols)
        raise(XG_INVALID_ARGUMENT, at + "input row stride below the row length");
    if (op == BO_SOFTMAX) {
        if (it.rows > 1 && it.out_row_stride < it.cols)
            raise(XG_INVALID_ARGUMENT, at + "output row stride below the row length");
    } else if (it.rows > 1 && it.out_row_stride < 1) {
        raise(XG_INVALID_ARGUMENT, at + "output stride must be positive");
    }
}

}  // namespace

extern "C" {

xg_status xg_batch_create(int32_t op, int32_t dtype, const xg_batch_item* items, int64_t n_items, xg_batch_t* out) {
    return guard([&] {
        if (!out) raise(XG_INVALID_ARGUMENT, "null batch handle");
        *out = nullptr;
        if (op < BO_SOFTMAX || op > BO_MAX) raise(XG_INVALID_ARGUMENT, "unknown batch op");
        if (n_items < 0 || (n_items > 0 && !items)) raise(XG_INVALID_ARGUMENT, "bad item list");
        if (n_items >= (int64_t(1) << 31)) raise(XG_UNSUPPORTED, "too many batch items");
        if (!batch_kernel_for(op, dtype))
            raise(XG_TYPE_ERROR, "batch op not available for this dtype (float32/float64; int32 for sum/amax)");
        auto b = std::make_unique<xg_batch_s>();
        b->op = op;
        b->dtype = dtype;
        check_hip(hipGetDevice(&b->device), "hipGetDevice");
        b->cus = device_cus();
        b->items.resize(size_t(n_items));
        int64_t rows_total = 0;
        for (int64_t i = 0; i < n_items; ++i) {
            const xg_batch_item& s = items[i];
            check_item(s, op, i);
            BItem& d = b->items[size_t(i)];
            d = batch_item(s.in, s.out, s.rows, s.cols, s.in_row_stride, s.out_row_stride, op, dtype);
            d.pad_ = 0;
            b->elements += s.rows * s.cols;
            rows_total += s.rows;
        }
        const double mean_cols = rows_total ? double(b->elements) / double(rows_total) : 0.0;
        b->cls = ClassKey{op, dtype, ilog2(double(b->elements)), ilog2(mean_cols)};
        xg_batch_tuning t = default_tuning();
        {
            std::lock_guard<std::mutex> lk(g_tune_mu);
            auto f = g_tune_cache.find(b->cls);
            if (f != g_tune_cache.end()) t = f->second;
        }
        if (hipMalloc(&b->d_counter, sizeof(uint32_t)) != hipSuccess) raise(XG_OUT_OF_MEMORY, "batch counter allocation failed");
        if (hipMalloc(&b->d_items, std::max<size_t>(1, b->items.size()) * sizeof(BItem)) != hipSuccess)
            raise(XG_OUT_OF_MEMORY, "batch item table allocation failed");
        if (!b->items.empty())
            check_hip(hipMemcpy(b->d_items, b->items.data(), b->items.size() * sizeof(BItem), hipMemcpyHostToDevice), "hipMemcpy");
        check_tuning(t);
        b->tune = t;
        b->build_schedule();
        *out = b.release();
    });
}

xg_status xg_batch_autotune(xg_batch_t b, xg_stream_t stream) {
    return guard([&] {
        if (!b) raise(XG_INVALID_ARGUMENT, "null batch");
        hipStream_t s = reinterpret_cast<hipStream_t>(stream);
        // Waves per CU up to what fits at once: the host-side schedule
        // assumes every wave of the grid is resident from the start.
        static const int kBlocks[] = {128, 256};
        static const int kWavesPerCu[] = {8, 12, 16, 20, 24, 32};
        static const int kElems[] = {1024, 2048, 4096, 8192};
        static const int kDyn[] = {0, 10, 30};
        hipEvent_t e0, e1;
        check_hip(hipEventCreate(&e0), "hipEventCreate");
        check_hip(hipEventCreate(&e1), "hipEventCreate");
        xg_batch_tuning best = b->tune;
        float best_ms = 1e30f;
        try {
            for (int te : kElems) {
                for (int blk : kBlocks) {
                    int resident = 0;
                    check_hip(hipOccupancyMaxActiveBlocksPerMultiprocessor(
                                  &resident, reinterpret_cast<const void*>(batch_kernel_for(b->op, b->dtype)), blk, 0),
                              "hipOccupancyMaxActiveBlocksPerMultiprocessor");
                    for (int wpc : kWavesPerCu) {
                      const int pc = wpc / (blk / 64);
                      if (pc < 1 || pc > std::max(1, resident)) continue;
                      for (int dp : kDyn) {
                        xg_batch_tuning t{blk, pc, te, dp};
                        b->apply(t);
                        b->enqueue(s);  // warm-up
                        check_hip(hipEventRecord(e0, s), "hipEventRecord");
                        for (int r = 0; r < 3; ++r) b->enqueue(s);
                        check_hip(hipEventRecord(e1, s), "hipEventRecord");
                        check_hip(hipEventSynchronize(e1), "hipEventSynchronize");
                        float ms = 0.f;
                        check_hip(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
                        if (ms < best_ms) {
                            best_ms = ms;
                            best = t;
                        }
                      }
                    }
                }
            }
        } catch (...) {
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
            throw;
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        b->apply(best);
        std::lock_guard<std::mutex> lk(g_tune_mu);
        g_tune_cache[b->cls] = best;
    });
}

xg_status xg_batch_get_tuning(xg_batch_t b, xg_batch_tuning* t) {
    return guard([&] {
        if (!b || !t) raise(XG_INVALID_ARGUMENT, "null argument");
        *t = b->tune;
    });
}

xg_status xg_batch_set_tuning(xg_batch_t b, const xg_batch_tuning* t) {
    return guard([&] {
        if (!b || !t) raise(XG_INVALID_ARGUMENT, "null argument");
        b->apply(*t);
    });
}

xg_status xg_batch_launch(xg_batch_t b, xg_stream_t stream) {
    return guard([&] {
        if (!b) raise(XG_INVALID_ARGUMENT, "null batch");
        b->enqueue(reinterpret_cast<hipStream_t>(stream));
        note_kernel("batch_rows_kernel");
    });
}

xg_status xg_batch_execute(xg_batch_t b, xg_stream_t stream) {
    return guard([&] {
        if (!b) raise(XG_INVALID_ARGUMENT, "null batch");
        if (!b->exec) b->build_graph();
        check_hip(hipGraphLaunch(b->exec, reinterpret_cast<hipStream_t>(stream)), "hipGraphLaunch");
        note_kernel("batch_rows_kernel");
    });
}

xg_status xg_batch_info(xg_batch_t b, int64_t* n_tasks, int64_t* n_elements) {
    return guard([&] {
        if (!b) raise(XG_INVALID_ARGUMENT, "null batch");
        if (n_tasks) *n_tasks = int64_t(b->tasks.size());
        if (n_elements) *n_elements = b->elements;
    });
}

void xg_batch_destroy(xg_batch_t b) { delete b; }

}  // extern "C"
