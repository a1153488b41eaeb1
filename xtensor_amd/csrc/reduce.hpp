// reduce.hpp — plans and host entry points of the axis-reduction kernels.
#pragma once

#include <hip/hip_runtime.h>

#include "planner.hpp"
#include "xg_internal.hpp"

namespace xg {

enum RedLayout : int32_t {
    RL_ROWS = 0,  // the contiguous (fast) dim is reduced: a thread group per output
    RL_COLS = 1   // the fast dim is kept: a thread per V adjacent outputs, rows split over waves
};

// Accumulation type of the reduction tree.
enum AccKind : int32_t {
    AK_F32C = 0,  // float sum with a TwoSum (Neumaier / Kahan-Babuska) compensation term
    AK_F64C,      // double sum with compensation
    AK_F32,       // float, for products, maxima, minima
    AK_F64,
    AK_I32,       // integers: exact two's-complement arithmetic of the value type
    AK_U32,
    AK_I64,
    AK_U64,
    AK_NUM
};

// RO_MAX / RO_MIN fold with xt::maximum / xt::minimum (select(a > b, a, b),
// a = accumulator: a NaN value enters the accumulator); RO_MAXNUM /
// RO_MINNUM with std::max / std::min (a NaN value never replaces it), for
// xt::minmax on floats.
enum RedOp : int32_t { RO_ADD = 0, RO_MUL, RO_MAX, RO_MIN, RO_MAXNUM, RO_MINNUM, RO_NUM };

// FZ_MEAN: value / divisor in the result type; FZ_MEAN_SQRT: sqrt of that
enum Finalize : int32_t { FZ_NONE = 0, FZ_MEAN = 1, FZ_SQRT = 2, FZ_MEAN_SQRT = 3 };

// internal pass codes of xt::minmax (its min and max folds)
constexpr int kRedMinNum = 100;
constexpr int kRedMaxNum = 101;

struct RedPlan {
    int32_t layout;
    int32_t rank;       // normalised dims (see below)
    int32_t acc;        // AccKind
    int32_t op;         // RedOp
    int32_t fin;        // Finalize
    int32_t index64;
    // RL_ROWS: dims = [kept 0..n_kept-1 | reduced outer | reduced fast]
    // RL_COLS: dims = [kept outer 0..n_kept-1 | reduced ... | kept fast]
    int32_t n_kept;     // leading kept dims
    int32_t n_red;      // reduced dims
    int64_t shape[kMaxRank];
    FastDiv div[kMaxRank];
    uint32_t n_out;     // number of output elements
    uint32_t lf;        // length of the last (fast) dim
    uint32_t nchunks_f; // ceil(lf / V)
    FastDiv div_nchunks_f;
    uint32_t red_count; // reduced elements per output (RL_COLS: rows; RL_ROWS: chunks)
    int32_t group;      // RL_ROWS threads per output (1..256, power of two)
    int32_t splits;     // partial results per output (1 = write the result directly)
    uint32_t per_split; // RL_ROWS chunks / RL_COLS rows handled by one split
    uint32_t col_tiles; // RL_COLS: ceil(nchunks_f / 64)
    FastDiv div_col_tiles;
    FastDiv div_splits;
    // finishing: acc -> value type V (vdt) -> [+initial] -> mean / sqrt (rdt) -> out dtype
    int32_t vdt;
    int32_t rdt;
    int32_t has_initial;
    int32_t out_dtype;
    uint64_t identity;  // start value of every accumulator, in acc representation
    uint64_t initial;   // xt::initial, in acc representation
    int32_t mdt;        // FZ_MEAN / FZ_MEAN_SQRT: type of value / divisor
    int32_t split_major;  // partials stored [split][output] (cols fast kernel), else [output][split]
    uint64_t div_bits;  // the divisor, in register form of mdt
    char* out_ptr;
    int64_t out_stride[kMaxRank];  // over the kept dims in normalised order (fast last for RL_COLS)
    void* partials;     // workspace: splits * n_out accumulators of 16 bytes
    // cols fast kernel with splits: one arrival counter per tile, zero
    // between launches; the last split of a tile to finish merges the tile's
    // partials (with `combine_group` lanes per output, as the combine kernel
    // would).  Null: a separate combine kernel merges them.
    uint32_t* counters;
    int32_t combine_group;
    LeafView leaves[kMaxLeaves];
    Program prog;       // element value mapped and converted to the accumulation type
    // Compact form for the compiled fast kernels: normalised rank <= 2 and
    // 32-bit offsets (rows: [outputs | fast] or [fast]; cols: [rows | fast]).
    // cst[l][k] = stride of leaf l along dim k.  f_* = their work split.
    int32_t compact;
    int32_t f_group;      // rows: lanes per output (1..64, power of two); cols: tile width in lanes (64, 128, 256)
    int32_t f_splits;     // partial results per output
    uint32_t f_per_split; // rows: chunks per split; cols: rows per split
    int32_t cst[kMaxLeaves + 1][4];
};

// Plan out = reduce(e).  Throws Error.  `V` is the vector width.
RedPlan plan_reduce(const xg_expr* e, const xg_reduce_desc* r, const xg_tensor* out, int V, int cus,
                    int* result_dtype, Shape* result_shape);

// Result dtype / shape only.
void reduce_result(const xg_expr* e, const xg_reduce_desc* r, int* dtype, Shape* shape);

// How the folded value of a reduction becomes its result: value type, the
// finishing step, the type and divisor of a mean, and the result type.  Host
// only and independent of the iteration space, so it also serves reductions
// that are evaluated in pieces and merged.
struct RedFinish {
    int32_t vdt;
    int32_t rdt;
    int32_t fin;       // Finalize
    int32_t mdt;
    uint64_t div_bits;
};
RedFinish reduce_finish(const xg_expr* e, const xg_reduce_desc* r);

#if !defined(__HIPCC_RTC__)  // host API (absent from run-time compiled sources)
hipError_t launch_reduce(RedPlan& p, int V, hipStream_t stream, const char** used);
// The split-combine launch alone: merges p.splits partials per output from
// p.partials (laid out as p.split_major says) and finishes the outputs.
hipError_t launch_reduce_combine(const RedPlan& p, hipStream_t stream);
#endif

// Arrival counters the cols fast kernel can merge its splits with: zeroed
// device memory of kReduceCounters words, left zeroed by every launch.
constexpr uint32_t kReduceCounters = 4096;

size_t reduce_workspace_bytes(const RedPlan& p);

// Waves per CU the compiled reduction kernels' work split aims at (default
// 16; measurement aid), for plans made after the call.
void set_reduce_waves_per_cu(int w);
// column fast kernel: workgroups per CU and the partials' byte budget
// multiplier (measurement aid; 1, 1 = the defaults)
void set_cols_split(int blocks_per_cu, int partial_mult);

struct RSpecEntry;
// compiled specialisation matching the plan, or nullptr
const RSpecEntry* find_rspec(const RedPlan& p, int V);
// True when a plan without an ahead-of-time program runs a run-time compiled
// one (compiling it on first use).
bool reduce_jit_ok(const RedPlan& p, int V);

}  // namespace xg
