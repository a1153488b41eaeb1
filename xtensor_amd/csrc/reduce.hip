// reduce.hip — axis reductions (sum, prod, amax, amin, mean, norms) on gfx950.
//
// Replaces xreducer's sequential per-output fold
// (reducers/xreducer.hpp:1801-1828) by a tree: every thread folds a strided
// subset of an output's elements, then threads combine with wave shuffles
// and LDS in a fixed order, so a result depends only on the shape of its
// reduction.  Two layouts cover any axis set:
//   rows: the contiguous dim is reduced; a group of 1..256 threads per
//         output reads its elements with vector loads;
//   cols: the contiguous dim is kept; each thread owns V adjacent outputs
//         and walks the reduced rows, four waves splitting the rows.
// Large reductions are split over several workgroups; their partial
// accumulators are combined by a second kernel in split order.
// Float sums carry a TwoSum compensation term (Neumaier / Kahan-Babuska
// summation with error-free pairwise combining); integer reductions use the
// exact two's-complement arithmetic of the value type.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>
#include <vector>

#include "jit.hpp"
#include "reduce_kernels.hpp"


#include "specs.inc"

namespace xg {

struct RSpecEntry {
    const char* name;
    int V, acc, op;
    SpecKey key;
    RedKernels k;
};

namespace {

template <class S> SpecKey rkey_of() {
    SpecKey k{};
    k.nops = S::nops;
    for (int i = 0; i < S::nops; ++i) k.ops[i] = S::ops[i];
    k.nleaves = S::nleaves;
    for (int l = 0; l < S::nleaves; ++l) {
        k.leaf_dtype[l] = S::leaf_dtype[l];
        k.leaf_mode[l] = S::leaf_mode[l];
    }
    k.result_reg = S::result_reg;
    return k;
}

std::vector<RSpecEntry>& rspec_table() {
    static std::vector<RSpecEntry> t;
    static std::once_flag once;
    std::call_once(once, [] {
#define XG_REGISTER_RSPEC(S) \
    t.push_back(RSpecEntry{#S, S::V, S::acc, S::op, rkey_of<S>(), spec_kernels_for<S::V, S::nregs, S, S::acc, S::op>()});
        XG_RSPEC_LIST(XG_REGISTER_RSPEC)
#undef XG_REGISTER_RSPEC
    });
    return t;
}

}  // namespace

// a compiled specialisation for (program, V, accumulator, op), or nullptr
const RSpecEntry* find_rspec(const RedPlan& p, int V) {
    SpecKey k{};
    k.nops = p.prog.nops;
    for (int i = 0; i < p.prog.nops; ++i) k.ops[i] = p.prog.ops[i];
    k.nleaves = p.prog.nleaves;
    for (int l = 0; l < p.prog.nleaves; ++l) {
        k.leaf_dtype[l] = uint8_t(p.leaves[l].dtype);
        k.leaf_mode[l] = uint8_t(lm_base(p.leaves[l].mode));
    }
    k.result_reg = p.prog.result_reg;
    for (auto& s : rspec_table())
        if (s.V == V && s.acc == p.acc && s.op == p.op && spec_equal(s.key, k)) return &s;
    return nullptr;
}

// Combine split partials: a thread per output when there are many outputs,
// a block per output when few outputs carry many partials each.
// the combine kernel: a block per output for many partials of few outputs
static bool combine_wide(const RedPlan& p) { return p.splits >= 1024 && p.n_out < 256; }

// lanes per output of the combine kernel: about 8 partials per lane, at most a wave
static int combine_lanes(int splits) {
    int T = 1;
    while (T < 64 && T * 8 < splits) T <<= 1;
    return T;
}

static hipError_t launch_combine(const RedPlan& p, hipStream_t stream) {
    const bool wide = combine_wide(p);
    KFn comb = combine_kernel(p.acc, p.op, wide);
    if (!comb) return hipErrorInvalidValue;
    RedPlan q = p;
    const uint32_t T = uint32_t(combine_lanes(p.splits));
    q.group = int(T);
    const uint64_t blocks = wide ? std::min<uint64_t>(q.n_out, 1u << 20)
                                 : std::min<uint64_t>((uint64_t(q.n_out) * T + 255) / 256, 1u << 20);
    hipLaunchKernelGGL(comb, dim3(uint32_t(std::max<uint64_t>(blocks, 1))), dim3(256), 0, stream, q);
    return hipGetLastError();
}

hipError_t launch_reduce_combine(const RedPlan& p, hipStream_t stream) { return launch_combine(p, stream); }

namespace {

uint64_t plan_elements(const RedPlan& p) {
    uint64_t n = 1;
    for (int k = 0; k < p.rank; ++k) n *= uint64_t(std::max<int64_t>(p.shape[k], 0));
    return n;
}

// the kernel template a plan runs: a fast one for compact plans
int jit_kind(const RedPlan& p) {
    if (p.compact) return p.layout == RL_ROWS ? (p.rank == 1 ? JR_ROWS_FAST1 : JR_ROWS_FAST2) : JR_COLS_FAST;
    return p.layout == RL_ROWS ? JR_ROWS : JR_COLS;
}

}  // namespace

bool reduce_jit_ok(const RedPlan& p, int V) {
    return !p.index64 && p.n_out > 0 && jit_wanted(plan_elements(p)) && jit_reduce(p, V, jit_kind(p)) != nullptr;
}

hipError_t launch_reduce(RedPlan& p, int V, hipStream_t stream, const char** used) {
    if (p.n_out == 0) return hipSuccess;
    RedKernels ks{};
    const JitKernel* jk = nullptr;
    const RSpecEntry* sp = p.index64 ? nullptr : find_rspec(p, V);
    if (sp) {
        ks = sp->k;
        if (used) *used = sp->name;
    } else if (reduce_jit_ok(p, V)) {  // a run-time compiled specialisation
        jk = jit_reduce(p, V, jit_kind(p));
        if (used) *used = jk->name.c_str();
    } else {
        if (V != 4) return hipErrorInvalidValue;  // dynamic kernels are built for V = 4
        if (!dyn_reduce_float(p.acc, p.op, &ks) && !dyn_reduce_int32(p.acc, p.op, &ks) &&
            !dyn_reduce_int64(p.acc, p.op, &ks))
            return hipErrorInvalidValue;
        if (used) *used = "dynamic";
    }
    auto go = [&](KFn f, uint32_t grid, uint32_t block, const RedPlan& q) -> hipError_t {
        if (jk) return jit_launch(jk, grid, block, &q, sizeof(RedPlan), stream);
        if (!f) return hipErrorInvalidValue;
        hipLaunchKernelGGL(f, dim3(grid), dim3(block), 0, stream, q);
        return hipGetLastError();
    };
    const bool fast = p.compact &&
                      (jk || (sp && (p.layout == RL_ROWS ? (p.rank == 1 ? ks.rows_fast1 : ks.rows_fast2) : ks.cols_fast)));
    if (fast) {
        // compiled fast kernels with their own split of the work
        RedPlan q = p;
        q.group = p.f_group;
        q.splits = p.f_splits;
        q.per_split = p.f_per_split;
        q.div_splits = make_fastdiv(uint32_t(p.f_splits));
        hipError_t err;
        if (q.layout == RL_ROWS) {
            const uint64_t items = uint64_t(q.n_out) * uint64_t(q.splits);
            const uint64_t per_block = 4 * (64 / uint64_t(q.group));
            const uint64_t blocks = (items + per_block - 1) / per_block;
            if (blocks >= (uint64_t(1) << 31)) return hipErrorInvalidValue;
            err = go(q.rank == 1 ? ks.rows_fast1 : ks.rows_fast2, uint32_t(blocks), 256, q);
        } else {
            const uint64_t tw = uint64_t(q.group);  // tile width in lanes
            if (tw != 64 && tw != 128 && tw != 256) return hipErrorInvalidValue;
            const uint64_t tiles = (q.nchunks_f + tw - 1) / tw;
            const uint64_t blocks = tiles * uint64_t(q.splits);
            if (blocks >= (uint64_t(1) << 31)) return hipErrorInvalidValue;
            // partials [split][output]: a workgroup's partials are one
            // contiguous run, not one 16-byte write per cache line (1.5-3 %
            // faster on [8192, 4096] float32 / int32 column sums, MI355X)
            q.split_major = 1;
            // The splits' partials are merged by a second launch.  Merging
            // them in the last-arriving workgroup of each tile instead
            // (arrival counters, device-scope stores and loads of the
            // partials) measured slower on MI355X: every tile finishes at
            // once and the merges then run on a few workgroups, latency-bound
            // (float32 [8192, 4096] column sum 31 -> 45 us, bf16 27 -> 81 us).
            q.counters = nullptr;
            err = go(ks.cols_fast, uint32_t(blocks), 512, q);
        }
        if (err != hipSuccess || q.splits == 1) return err;
        return launch_combine(q, stream);
    }
    hipError_t err;
    if (p.layout == RL_ROWS) {
        uint64_t items = uint64_t(p.n_out) * p.splits;
        uint64_t per_block = 256 / p.group;
        uint64_t blocks = (items + per_block - 1) / per_block;
        blocks = std::min<uint64_t>(std::max<uint64_t>(blocks, 1), 1u << 20);
        err = go(p.index64 ? ks.rows64 : ks.rows32, uint32_t(blocks), 256, p);
    } else {
        uint64_t items = uint64_t(p.n_out / p.lf) * p.col_tiles * p.splits;
        uint64_t blocks = std::min<uint64_t>(std::max<uint64_t>(items, 1), 1u << 20);
        err = go(p.index64 ? ks.cols64 : ks.cols32, uint32_t(blocks), 256, p);
    }
    if (err != hipSuccess || p.splits == 1) return err;
    return launch_combine(p, stream);
}

}  // namespace xg
