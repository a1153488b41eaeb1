// ew.hip — elementwise assignment kernels (out = f(leaves...)) for gfx950.
//
// Replaces xtensor's linear / strided-loop / stepper assigners
// (core/xassign.hpp:644-695, 701-849, 855-1358): instead of a host loop
// over a stepper, the planner hands over a normalised iteration space
// (outer dims x one inner row) and every work-item evaluates the micro-op
// program on V consecutive elements of one inner row.  Leaves contiguous
// along the row are read with 16-byte vector loads, leaves broadcast along
// the row (stride 0) with one load per work-item, others element by element;
// the output is written the same way.
#include <hip/hip_runtime.h>

#include <mutex>
#include <vector>

#include "chunk.hpp"
#include "spec_eval.hpp"
#include "ew.hpp"
#include "ew_kernels.hpp"
#include "jit.hpp"
#include "planner.hpp"

namespace xg {

template <int V, bool I64>
__device__ __forceinline__ void store_out(const EwPlan& plan, int mode, int dt, const uint32_t* idx, uint32_t j0,
                                          int nvalid, const Reg (&val)[V]) {
    const int esz = dtype_size(dt);
    auto off = unit_offset<I64>(plan.out_stride, idx, plan.rank, j0);
    char* p = plan.out_ptr + int64_t(off) * esz;
    if (mode == LM_CONTIG && nvalid == V) {
        store_vec<V>(p, esz, dt, val);
    } else {
        const int64_t os = plan.out_stride[plan.rank - 1];
#pragma unroll
        for (int v = 0; v < V; ++v)
            if (v < nvalid) store_raw(p + int64_t(v) * os * esz, esz, to_storage(val[v], dt));
    }
}

// The generic interpreter for elementwise plans of vector width 1: one
// element per work item and the program's register file in LDS, indexed
// [register][thread] (a register array indexed by run-time operands would
// live in scratch memory).  Each micro-op is one out-of-line call.
template <bool I64>
__global__ __launch_bounds__(256) void ew_lds_kernel(const EwPlan plan_arg) {
    const EwPlan& plan = kernel_plan<DynSpec>(plan_arg);
    __shared__ Reg regs[kMaxRegs][256];
    const int tid = int(threadIdx.x);
    const uint32_t nunits = uint32_t(plan.units);
    const uint32_t step = gridDim.x * 256u;
    for (uint32_t u = blockIdx.x * 256u + uint32_t(tid); u < nunits; u += step) {
        const uint32_t row = fastdiv(u, plan.div_chunks);
        const uint32_t j0 = u - row * plan.nchunks;
        uint32_t idx[kMaxRank];
        uint32_t r = row;
#pragma unroll
        for (int k = kMaxRank - 2; k >= 0; --k) {
            if (k <= plan.rank - 2) {
                const uint32_t q = fastdiv(r, plan.div[k]);
                idx[k] = r - q * uint32_t(plan.shape[k]);
                r = q;
            } else {
                idx[k] = 0;
            }
        }
        for (int pc = 0; pc < plan.prog.nops; ++pc) {
            const MicroOp m = plan.prog.ops[pc];
            Reg v[1];
            if (m.op == MO_LOAD) {
                const LeafView& lv = plan.leaves[m.aux];
                load_leaf<1, I64>(lv, lv.mode, lv.dtype, idx, plan.rank, j0, 1, v);
            } else if (m.op == MO_GATHER) {
                const LeafView& lv = plan.leaves[m.aux];
                const Reg off[1] = {regs[m.a][tid]};
                gather_vals<1>(lv.ptr, lv.dtype, lv.stride[0], lv.stride[1], 1, off, v);
            } else {
                v[0] = exec_scalar_dyn(m, regs[m.a][tid], regs[m.b][tid], regs[m.c][tid], plan.prog.consts);
            }
            regs[m.dst][tid] = v[0];
        }
        const Reg res[1] = {regs[plan.prog.result_reg][tid]};
        store_out<1, I64>(plan, lm_base(plan.out_mode), plan.out_dtype, idx, j0, 1, res);
    }
}

}  // namespace xg

// compiled specialisations (generated from specs_list.cpp at build time)
#include "specs.inc"

namespace xg {

namespace {

using EwFn = void (*)(const EwPlan);
using EwMultiFn = void (*)(const CompactPlan*, const uint64_t*, uint32_t);

struct SpecEntry {
    const char* name;
    int V;
    SpecKey key;
    EwFn by_rank[4];          // compiled kernels for normalised rank 1..4
    EwMultiFn multi_by_rank[4];  // fused many-expression kernels, rank 1..4
    EwFn wide_by_rank[4];     // four units per thread, for large launches
};

template <class S> SpecKey key_of() {
    SpecKey k{};
    k.nops = S::nops;
    for (int i = 0; i < S::nops; ++i) k.ops[i] = S::ops[i];
    k.nleaves = S::nleaves;
    for (int l = 0; l < S::nleaves; ++l) {
        k.leaf_dtype[l] = S::leaf_dtype[l];
        k.leaf_mode[l] = S::leaf_mode[l];
    }
    k.out_dtype = uint8_t(S::out_dtype);
    k.out_mode = uint8_t(S::out_mode);
    k.result_reg = S::result_reg;
    return k;
}

std::vector<SpecEntry>& spec_table() {
    static std::vector<SpecEntry> t;
    static std::once_flag once;
    std::call_once(once, [] {
#define XG_REGISTER_SPEC(S)                                                                                   \
    t.push_back(SpecEntry{#S, S::V, key_of<S>(),                                                              \
                          {ew_spec_kernel<S::V, S::nregs, S, 1>, ew_spec_kernel<S::V, S::nregs, S, 2>,          \
                           ew_spec_kernel<S::V, S::nregs, S, 3>, ew_spec_kernel<S::V, S::nregs, S, 4>},         \
                          {ew_multi_kernel<S::V, S::nregs, S, 1>, ew_multi_kernel<S::V, S::nregs, S, 2>,        \
                           ew_multi_kernel<S::V, S::nregs, S, 3>, ew_multi_kernel<S::V, S::nregs, S, 4>},       \
                          {ew_spec_multi_unit_kernel<S::V, S::nregs, S, 1, 4>,                                  \
                           ew_spec_multi_unit_kernel<S::V, S::nregs, S, 2, 4>,                                  \
                           ew_spec_multi_unit_kernel<S::V, S::nregs, S, 3, 4>,                                  \
                           ew_spec_multi_unit_kernel<S::V, S::nregs, S, 4, 4>}});
        XG_SPEC_LIST(XG_REGISTER_SPEC)
#undef XG_REGISTER_SPEC
    });
    return t;
}

int g_force_dynamic_all = 0;              // every thread (testing aid)
thread_local int t_force_dynamic = 0;     // the calling thread only
int g_units_per_thread = 0;  // 0: choose by size; 1 or 4: forced (measurement aid)

bool force_dynamic() { return g_force_dynamic_all || t_force_dynamic; }

}  // namespace

void ew_set_force_dynamic(int on) { g_force_dynamic_all = on; }

int ew_set_thread_force_dynamic(int on) {
    const int prev = t_force_dynamic;
    t_force_dynamic = on ? 1 : 0;
    return prev;
}

bool ew_jit_ok(const EwPlan& p, int V) {
    return !force_dynamic() && p.compact && jit_wanted(p.units * uint64_t(V)) && jit_ew(p, V) != nullptr;
}
void ew_set_units_per_thread(int upt) { g_units_per_thread = (upt == 1 || upt == 4) ? upt : 0; }

const char* ew_find_spec(const EwPlan& p, int V) {
    if (force_dynamic() || !p.compact) return nullptr;
    SpecKey k = spec_key(p);
    for (auto& s : spec_table())
        if (s.V == V && spec_equal(s.key, k)) return s.name;
    return nullptr;
}

// one work unit per thread (the loop in the kernels only runs beyond 2^31 blocks)
int ew_grid(uint64_t units, int cus) {
    (void)cus;
    uint64_t blocks = (units + 255) / 256;
    if (blocks > 0x7fffffffull) blocks = 0x7fffffffull;
    if (blocks < 1) blocks = 1;
    return int(blocks);
}

int ew_spec_index(const EwPlan& p, int V) {
    if (force_dynamic() || !p.compact) return -1;
    SpecKey k = spec_key(p);
    auto& t = spec_table();
    for (size_t i = 0; i < t.size(); ++i)
        if (t[i].V == V && spec_equal(t[i].key, k)) return int(i);
    return -1;
}

const char* ew_spec_name(int index) {
    auto& t = spec_table();
    return index >= 0 && size_t(index) < t.size() ? t[size_t(index)].name : nullptr;
}

CompactPlan make_compact(const EwPlan& p) {
    CompactPlan c{};
    for (int k = 0; k < 4; ++k) {
        c.shape32[k] = p.shape32[k];
        c.div[k] = p.div[k];
    }
    c.div_chunks = p.div_chunks;
    c.nchunks = p.nchunks;
    c.inner = p.inner;
    c.units = p.units;
    for (int l = 0; l <= kMaxLeaves; ++l)
        for (int k = 0; k < 4; ++k) c.cst[l][k] = p.cst[l][k];
    for (int l = 0; l < kMaxLeaves; ++l) c.leaf[l] = p.leaves[l].ptr;
    c.out_ptr = p.out_ptr;
    for (int i = 0; i < kMaxConsts; ++i) c.consts[i] = p.prog.consts[i];
    return c;
}

hipError_t launch_ew_multi(int spec_index, int rank, const CompactPlan* d_plans, const uint64_t* d_prefix, uint32_t n,
                           uint64_t total_units, hipStream_t stream) {
    auto& t = spec_table();
    if (spec_index < 0 || size_t(spec_index) >= t.size() || rank < 1 || rank > 4 || n == 0) return hipErrorInvalidValue;
    if (total_units == 0) return hipSuccess;
    dim3 grid(ew_grid(total_units, 0)), block(256);
    hipLaunchKernelGGL(t[size_t(spec_index)].multi_by_rank[rank - 1], grid, block, 0, stream, d_plans, d_prefix, n);
    return hipGetLastError();
}

hipError_t launch_ew(const EwPlan& p, int V, int cus, hipStream_t stream, const char** used) {
    if (p.units == 0) return hipSuccess;
    dim3 grid(ew_grid(p.units, cus)), block(256);
    if (!force_dynamic() && p.compact) {
        SpecKey k = spec_key(p);
        for (auto& s : spec_table()) {
            if (s.V == V && spec_equal(s.key, k)) {
                if (used) *used = s.name;
                // One unit per thread unless forced: four units per thread
                // measured slower on MI355X (lower occupancy: 2.74 vs 2.93 TB/s
                // on the strided-view affine workload, 5.3 vs 5.6 TB/s on the
                // contiguous int32 bit-mix).
                const bool wide = g_units_per_thread == 4;
                if (wide) {
                    dim3 g4(ew_grid((p.units + 3) / 4, cus));
                    hipLaunchKernelGGL(s.wide_by_rank[p.rank - 1], g4, block, 0, stream, p);
                } else {
                    hipLaunchKernelGGL(s.by_rank[p.rank - 1], grid, block, 0, stream, p);
                }
                return hipGetLastError();
            }
        }
    }
    if (ew_jit_ok(p, V)) {  // a run-time compiled specialisation
        const JitKernel* k = jit_ew(p, V);
        if (used) *used = k->name.c_str();
        return jit_launch(k, grid.x, 256, &p, sizeof(EwPlan), stream);
    }
    if (V != 1) return hipErrorInvalidValue;  // the generic interpreter takes one element per lane
    if (used) *used = "dynamic";
    if (p.index64) hipLaunchKernelGGL((ew_lds_kernel<true>), grid, block, 0, stream, p);
    else hipLaunchKernelGGL((ew_lds_kernel<false>), grid, block, 0, stream, p);
    return hipGetLastError();
}

}  // namespace xg
